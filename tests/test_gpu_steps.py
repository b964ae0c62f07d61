"""The step-level C ABI (hprqp_ws_*, hprqp_step_*, hprqp_kkt_reduce,
hprqp_mnorm, hprqp_sigma_dots, hprqp_restart): solve()'s loop run on the host
(hprqp.steploop, the restart and sigma logic of src/algorithm.jl:323-555 in
Python) over the device operations must give hprqp_solve's result bit for
bit -- the same iteration count, restarts, final sigma, and identical x and
y, with the row-product shapes the autotuner chose for the model shared --
on every path main_update takes (standard, noC with diagonal Q, LASSO,
LP, the QAP operator, no constraints), both through hprqp_step and through
the individual half-step entry points."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import fixture_path

pytestmark = pytest.mark.gpu


def _model(case):
    from hprqp import build_from_mat, build_from_mps, build_from_QAbc, instances

    if case == "random":
        return instances.random_sparse_qp(2000, 1000, seed=0), "QP"
    if case == "aug2d_noC_diagQ":
        return build_from_mps(fixture_path("AUG2D.mps"), verbose=False), "QP"
    if case == "lasso":
        return instances.lasso(200, 4000, mean_nnz=60, seed=1), "LASSO"
    if case == "qap_esc64a":
        return build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False), "QAP"
    if case == "lp":
        rng = np.random.default_rng(3)
        n, m = 60, 30
        A = sp.random(m, n, density=0.2, random_state=4, format="csr") + sp.eye(m, n)
        b = A @ rng.uniform(0, 1, n)
        return build_from_QAbc(sp.csr_matrix((n, n)), rng.uniform(0.1, 1.0, n), A, b, b + 1.0, np.zeros(n),
                               np.full(n, 5.0), verbose=False), "QP"
    if case == "box_m0":
        rng = np.random.default_rng(5)
        n = 80
        R = sp.random(40, n, density=0.1, random_state=6, format="csr")
        Q = (R.T @ R + sp.diags(rng.uniform(0.01, 1.0, n))).tocsr()
        return build_from_QAbc(Q, rng.normal(size=n), sp.csr_matrix((0, n)), np.zeros(0), np.zeros(0),
                               np.full(n, -1.0), np.full(n, 1.0), verbose=False), "QP"
    raise ValueError(case)


@pytest.mark.parametrize("half_steps", [False, True], ids=["hprqp_step", "half_steps"])
@pytest.mark.parametrize("case", ["random", "aug2d_noC_diagQ", "lasso", "qap_esc64a", "lp", "box_m0"])
def test_host_loop_reproduces_solve_bit_for_bit(case, half_steps):
    from hprqp import HPRQP_parameters
    from hprqp.solver import DeviceModel
    from hprqp.steploop import solve_by_steps

    model, ptype = _model(case)
    params = HPRQP_parameters(verbose=False, problem_type=ptype)
    if case == "qap_esc64a" and half_steps:
        # the half-step entries are the reference's four half-steps; the
        # one-apply QAP form fuses them differently, so compare like with like
        params.qap_recompute_Qw = True
    # one uploaded model: the row-product shapes the autotuner picks (by
    # timing) are cached in it, so both solves sum in the same order
    dm = DeviceModel(model)
    try:
        ref = dm.solve(params)
        got = solve_by_steps(model, params, half_steps=half_steps, device_model=dm)
    finally:
        dm.close()
    print(f"{case}: solve iter={ref.iter} restarts={ref.restarts} sigma={ref.sigma!r}; host loop iter={got.iter} "
          f"restarts={got.restarts} sigma={got.sigma!r} obj {got.primal_obj:.12e}")
    assert got.status == ref.status == "OPTIMAL"
    assert got.iter == ref.iter and got.restarts == ref.restarts
    assert got.sigma == ref.sigma
    assert got.primal_obj == ref.primal_obj and got.residuals == ref.residuals
    np.testing.assert_array_equal(got.x, ref.x)
    np.testing.assert_array_equal(got.y, ref.y)
    np.testing.assert_array_equal(got.z, ref.z)


def test_workspace_info_and_path_checks():
    """hprqp_ws_info reports what solve decides before its loop; a half-step
    that is not on the problem's path is refused with HPRQP_E_STATE, and a
    workspace is tied to its context."""
    from hprqp import HPRQP_parameters, _lib
    from hprqp.solver import DeviceModel
    from hprqp.steploop import Workspace

    model, _ = _model("aug2d_noC_diagQ")
    dm = DeviceModel(model)
    ref = dm.solve(HPRQP_parameters(verbose=False))
    ws = Workspace(dm, HPRQP_parameters(verbose=False))
    try:
        info = ws.info
        assert info.path == _lib.PATH_NOC and info.noC == 1 and info.Q_is_diag == 1
        assert info.lambda_max_A == ref.lambda_max_A and info.lambda_max_Q == ref.lambda_max_Q
        assert info.sigma0 == ref.sigma0 and info.power_iter_A == ref.power_iter_A
        with pytest.raises(_lib.HPRQPError) as ei:
            ws.step("zxw1", info.sigma0, 0.5, info.lambda_max_A, info.lambda_max_Q, False)
        assert ei.value.code == -4 and "not on this problem's path" in str(ei.value)
        with pytest.raises(_lib.HPRQPError):
            ws.step("lp_zx", info.sigma0, 0.5, info.lambda_max_A, info.lambda_max_Q, False)
        with pytest.raises(_lib.HPRQPError):
            ws.download(_lib.VEC_X_BAR, info.n + 1)
        other = _lib.Context(0)
        try:
            rc = ws.lib.hprqp_kkt_reduce(other.h, ws.h, 0, C.byref(_lib.hprqp_kkt_out()))
            assert rc == -4
        finally:
            other.close()
        # the workspace still works: one whole step, then the residuals
        ws.step("main", info.sigma0, 0.5, info.lambda_max_A, info.lambda_max_Q, True)
        assert np.isfinite(ws.kkt(1).kkt)
    finally:
        ws.close()
        dm.close()
