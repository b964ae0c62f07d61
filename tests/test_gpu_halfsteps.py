"""Each fused half-step kernel against the oracle's half-step on the same
state, partial and full write sets.

A workspace (hprqp_ws_*, scaling off so its data are the model's) is run for
a few steps to reach a non-trivial state; the whole state is read back
(hprqp_ws_download), one half-step runs on the GPU through its own entry
point (hprqp_step_*), and the oracle's function for that half-step
(oracle_halfstep: the reference's CPU loops, src/kernels.jl:2117-2832) runs on
the state read back before it.  Every vector the half-step writes must
agree to rounding (the products' summation order differs: 1e-11 of the
vector's scale); vectors it must not write are checked unchanged.  Paths:
standard (z/x/w1, tempv + y, w2), noC with a non-diagonal and with a
diagonal Q (z/x/w with the Halpern Qw update, then y), LASSO, and LP (z/x,
then y, whose partial step writes y_bar on the GPU as the reference's GPU
kernel does, not in the CPU loop)."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import fixture_path

pytestmark = pytest.mark.gpu

VEC = {"x_bar": 0, "y_bar": 1, "z_bar": 2, "w_bar": 3, "ATdy": 4, "x": 7, "w": 8, "y": 9, "ATy": 10, "Qw": 11,
       "x_hat": 12, "dx": 13, "dw": 14, "dy": 15, "s": 16, "ATy_bar": 17, "tempv": 18, "Qw_bar": 19,
       "last_x": 20, "last_y": 21, "last_w": 22, "last_ATy": 23, "last_Qw": 24, "fact1": 25, "fact2": 26, "fact": 27}
M_VECS = {"y", "y_bar", "dy", "s", "last_y"}


def _model(case):
    from hprqp import build_from_mps, build_from_QAbc, instances

    rng = np.random.default_rng(3)
    if case == "standard":
        return instances.random_sparse_qp(300, 150, seed=4), "QP"
    if case == "noc":
        n, m = 200, 60
        R = sp.random(40, n, density=0.05, random_state=5, format="csr")
        Q = (R.T @ R + sp.diags(rng.uniform(0.1, 1.0, n))).tocsr()
        A = sp.random(m, n, density=0.05, random_state=6, format="csr") + sp.eye(m, n)
        b = A @ rng.normal(size=n)
        l = np.full(n, -np.inf)
        u = np.full(n, np.inf)
        l[:10], u[:10] = -1.0, 1.0  # 95 % free: the noC path
        return build_from_QAbc(Q, rng.normal(size=n), A, b, b + 1.0, l, u, verbose=False), "QP"
    if case == "noc_diag":
        return build_from_mps(fixture_path("AUG2D.mps"), verbose=False), "QP"
    if case == "lasso":
        return instances.lasso(60, 500, mean_nnz=30, seed=2), "LASSO"
    n, m = 80, 30
    A = sp.random(m, n, density=0.2, random_state=7, format="csr") + sp.eye(m, n)
    b = A @ rng.uniform(0, 1, n)
    return build_from_QAbc(sp.csr_matrix((n, n)), rng.uniform(0.1, 1.0, n), A, b, b + 1.0, np.zeros(n),
                           np.full(n, 5.0), verbose=False), "QP"


# GPU entry points per oracle half-step, and the vectors each writes
# (partial, full)
STEPS = {
    "standard": [("zxw1", ["zxw1"], ["x", "x_hat", "w_bar", "Qw"], ["dx", "x_bar", "z_bar"]),
                 ("tempv_y", ["tempv", "y"], ["tempv", "y", "y_bar"], ["s", "dy"]),
                 ("w2", ["w2"], ["w", "ATy"], ["w_bar", "dw", "ATdy", "ATy_bar"])],
    "noc": [("noc_zxw", ["noc_zxw", "noc_tempv"], ["ATy", "x", "x_hat", "w", "w_bar", "Qw", "Qw_bar", "tempv"],
             ["dx", "x_bar", "z_bar", "dw"]),
            ("noc_y", ["y"], ["y", "y_bar"], ["s", "dy"])],
    "lasso": [("lasso", ["lasso"], ["x", "x_hat", "w", "w_bar", "Qw"], ["dx", "x_bar", "z_bar", "dw"])],
    "lp": [("lp_zx", ["lp_zx"], ["ATy", "x", "x_hat"], ["dx", "x_bar", "z_bar"]),
           ("lp_y", ["lp_y"], ["y"], ["s", "dy", "y_bar"])],
}
STEPS["noc_diag"] = STEPS["noc"]


@pytest.mark.parametrize("full", [False, True], ids=["partial", "full"])
@pytest.mark.parametrize("case", ["standard", "noc", "noc_diag", "lasso", "lp"])
def test_half_step_kernels_match_oracle(case, full):
    from hprqp import HPRQP_parameters
    from hprqp.solver import DeviceModel
    from hprqp.steploop import Workspace
    from oracle import oracle

    model, ptype = _model(case)
    params = HPRQP_parameters(verbose=False, problem_type=ptype, use_Ruiz_scaling=False,
                              use_Pock_Chambolle_scaling=False)
    dm = DeviceModel(model)
    ws = Workspace(dm, params)
    try:
        info = ws.info
        n, m = info.n, info.m
        assert (n, m) == (model.n, model.m)
        sigma, lamA, lamQ = info.sigma0, info.lambda_max_A, info.lambda_max_Q
        diag = bool(info.Q_is_diag)
        if diag:
            ws.update_Q_factors(sigma)
        for k in range(23):  # a state away from the start
            ws.step("main", sigma, 1.0 / (k + 2.0), lamA, lamQ, k % 5 == 4)
        h1 = 1.0 / 25.0

        def snapshot():
            return {name: ws.download(code, m if name in M_VECS else n) for name, code in VEC.items()}

        for oname, entries, writes, full_writes in STEPS[case]:
            before = snapshot()
            for e in entries:
                ws.step(e, sigma, h1, lamA, lamQ, full)
            after = snapshot()
            st = {k: before[k] if k in before else np.zeros(m if k in oracle.STATE_M else n)
                  for k in oracle.STATE_N + oracle.STATE_M}
            ref = oracle.halfstep(model, oname, st, sigma, lamA, lamQ, h1, full, Q_is_diag=diag,
                                  lam=getattr(model, "lam", 0.0))
            written = writes + (full_writes if full else [])
            for k in written:
                scale = 1.0 + np.max(np.abs(ref[k])) if ref[k].size else 1.0
                err = np.max(np.abs(after[k] - ref[k])) / scale if ref[k].size else 0.0
                print(f"{case} {oname} {'full' if full else 'partial'}: {k} max rel diff {err:.2e}")
                assert err < 1e-11, (oname, k, err)
            for k in VEC:
                if k in written:
                    continue
                if oname == "lp_y" and k == "y_bar":
                    # the GPU LP y step writes y_bar in partial steps too, as
                    # the reference's GPU kernel does (src/kernels.jl:858-859)
                    np.testing.assert_allclose(after[k], oracle.halfstep(model, oname, st, sigma, lamA, lamQ, h1,
                                                                         True, Q_is_diag=diag)[k], rtol=0,
                                               atol=1e-11 * (1 + np.max(np.abs(after[k]))))
                    continue
                np.testing.assert_array_equal(after[k], before[k], err_msg=f"{oname} changed {k}")
    finally:
        ws.close() if hasattr(ws, "close") else None
        dm.close()
