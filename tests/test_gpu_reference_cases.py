"""The reference's end-to-end testsets (test/runtests.jl) that the parity
suite does not already cover: parameter tolerances (:348-380), result fields
(:382-420), a dense Q given to build_from_QAbc (:283-313), the iteration and
time limits (status strings, src/algorithm.jl:558-576), every row-product
shape on both MPS fixtures (:704-788, the spmv_mode_Q x spmv_mode_A
combinations), and an invalid device number (:790-914; here it raises instead
of falling back to a CPU path)."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import fixture_path

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tol", [1e-4, 1e-6, 1e-8])
def test_parameter_tolerances(tol):
    from hprqp import HPRQP_parameters, build_from_mps, optimize

    for name in ("model.mps", "AUG2D.mps"):
        model = build_from_mps(fixture_path(name), verbose=False)
        res = optimize(model, HPRQP_parameters(verbose=False, stoptol=tol))
        print(f"{name} stoptol={tol:g}: {res.status} iter={res.iter} kkt={res.residuals:.2e}")
        assert res.status == "OPTIMAL"
        assert res.residuals < tol


def test_tighter_tolerance_takes_no_fewer_iterations():
    from hprqp import HPRQP_parameters, build_from_mps, optimize

    model = build_from_mps(fixture_path("AUG2D.mps"), verbose=False)
    its = [optimize(model, HPRQP_parameters(verbose=False, stoptol=t)).iter for t in (1e-4, 1e-6, 1e-8)]
    assert its[0] <= its[1] <= its[2], its


def test_result_fields():
    from hprqp import HPRQP_parameters, build_from_mps, optimize

    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    res = optimize(model, HPRQP_parameters(verbose=False))
    assert res.status == "OPTIMAL"
    assert res.x.shape == (model.n,) and res.z.shape == (model.n,) and res.w.shape == (model.n,)
    assert res.y.shape == (model.m,)
    assert 0 <= res.iter_4 <= res.iter_6 <= res.iter
    assert 0.0 <= res.power_time <= res.time and res.time_4 <= res.time_6 <= res.time + 1e-12
    assert np.isfinite(res.primal_obj) and np.isfinite(res.gap) and res.residuals < 1e-6
    # closed form of model.mps: x = (1.5, 2.5), objective -8.5
    np.testing.assert_allclose(res.x, [1.5, 2.5], atol=1e-5)
    assert abs(res.primal_obj + 8.5) < 1e-5


def test_dense_Q_matches_sparse_Q():
    from hprqp import HPRQP_parameters, build_from_QAbc, optimize

    rng = np.random.default_rng(7)
    n, m = 40, 15
    R = rng.standard_normal((n, n))
    Q = R.T @ R / n + 0.1 * np.eye(n)
    A = rng.standard_normal((m, n))
    x0 = rng.uniform(0.0, 1.0, n)
    b = A @ x0
    c = rng.standard_normal(n)
    l, u = np.zeros(n), np.full(n, 2.0)
    dense = optimize(build_from_QAbc(Q, c, A, b - 0.1, b + 0.1, l, u, verbose=False), HPRQP_parameters(verbose=False))
    sparse = optimize(build_from_QAbc(sp.csc_matrix(Q), c, sp.csr_matrix(A), b - 0.1, b + 0.1, l, u, verbose=False),
                      HPRQP_parameters(verbose=False))
    assert dense.status == sparse.status == "OPTIMAL"
    assert dense.iter == sparse.iter
    assert abs(dense.primal_obj - sparse.primal_obj) <= 1e-12 * (1 + abs(sparse.primal_obj))


def test_iteration_and_time_limits():
    from hprqp import HPRQP_parameters, build_from_mps, optimize

    model = build_from_mps(fixture_path("AUG2D.mps"), verbose=False)
    res = optimize(model, HPRQP_parameters(verbose=False, max_iter=25))
    assert res.status == "MAX_ITER" and res.iter == 25
    res = optimize(model, HPRQP_parameters(verbose=False, time_limit=1e-9))
    assert res.status == "TIME_LIMIT"


@pytest.mark.parametrize("lanes", [1, 64, 256, -256])
@pytest.mark.parametrize("name", ["AUG2D.mps", "AUG3DCQP.mps"])
def test_fixtures_under_every_row_product_shape(name, lanes):
    from hprqp import HPRQP_parameters, build_from_mps, optimize

    model = build_from_mps(fixture_path(name), verbose=False)
    ref = optimize(model, HPRQP_parameters(verbose=False))
    res = optimize(model, HPRQP_parameters(verbose=False, spmv_lanes=lanes))
    assert res.status == "OPTIMAL" and res.residuals < 1e-6
    assert abs(res.primal_obj - ref.primal_obj) <= 1e-6 * (1 + abs(ref.primal_obj))
    step = 10 if ref.iter < 1000 else 100
    assert abs(res.iter - ref.iter) <= max(0.1 * ref.iter, step)


def test_invalid_device_number_raises():
    from hprqp import HPRQP_parameters, build_from_mps, optimize
    from hprqp._lib import HPRQPError

    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    with pytest.raises((HPRQPError, ValueError, RuntimeError)):
        optimize(model, HPRQP_parameters(verbose=False, device_number=999))
