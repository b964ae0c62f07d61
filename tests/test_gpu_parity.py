"""GPU parity: libhprqp (FP64 plain form) against the CPU oracle.

Bar (from the reference's CPU path, restated by the oracle): status OPTIMAL,
relative KKT < 1e-6, objective within 1e-6 relative, iteration count within
10% (at least one residual-check step, since termination is only tested at
check/print iterations: every 10 iterations below 1000, every 100 below
10^4).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import fixture_path

pytestmark = pytest.mark.gpu


def _iter_ok(a, b):
    step = 10 if b < 1000 else (100 if b < 10_000 else 1000)
    return abs(a - b) <= max(0.10 * b, step)


def _solve_both(model, problem_type="QP", **kw):
    from hprqp import HPRQP_parameters, optimize
    from oracle import oracle

    params = HPRQP_parameters(verbose=False, problem_type=problem_type, **kw)
    res = optimize(model, params)
    ref = oracle.solve(model, problem_type=problem_type,
                       **{k: v for k, v in kw.items() if k in ("max_iter", "stoptol", "check_iter")})
    print(f"gpu: {res.status} iter={res.iter} obj={res.primal_obj:.12e} kkt={res.residuals:.3e} "
          f"lamA={res.lambda_max_A:.6e} lamQ={res.lambda_max_Q:.6e} restarts={res.restarts}")
    print(f"cpu: {ref.status} iter={ref.iter} obj={ref.primal_obj:.12e} kkt={ref.residuals:.3e} "
          f"lamA={ref.lambda_max_A:.6e} lamQ={ref.lambda_max_Q:.6e} restarts={ref.restarts}")
    return res, ref


def _check(res, ref, obj_tol=1e-6):
    assert res.status == "OPTIMAL"
    assert res.residuals < 1e-6
    assert ref.status == "OPTIMAL"
    rel = abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj))
    assert rel < obj_tol, rel
    assert _iter_ok(res.iter, ref.iter), (res.iter, ref.iter)
    assert res.Q_is_diag == ref.Q_is_diag
    assert res.noC == ref.noC
    # the power method is seeded identically: lambda_max agrees to its tolerance
    assert abs(res.lambda_max_A - ref.lambda_max_A) <= 1e-3 * max(1.0, ref.lambda_max_A)
    assert abs(res.lambda_max_Q - ref.lambda_max_Q) <= 1e-3 * max(1.0, ref.lambda_max_Q)
    # and it stops on the same iteration (one either way for an error within
    # rounding of the tolerance)
    print(f"power iterations: gpu A {res.power_iter_A} Q {res.power_iter_Q}; "
          f"cpu A {ref.power_iter_A} Q {ref.power_iter_Q}")
    assert abs(res.power_iter_A - ref.power_iter_A) <= 1, (res.power_iter_A, ref.power_iter_A)
    assert abs(res.power_iter_Q - ref.power_iter_Q) <= 1, (res.power_iter_Q, ref.power_iter_Q)


def test_model_mps():
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    res, ref = _solve_both(model)
    _check(res, ref)
    assert abs(res.primal_obj - (-8.5)) < 1e-5
    np.testing.assert_allclose(res.x, [1.5, 2.5], atol=1e-4)


def test_aug2d_noC_diagonal_Q():
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("AUG2D.mps"), verbose=False)
    res, ref = _solve_both(model)
    _check(res, ref)
    assert res.noC and res.Q_is_diag
    # known optimum from an equality-KKT solve (and Maros-Meszaros 1.6874118e+06)
    assert abs(res.primal_obj - 1687411.752897) / 1687411.752897 < 1e-6


def test_aug3dcqp_diagonal_Q():
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("AUG3DCQP.mps"), verbose=False)
    res, ref = _solve_both(model)
    _check(res, ref)


def test_random_sparse_qp_small():
    from hprqp import instances

    model = instances.random_sparse_qp(2000, 1000, seed=0)
    res, ref = _solve_both(model)
    _check(res, ref)
    assert not res.Q_is_diag and not res.noC


def test_grid_qp_small():
    from hprqp import instances

    model = instances.grid_qp(60, seed=0)
    res, ref = _solve_both(model)
    _check(res, ref)


@pytest.mark.parametrize("lanes", [1, 4, 64, 256, -64, -256])
def test_every_spmv_kernel_shape_gives_the_same_solution(lanes):
    """Vector kernels (1..64 lanes per row) and CSR-stream kernels (rows per
    block) change only the summation order inside a row."""
    from hprqp import instances

    model = instances.random_sparse_qp(1500, 700, seed=11)
    res, ref = _solve_both(model, spmv_lanes=lanes)
    _check(res, ref)


@pytest.mark.parametrize("case", ["grid", "aug2d", "aug3dcqp", "random"])
def test_value_table_form_matches_oracle(case):
    """The FP64 iteration with Q, A and A^T in value-table form (one byte per
    nonzero into at most 256 unscaled values, the scaling as row and column
    factors), forced: the oracle's solution, iterations and restarts; a
    matrix with more than 256 distinct values keeps its plain shape (the
    random QP's Q), and stopped part-way the iterates are the plain form's
    to rounding."""
    from hprqp import HPRQP_parameters, build_from_mps, instances, optimize

    model = {"grid": lambda: instances.grid_qp(60, seed=0),
             "aug2d": lambda: build_from_mps(fixture_path("AUG2D.mps"), verbose=False),
             "aug3dcqp": lambda: build_from_mps(fixture_path("AUG3DCQP.mps"), verbose=False),
             "random": lambda: instances.random_sparse_qp(1500, 700, seed=11)}[case]()
    res, ref = _solve_both(model, spmv_lanes=-1256)
    print(f"{case}: shapes {res.lanes}")
    _check(res, ref)
    assert abs(res.iter - ref.iter) <= 10 and res.restarts == ref.restarts
    if case == "grid":
        assert res.lanes[0] == res.lanes[1] == res.lanes[2] == -1256
    if case == "random":
        assert res.lanes[0] != -1256
    its = 200
    a = optimize(model, HPRQP_parameters(verbose=False, max_iter=its, spmv_lanes=-1256))
    b = optimize(model, HPRQP_parameters(verbose=False, max_iter=its, spmv_lanes=-256))
    dx = np.max(np.abs(a.x - b.x)) / (1 + np.max(np.abs(b.x)))
    print(f"{case}: stopped at {its}: x diff value-table vs plain {dx:.2e}")
    assert a.restarts == b.restarts and dx < 1e-9


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
@pytest.mark.parametrize("case", ["grid", "random", "aug3dcqp", "lasso", "lasso_chunked", "lasso_wide"])
def test_reduced_precision_reaches_fp64_accuracy(precision, case):
    """Mixed (FP32 matrices and data, FP64 iterates) and FP32 (+FP64 polish)
    forms: the KKT residual is evaluated in FP64 against the FP64 data, so
    the same 1e-6 relative KKT bar applies; the objective must agree with
    the FP64 oracle within 1e-6 relative.  lasso_chunked forces the
    column-chunked row product (FP32 values on FP64 vectors, and all-FP32).
    lasso_wide (95 x 7538) drives sigma to ~1e5, where FP32 iterates drift
    off the solution: both reduced precisions must converge within twice
    the oracle's iterations (LASSO runs the FP32-data, FP64-iterate form)."""
    from hprqp import HPRQP_parameters, build_from_mps, instances, optimize
    from oracle import oracle

    ptype = "QP"
    if case == "grid":
        model = instances.grid_qp(80, seed=2)
    elif case == "random":
        model = instances.random_sparse_qp(2000, 1000, seed=4)
    elif case == "aug3dcqp":
        model = build_from_mps(fixture_path("AUG3DCQP.mps"), verbose=False)
    elif case == "lasso":
        model = instances.lasso(200, 4000, mean_nnz=60, seed=1)
        ptype = "LASSO"
    elif case == "lasso_wide":
        model = instances.lasso(95, 7538, mean_nnz=55.0, seed=281397143)
        ptype = "LASSO"
    else:
        model = instances.lasso(120, 30_000, mean_nnz=500, seed=3)
        ptype = "LASSO"
    lanes = 1 << 20 if case == "lasso_chunked" else 0
    res = optimize(model, HPRQP_parameters(verbose=False, precision=precision, problem_type=ptype, spmv_lanes=lanes))
    ref = oracle.solve(model, problem_type=ptype)
    rel = abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj))
    print(f"{precision} {case}: {res.status} iter={res.iter} polish={res.polish_iter} kkt={res.residuals:.2e} "
          f"rel_obj_err_vs_fp64={rel:.2e} (oracle iter={ref.iter})")
    assert res.status == "OPTIMAL" and res.residuals < 1e-6
    assert rel < 1e-6
    if case == "lasso_wide":
        assert res.iter <= 2 * ref.iter


def test_lasso_small():
    from hprqp import instances

    model = instances.lasso(200, 4000, mean_nnz=60, seed=1)
    res, ref = _solve_both(model, problem_type="LASSO")
    _check(res, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lam_factor", [1.01, 0.995])
def test_lasso_closed_form_at_the_zero_threshold(lam_factor, precision):
    """lam just above ||A^T b||_inf: x = 0; just below: one nonzero with a
    closed form.  The soft-threshold prox must land on exact zeros."""
    from lasso_closed_form import closed_form

    from hprqp import instances

    model = instances.lasso(60, 900, mean_nnz=40, seed=7, lam_factor=lam_factor)
    x, f = closed_form(model.Q.A, model.Q.b, model.lam)
    res, ref = _solve_both(model, problem_type="LASSO", precision=precision)
    if precision == "fp64":
        _check(res, ref)
    assert res.status == "OPTIMAL" and res.residuals < 1e-6
    assert abs(res.primal_obj - f) / (1 + abs(f)) < 1e-6
    np.testing.assert_allclose(res.x, x, atol=1e-4)
    assert np.count_nonzero(res.x) == np.count_nonzero(x)


def test_lasso_long_rows_column_chunked():
    """Rows of ~500 nonzeros over 30,000 columns (eight column chunks, the
    last ragged; log-normal row lengths, so some rows pass 64 nonzeros in a
    chunk and some are empty there) with the column-chunked row product forced
    for every matrix: parity with the oracle, and the same solve as the
    autotuned shapes."""
    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.lasso(120, 30_000, mean_nnz=500, seed=3)
    res, ref = _solve_both(model, problem_type="LASSO", spmv_lanes=1 << 20)
    _check(res, ref)
    assert res.lanes[0] == 1 << 20
    auto = optimize(model, HPRQP_parameters(verbose=False, problem_type="LASSO"))
    rel = abs(res.primal_obj - auto.primal_obj) / (1.0 + abs(auto.primal_obj))
    print(f"autotuned shapes {auto.lanes}: iter={auto.iter} rel obj {rel:.2e}")
    assert rel < 1e-9 and abs(res.iter - auto.iter) <= 10


def test_qap_esc64a():
    from hprqp import build_from_mat

    model = build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False)
    res, ref = _solve_both(model, problem_type="QAP")
    _check(res, ref)


def test_lp_empty_Q():
    from hprqp import build_from_QAbc

    rng = np.random.default_rng(3)
    n, m = 60, 30
    A = sp.random(m, n, density=0.2, random_state=4, format="csr") + sp.eye(m, n)
    x0 = rng.uniform(0, 1, n)
    b = A @ x0
    c = rng.uniform(0.1, 1.0, n)
    model = build_from_QAbc(sp.csr_matrix((n, n)), c, A, b, b + 1.0, np.zeros(n), np.full(n, 5.0), verbose=False)
    res, ref = _solve_both(model)
    _check(res, ref)


def _box_qp_no_constraints():
    from hprqp import build_from_QAbc

    rng = np.random.default_rng(5)
    n = 80
    R = sp.random(40, n, density=0.1, random_state=6, format="csr")
    Q = (R.T @ R + sp.diags(rng.uniform(0.01, 1.0, n))).tocsr()
    c = rng.normal(size=n)
    return build_from_QAbc(Q, c, sp.csr_matrix((0, n)), np.zeros(0), np.zeros(0),
                           np.full(n, -1.0), np.full(n, 1.0), verbose=False)


def test_box_qp_without_constraints():
    """m = 0: the y half-step, the A/Aᵀ products and the primal residual are
    skipped (src/kernels.jl:2246, 2586, 2686); only the box projection and
    Q remain."""
    model = _box_qp_no_constraints()
    assert model.m == 0
    res, ref = _solve_both(model)
    _check(res, ref)
    assert res.err_Rp == 0.0
    np.testing.assert_allclose(res.x, ref.x, atol=1e-4)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_box_qp_without_constraints_reduced_precision(precision):
    from hprqp import HPRQP_parameters, optimize
    from oracle import oracle

    model = _box_qp_no_constraints()
    res = optimize(model, HPRQP_parameters(verbose=False, precision=precision))
    ref = oracle.solve(model)
    rel = abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj))
    print(f"{precision} box QP m=0: {res.status} iter={res.iter} kkt={res.residuals:.2e} rel_obj_err={rel:.2e}")
    assert res.status == "OPTIMAL" and res.residuals < 1e-6
    assert rel < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32"])
def test_every_bound_kind(precision):
    """Free, lower-only, upper-only, boxed and fixed variables with =, <=, >=
    and ranged rows in one instance: the projections onto [l, u] and
    [AL, AU] with infinite ends, against the oracle and against the exact
    minimiser on the active set."""
    from bounds_case import active_set_certificate, every_bound_kind_qp

    from hprqp import build_from_QAbc

    Q, c, A, AL, AU, l, u = every_bound_kind_qp()
    model = build_from_QAbc(Q, c, A, AL, AU, l, u, verbose=False)
    res, ref = _solve_both(model, precision=precision)
    if precision == "fp64":
        _check(res, ref)
    else:
        assert res.status == "OPTIMAL" and res.residuals < 1e-6
        assert abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj)) < 1e-6
    xs, ok = active_set_certificate(Q, c, A, AL, AU, l, u, res.x, tol=1e-4)
    assert ok
    fstar = 0.5 * xs @ (Q @ xs) + c @ xs
    assert abs(res.primal_obj - fstar) / (1 + abs(fstar)) < 1e-6
    np.testing.assert_allclose(res.x, xs, atol=1e-3)


@pytest.mark.parametrize("lanes", [0, 1, 64, -256])
def test_partially_quadratic_empty_Q_rows(lanes):
    """Q with 200 empty rows out of 400 (variables that enter only linearly)
    under the autotuned and forced row-product shapes."""
    from bounds_case import partially_quadratic_qp

    from hprqp import build_from_QAbc

    Q, c, A, AL, AU, l, u = partially_quadratic_qp()
    model = build_from_QAbc(Q, c, A, AL, AU, l, u, verbose=False)
    res, ref = _solve_both(model, spmv_lanes=lanes)
    _check(res, ref)
    assert not res.Q_is_diag


def test_bound_edges_mps():
    """An MPS file with every RANGES sign and bound type (tests/golden/
    bound_edges.mps), read, formulated and solved: against the oracle and the
    hand-derived optimum -21/16 (x5 and x7 are not unique)."""
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("bound_edges.mps"), verbose=False)
    res, ref = _solve_both(model)
    _check(res, ref)
    assert abs(res.primal_obj - (-21 / 16)) < 1e-5
    np.testing.assert_allclose(res.x[[0, 1, 2, 3, 5, 7]], [-1, 1.75, 2.5, -0.75, 3.25, 0], atol=1e-4)


def test_portfolio_small():
    from hprqp import instances

    model, _ = instances.portfolio(300, 10, seed=0)
    res, ref = _solve_both(model)
    _check(res, ref)


def test_warm_start_at_optimum_gives_iter_0():
    """test/runtests.jl:209-246: x and y warm start at the optimum -> iter == 0."""
    from hprqp import HPRQP_parameters, build_from_mps, optimize

    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    cold = optimize(model, HPRQP_parameters(verbose=False))
    warm = optimize(model, HPRQP_parameters(verbose=False, initial_x=cold.x, initial_y=cold.y))
    assert warm.status == "OPTIMAL"
    assert warm.iter == 0
    xonly = optimize(model, HPRQP_parameters(verbose=False, initial_x=cold.x))
    assert xonly.status == "OPTIMAL" and xonly.iter <= cold.iter


@pytest.mark.parametrize("start", ["x_only", "y_only", "perturbed_xy"])
def test_warm_start_matches_oracle(start):
    """The reference's three warm starts (test/runtests.jl:190-245), each
    given to the GPU solve and to the oracle: the starts are mapped into the
    scaled space the same way, so the runs agree in iterations and
    objective.  Perturbed starts keep the run away from iteration 0."""
    from hprqp import HPRQP_parameters, instances, optimize
    from oracle import oracle

    model = instances.random_sparse_qp(1500, 700, seed=17)
    cold = oracle.solve(model)
    rng = np.random.default_rng(1)
    x0 = cold.x + 1e-2 * rng.standard_normal(model.n)
    y0 = cold.y + 1e-2 * rng.standard_normal(model.m)
    kw = {"x_only": dict(initial_x=x0), "y_only": dict(initial_y=y0),
          "perturbed_xy": dict(initial_x=x0, initial_y=y0)}[start]
    res = optimize(model, HPRQP_parameters(verbose=False, **kw))
    ref = oracle.solve(model, **kw)
    print(f"{start}: gpu iter={res.iter} obj={res.primal_obj:.12e}; oracle iter={ref.iter} "
          f"obj={ref.primal_obj:.12e}; cold iter={cold.iter}")
    _check(res, ref)
    if start == "perturbed_xy":
        assert res.iter < cold.iter


def test_noC_warm_start_with_initial_x_leaves_Qw_unset():
    """The noC path with initial_x: process_initial_points!
    sets x and w from initial_x but not Qw / last_Qw (src/algorithm.jl:1797-1802),
    which the noC step carries by its Halpern recurrence from the last restart
    (:544-547), so the run starts from Qw = 0 until its first restart.  The GPU
    reproduces that: on AUG2D (noC, diagonal Q) warm-started from a perturbed
    optimum with x only, it takes the oracle's iterations and restarts, and
    stopped part-way, before the first restart, its iterates are the
    oracle's."""
    from hprqp import HPRQP_parameters, build_from_mps, optimize
    from oracle import oracle

    model = build_from_mps(fixture_path("AUG2D.mps"), verbose=False)
    cold = oracle.solve(model)
    rng = np.random.default_rng(11)
    x0 = cold.x + 1e-2 * rng.standard_normal(model.n)
    res = optimize(model, HPRQP_parameters(verbose=False, initial_x=x0))
    ref = oracle.solve(model, initial_x=x0)
    print(f"noC + initial_x: gpu iter={res.iter} restarts={res.restarts} obj={res.primal_obj:.12e}; oracle "
          f"iter={ref.iter} restarts={ref.restarts} obj={ref.primal_obj:.12e}; cold {cold.iter}")
    assert res.noC and ref.noC and res.Q_is_diag
    _check(res, ref)
    assert res.iter == ref.iter and res.restarts == ref.restarts
    its = 60
    part = optimize(model, HPRQP_parameters(verbose=False, initial_x=x0, max_iter=its))
    rpart = oracle.solve(model, initial_x=x0, max_iter=its)
    dx = np.max(np.abs(part.x - rpart.x)) / (1 + np.max(np.abs(rpart.x)))
    dy = np.max(np.abs(part.y - rpart.y)) / (1 + np.max(np.abs(rpart.y)))
    print(f"stopped at {its}: x diff {dx:.2e}, y diff {dy:.2e}, restarts {part.restarts} / {rpart.restarts}")
    assert part.restarts == rpart.restarts == 0 and dx < 1e-9 and dy < 1e-9


@pytest.mark.parametrize("case", ["random", "grid", "lasso", "qap", "aug2d"])
def test_trajectory_matches_oracle_mid_solve(case):
    """Stopped part-way (max_iter = 300, before convergence), the GPU's
    iterates equal the oracle's to rounding: the same steps, restarts and
    sigma updates, not only the same end point.  Summation order differs
    (row products, reductions), so rounding drifts over 300 iterations:
    measured up to 1.3e-9 relative (random, grid) and 2e-14 (LASSO); a wrong
    term or sign moves the iterates by orders of magnitude more.  AUG2D
    (diagonal Q, n = 20,200) takes its sigma updates from the device's
    log-grid search, the oracle from the reference's golden-section search
    on the host: the same minimiser to 1e-13."""
    from hprqp import HPRQP_parameters, build_from_mat, build_from_mps, instances, optimize
    from oracle import oracle

    ptype = {"lasso": "LASSO", "qap": "QAP"}.get(case, "QP")
    model = {"random": lambda: instances.random_sparse_qp(2000, 1000, seed=0),
             "grid": lambda: instances.grid_qp(60, seed=0),
             "lasso": lambda: instances.lasso(200, 4000, mean_nnz=60, seed=1),
             "qap": lambda: build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False),
             "aug2d": lambda: build_from_mps(fixture_path("AUG2D.mps"), verbose=False)}[case]()
    # esc64a is within 2% of the tolerance by iteration 300: stop it earlier
    its = 200 if case == "qap" else 300
    res = optimize(model, HPRQP_parameters(verbose=False, max_iter=its, problem_type=ptype))
    ref = oracle.solve(model, max_iter=its, problem_type=ptype)
    assert res.status == ref.status == "MAX_ITER" and res.iter == ref.iter == its
    assert res.restarts == ref.restarts
    dx = np.max(np.abs(res.x - ref.x)) / (1 + np.max(np.abs(ref.x)))
    dy = np.max(np.abs(res.y - ref.y)) / (1 + np.max(np.abs(ref.y))) if model.m else 0.0
    print(f"{case}: restarts {res.restarts}, x diff {dx:.2e}, y diff {dy:.2e}, kkt {res.residuals:.3e} / {ref.residuals:.3e}")
    assert dx < 1e-7 and dy < 1e-7
    assert abs(res.residuals - ref.residuals) <= 1e-5 * ref.residuals


def test_csr_spmv_matches_scipy():
    import ctypes as C

    from hprqp import _lib

    ctx = _lib.context(0)
    rng = np.random.default_rng(0)
    for (m, n, dens) in ((1, 1, 1.0), (37, 53, 0.2), (1000, 700, 0.01), (300, 5000, 0.3)):
        M = sp.random(m, n, density=dens, random_state=rng, format="csr")
        M.sort_indices()
        rp = M.indptr.astype(np.int32)
        ci = M.indices.astype(np.int32)
        v = M.data.astype(np.float64)
        h = C.c_void_p()
        _lib.check(ctx.h, ctx.lib.hprqp_csr_create(ctx.h, m, n, M.nnz, _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(v), 0,
                                                   1, C.byref(h)))
        x = rng.standard_normal(n)
        yv = rng.standard_normal(m)
        xb = _lib.DeviceBuffer.from_array(ctx, x)
        yb = _lib.DeviceBuffer(ctx, m)
        _lib.check(ctx.h, ctx.lib.hprqp_csr_spmv(ctx.h, h, 0, xb.h, yb.h))
        np.testing.assert_allclose(yb.download(), M @ x, rtol=1e-12, atol=1e-12)
        yb2 = _lib.DeviceBuffer.from_array(ctx, yv)
        xo = _lib.DeviceBuffer(ctx, n)
        _lib.check(ctx.h, ctx.lib.hprqp_csr_spmv(ctx.h, h, 1, yb2.h, xo.h))
        np.testing.assert_allclose(xo.download(), M.T @ yv, rtol=1e-12, atol=1e-12)
        ctx.lib.hprqp_csr_destroy(ctx.h, h)


def test_qop_apply_matches_oracle():
    from hprqp import build_from_mat, instances, to_gpu
    from oracle import oracle

    rng = np.random.default_rng(1)
    qap = build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False)
    las = instances.lasso(50, 800, mean_nnz=30, seed=2)
    rq = instances.random_sparse_qp(500, 100, seed=3)
    for model in (qap, las, rq):
        x = rng.standard_normal(model.n)
        dev = to_gpu(model.Q, model.n)
        got = dev.Qmap(x)
        want = oracle.qmap(model.Q, x)
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-9 * np.max(np.abs(want)))
        n, t, name = dev.info()
        assert n == model.n
        dev.close()


def _qap_dense(Q, X):
    return 2.0 * (Q.A @ X @ Q.B - Q.S @ X - X @ Q.T)


def _qap_cases():
    """QAP operators: esc64a (full S, T; integer data), the 128 x 128
    flow/distance instance (diagonal S, T; F/2 and D exact in BF16) and two
    sizes that are not multiples of the 16 x 16 tile with Gaussian (not
    BF16-exact) matrices, one with full and one with diagonal shifts."""
    from hprqp import build_from_mat, instances
    from hprqp.model import QAP_Q_operator

    rng = np.random.default_rng(11)
    out = [("esc64a", build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False).Q),
           ("flow128", instances.qap(128, seed=0)[0].Q)]
    for nq, full in ((45, True), (70, False)):
        A = rng.standard_normal((nq, nq))
        B = rng.standard_normal((nq, nq))
        if full:
            S, T = rng.standard_normal((nq, nq)), rng.standard_normal((nq, nq))
        else:
            S, T = np.diag(rng.standard_normal(nq)), np.diag(rng.standard_normal(nq))
        out.append((f"gauss{nq}", QAP_Q_operator(A, B, S, T)))
    return out


def test_qap_fp64_mfma_matches_closed_form():
    """FP64 MFMA operator (f64 16x16x4) against vec(2(AXB - SX - XT))."""
    from hprqp import to_gpu

    rng = np.random.default_rng(2)
    for name, Q in _qap_cases():
        nq = Q.A.shape[0]
        X = rng.standard_normal((nq, nq))
        dev = to_gpu(Q, nq * nq)
        got = dev.Qmap(X.ravel(order="F"))
        want = _qap_dense(Q, X).ravel(order="F")
        scale = np.max(np.abs(want))
        err = np.max(np.abs(got - want)) / scale
        print(f"{name}: fp64 max rel err {err:.2e}")
        assert err < 1e-13
        dev.close()


def test_qap_split_bf16x3_accuracy():
    """Split-BF16 three-term MFMA product (FP32 accumulation) against the FP64
    closed form, error relative to the magnitude M = 2(|A||X||B| + |S||X| +
    |X||T|): BF16-exact coefficients (three-piece data split) reach FP32
    accuracy; general coefficients (two-piece splits of both) 16 bits."""
    from hprqp import to_gpu

    rng = np.random.default_rng(3)
    for name, Q in _qap_cases():
        nq = Q.A.shape[0]
        X = rng.uniform(0.0, 1.0, (nq, nq)) / nq
        dev = to_gpu(Q, nq * nq)
        got = dev.Qmap(X.ravel(order="F"), precision="mixed")
        want = _qap_dense(Q, X).ravel(order="F")
        aX = np.abs(X)
        M = (2.0 * (np.abs(Q.A) @ aX @ np.abs(Q.B) + np.abs(Q.S) @ aX + aX @ np.abs(Q.T))).ravel(order="F")
        err = np.max(np.abs(got - want) / M)
        exact = name in ("esc64a", "flow128")
        print(f"{name}: split-bf16x3 max err/M {err:.2e} ({'exact' if exact else 'general'} coefficients)")
        assert err < (1e-5 if exact else 2e-4)
        dev.close()


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_qap_reduced_precision_solve(precision):
    """QAP relaxations with the split-BF16 x3 operator in the iteration: the
    FP64 KKT residual still decides termination (1e-6) and the objective
    agrees with the FP64 oracle within 1e-6 relative."""
    from hprqp import HPRQP_parameters, build_from_mat, instances, optimize
    from oracle import oracle

    for model in (build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False), instances.qap(48, seed=1)[0]):
        res = optimize(model, HPRQP_parameters(verbose=False, precision=precision, problem_type="QAP"))
        ref = oracle.solve(model, problem_type="QAP")
        rel = abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj))
        print(f"{precision} QAP n={model.n}: {res.status} iter={res.iter} polish={res.polish_iter} "
              f"kkt={res.residuals:.2e} rel_obj_err_vs_fp64={rel:.2e} (oracle iter={ref.iter})")
        assert res.status == "OPTIMAL" and res.residuals < 1e-6
        assert rel < 1e-6


@pytest.mark.parametrize("lanes", [1 << 20, 0])
def test_empty_A_rows_with_the_column_chunked_shape(lanes):
    """A model handed to the C ABI with m > 0 rows and no nonzeros in A (the
    host package's qp_formulation would delete such rows; hprqp_model_create
    does not).  The column-chunked shape loads entry 0 of the value and index
    arrays speculatively, and an empty matrix has none, so an empty matrix
    takes the one-lane shape whatever shape is forced.  The reference's power
    method gives NaN for lambda_max(A A^T) = 0 (0/0 in its error test), and so
    do the oracle and this solve; A^T y stays exactly 0, so x follows the box
    QP's trajectory, which the oracle's matches for the 20 iterations before
    the first sigma update."""
    from hprqp import HPRQP_parameters
    from hprqp.model import QP_info_cpu
    from hprqp.solver import DeviceModel
    from oracle import oracle

    n, rng = 40, np.random.default_rng(0)
    R = sp.random(20, n, density=0.2, random_state=1, format="csr")
    model = QP_info_cpu(Q=(R.T @ R + sp.identity(n)).tocsr(), c=rng.normal(size=n), A=sp.csc_matrix((3, n)),
                        AL=-np.ones(3), AU=np.ones(3), l=-np.ones(n), u=np.ones(n))
    dm = DeviceModel(model)
    try:
        res = dm.solve(HPRQP_parameters(verbose=False, max_iter=20, spmv_lanes=lanes))
    finally:
        dm.close()
    ref = oracle.solve(model, max_iter=20)
    print(f"lanes {lanes}: gpu {res.status} iter={res.iter} lamA={res.lambda_max_A} x[:3]={res.x[:3]}; "
          f"oracle {ref.status} iter={ref.iter} lamA={ref.lambda_max_A}")
    assert res.status == ref.status == "MAX_ITER" and res.iter == ref.iter == 20
    assert np.isnan(res.lambda_max_A) and np.isnan(ref.lambda_max_A)
    assert res.power_iter_A == ref.power_iter_A == 5000
    assert np.all(np.isfinite(res.x))
    np.testing.assert_allclose(res.x, ref.x, rtol=0, atol=1e-12)


def test_history_and_stop_at_a_print_step():
    """Print steps between two check iterations have their residuals read back
    one iteration late (only the termination test acts on them).  A solve
    that stops on such a step (660, below 1,000 where every tenth iteration
    is a print step) must stop there with that step's iterate, and the
    residual history -- every print and check step, in order -- must match
    the oracle's row for row."""
    from hprqp import HPRQP_parameters, instances, optimize
    from oracle import oracle

    model = instances.random_sparse_qp(400, 200, seed=0)
    res = optimize(model, HPRQP_parameters(verbose=False))
    ref = oracle.solve(model)
    print(f"gpu iter={res.iter} rows={len(res.history)}; cpu iter={ref.iter} rows={len(ref.history)}")
    assert res.iter == ref.iter and res.iter % 100 != 0 and res.iter < 1000
    _check(res, ref)
    h, g = res.history, ref.history
    assert h.shape == g.shape
    np.testing.assert_array_equal(h[:, 0], g[:, 0])
    np.testing.assert_array_equal(h[:, 7], g[:, 7])
    for col in (1, 2, 3):
        np.testing.assert_allclose(h[:, col], g[:, col], rtol=1e-6, atol=1e-12)
    # sigma: golden-section minimisers of functions formed from GPU and CPU
    # dot products (3e-8 apart)
    np.testing.assert_allclose(h[:, 6], g[:, 6], rtol=1e-6)
    np.testing.assert_allclose(res.x, ref.x, atol=1e-8 * (1 + np.max(np.abs(ref.x))))
    np.testing.assert_allclose(res.y, ref.y, atol=1e-8 * (1 + np.max(np.abs(ref.y))))
