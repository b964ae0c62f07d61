"""Pinning the CPU oracle (oracle/hprqp_oracle.c) against independent checks.

Each oracle function is checked against something that does not retype its
formula: scipy/numpy (SpMV, dense products, eigenvalues, scalar minimisation),
closed-form optima, brute-force active-set enumeration on tiny QPs, KKT
conditions recomputed from the unscaled solution, and the reference's own
fixtures and known answers (model.mps, AUG2D, AUG3DCQP, esc64a).

Unpinned: the exact iteration count/restart sequence of the Julia reference
(its power method starts from Julia's MersenneTwister(1) randn stream, and
Julia is not available), so iteration parity is judged against this oracle.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import minimize_scalar

from conftest import fixture_path


def _oracle():
    from oracle import oracle

    return oracle


# ---------------------------------------------------------------- kernels
def test_csc_mul_matches_scipy():
    o = _oracle()
    rng = np.random.default_rng(0)
    for (m, n, d) in ((1, 1, 1.0), (7, 13, 0.3), (200, 90, 0.05), (3, 400, 0.5)):
        M = sp.random(m, n, density=d, random_state=rng, format="csc")
        x = rng.standard_normal(n)
        np.testing.assert_allclose(o.csc_mul(M, x), M @ x, rtol=1e-13, atol=1e-13)


def test_qmap_sparse_lasso_qap_match_closed_forms():
    from hprqp import build_from_mat, instances

    o = _oracle()
    rng = np.random.default_rng(1)
    Q = sp.random(50, 50, density=0.2, random_state=rng)
    Q = (Q + Q.T).tocsc()
    x = rng.standard_normal(50)
    np.testing.assert_allclose(o.qmap(Q, x), Q @ x, rtol=1e-13, atol=1e-13)
    las = instances.lasso(30, 200, mean_nnz=20, seed=0)
    x = rng.standard_normal(200)
    A = las.Q.A.toarray()
    np.testing.assert_allclose(o.qmap(las.Q, x), A.T @ (A @ x), rtol=1e-12, atol=1e-12)
    qap = build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False)
    x = rng.standard_normal(qap.n)
    q = qap.Q
    X = x.reshape(q.n, q.n, order="F")
    want = (2.0 * (q.A @ X @ q.B - q.S @ X - X @ q.T)).ravel(order="F")
    np.testing.assert_allclose(o.qmap(q, x), want, rtol=1e-12, atol=1e-9 * np.abs(want).max())
    # operator form vec(AXB) = (B^T kron A) vec(X) on a small case
    n = 5
    Aq, Bq, Sq, Tq = (rng.standard_normal((n, n)) for _ in range(4))
    from hprqp import QAP_Q_operator

    op = QAP_Q_operator(Aq, Bq, Sq, Tq)
    K = 2.0 * (np.kron(Bq.T, Aq) - np.kron(np.eye(n), Sq) - np.kron(Tq.T, np.eye(n)))
    x = rng.standard_normal(n * n)
    np.testing.assert_allclose(o.qmap(op, x), K @ x, rtol=1e-12, atol=1e-12)


def test_randn_stream_statistics():
    o = _oracle()
    z = o.randn(1, 200_000)
    assert abs(z.mean()) < 0.01
    assert abs(z.std() - 1.0) < 0.01
    assert np.all(o.randn(1, 10) == z[:10])  # counter-based: prefix stable
    assert not np.all(o.randn(2, 10) == z[:10])


def test_power_iteration_matches_eigvalsh():
    from hprqp import QP_info_cpu

    o = _oracle()
    rng = np.random.default_rng(2)
    A = sp.random(40, 70, density=0.2, random_state=rng, format="csc")
    R = sp.random(30, 70, density=0.2, random_state=rng)
    Q = (R.T @ R).tocsc()
    model = QP_info_cpu(Q, np.zeros(70), A, np.zeros(40), np.zeros(40), np.zeros(70), np.zeros(70))
    lam_a, _ = o.power_iteration(model, 0)
    lam_q, _ = o.power_iteration(model, 1)
    ea = np.linalg.eigvalsh((A @ A.T).toarray()).max()
    eq = np.linalg.eigvalsh(Q.toarray()).max()
    assert abs(lam_a - ea) / ea < 1e-3
    assert abs(lam_q - eq) / eq < 1e-3


@pytest.mark.parametrize("abcd", [(1.0, 2.0, 0.5, 3.0), (1e-6, 5.0, 2.0, 0.1), (10.0, 1e-3, 0.0, 0.0),
                                  (0.3, 0.3, 1e3, 1e2)])
def test_golden_matches_scipy(abcd):
    o = _oracle()
    a, b, c, d = abcd
    f = lambda x: a * x + b / x + c * x * x / (1 + d * x)
    got = o.golden(a, b, c, d)
    ref = minimize_scalar(f, bounds=(1e-12, 1e6), method="bounded", options={"xatol": 1e-12})
    assert f(got) <= ref.fun * (1 + 1e-9) + 1e-12
    if c == 0.0:
        assert abs(got - np.sqrt(b / a)) / np.sqrt(b / a) < 1e-6


def test_golden_Q_diag_matches_scipy():
    o = _oracle()
    rng = np.random.default_rng(3)
    Qd = rng.uniform(0.1, 2.0, 30)
    c = rng.standard_normal(30)
    d = c * rng.uniform(0.5, 1.5, 30)
    a, b = 0.7, 1.3
    f = lambda x: a * x + b / x + x * x * np.dot(c, d / (1 + x * Qd))
    got = o.golden_Q_diag(a, b, Qd, c, d)
    ref = minimize_scalar(f, bounds=(1e-9, 1e6), method="bounded", options={"xatol": 1e-12})
    assert f(got) <= ref.fun * (1 + 1e-9) + 1e-12


# --------------------------------------------------------- full solves
def _kkt_original(model, x, y, z):
    """Relative KKT residuals of (x, y, z) in the original space, computed
    independently of the oracle."""
    Q = model.Q if sp.issparse(model.Q) else None
    Qx = Q @ x if Q is not None else model.Q.Qmap(x)
    A = model.A
    Ax = A @ x if model.m else np.zeros(0)
    rp = np.clip(Ax, model.AL, model.AU) - Ax if model.m else np.zeros(0)
    aty = A.T @ y if model.m else np.zeros_like(x)
    rd = Qx + model.c - aty - z
    box = np.maximum(model.l - x, 0) + np.maximum(x - model.u, 0)
    bnorm = np.max(np.abs(np.concatenate([np.where(np.isinf(model.AL), 0, model.AL),
                                          np.where(np.isinf(model.AU), 0, model.AU)]))) if model.m else 0.0
    ep = max(np.max(np.abs(rp), initial=0.0), np.max(box, initial=0.0)) / (1 + max(bnorm, np.max(np.abs(Ax), initial=0)))
    ed = np.max(np.abs(rd)) / (1 + max(np.max(np.abs(model.c)), np.max(np.abs(aty)), np.max(np.abs(Qx))))
    return ep, ed


def test_model_mps_closed_form():
    """min x1^2+x2^2-3x1-5x2, x1+2x2<=10, 3x1+x2<=12, x>=0: the unconstrained
    minimiser (1.5, 2.5) is feasible, so the optimum is -8.5 with y = 0."""
    from hprqp import build_from_mps

    o = _oracle()
    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    r = o.solve(model)
    assert r.status == "OPTIMAL" and r.residuals < 1e-6
    assert abs(r.primal_obj + 8.5) < 1e-5
    np.testing.assert_allclose(r.x, [1.5, 2.5], atol=1e-4)
    np.testing.assert_allclose(r.y, [0.0, 0.0], atol=1e-4)


def test_aug2d_known_objective():
    from hprqp import build_from_mps

    o = _oracle()
    model = build_from_mps(fixture_path("AUG2D.mps"), verbose=False)
    assert (model.n, model.m, model.obj_constant) == (20200, 10000, 9900.0)
    r = o.solve(model)
    assert r.status == "OPTIMAL" and r.residuals < 1e-6
    assert r.noC and r.Q_is_diag
    assert abs(r.primal_obj - 1687411.752897) / 1687411.752897 < 1e-6
    ep, ed = _kkt_original(model, r.x, r.y, r.z)
    assert ep < 1e-5 and ed < 1e-5


def test_aug3dcqp_known_objective():
    """Maros-Meszaros AUG3DCQP optimum 9.9336215e+02 (public value)."""
    from hprqp import build_from_mps

    o = _oracle()
    model = build_from_mps(fixture_path("AUG3DCQP.mps"), verbose=False)
    r = o.solve(model)
    assert r.status == "OPTIMAL" and r.residuals < 1e-6
    assert abs(r.primal_obj - 993.36215) / 993.36215 < 2e-6


def _brute_force_qp(Q, c, A, AL, AU, l, u):
    """Enumerate active sets of a tiny convex QP; return the best KKT point."""
    n, m = len(c), A.shape[0]
    best = None
    for vs in itertools.product((0, 1, 2), repeat=n):          # 0 free, 1 at l, 2 at u
        for cs in itertools.product((0, 1, 2), repeat=m):      # 0 inactive, 1 at AL, 2 at AU
            if any(v == 1 and not np.isfinite(l[i]) or v == 2 and not np.isfinite(u[i]) for i, v in enumerate(vs)):
                continue
            if any(s == 1 and not np.isfinite(AL[i]) or s == 2 and not np.isfinite(AU[i]) for i, s in enumerate(cs)):
                continue
            rows, rhs = [], []
            for i, v in enumerate(vs):
                if v:
                    e = np.zeros(n)
                    e[i] = 1
                    rows.append(e)
                    rhs.append(l[i] if v == 1 else u[i])
            for i, s in enumerate(cs):
                if s:
                    rows.append(A[i])
                    rhs.append(AL[i] if s == 1 else AU[i])
            k = len(rows)
            E = np.array(rows).reshape(k, n)
            K = np.block([[Q, E.T], [E, np.zeros((k, k))]])
            try:
                sol = np.linalg.lstsq(K, np.concatenate([-c, rhs]), rcond=None)[0]
            except np.linalg.LinAlgError:
                continue
            x = sol[:n]
            if np.linalg.norm(K @ sol - np.concatenate([-c, rhs])) > 1e-8:
                continue
            if np.any(x < l - 1e-9) or np.any(x > u + 1e-9):
                continue
            Ax = A @ x
            if np.any(Ax < AL - 1e-9) or np.any(Ax > AU + 1e-9):
                continue
            f = 0.5 * x @ Q @ x + c @ x
            if best is None or f < best[0] - 1e-12:
                best = (f, x)
    return best


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_tiny_qp_matches_brute_force(seed):
    from hprqp import build_from_QAbc

    o = _oracle()
    rng = np.random.default_rng(seed)
    n, m = 4, 2
    R = rng.standard_normal((n, n))
    Q = R @ R.T + 0.1 * np.eye(n)
    c = rng.standard_normal(n)
    A = rng.standard_normal((m, n))
    x0 = rng.uniform(-0.5, 0.5, n)
    AL = A @ x0 - rng.uniform(0, 1, m)
    AU = A @ x0 + rng.uniform(0, 1, m)
    AL[0] = AU[0] = (A @ x0)[0]
    l, u = np.full(n, -1.0), np.full(n, 1.0)
    fbest, xbest = _brute_force_qp(Q, c, A, AL, AU, l, u)
    model = build_from_QAbc(Q, c, A, AL, AU, l, u, verbose=False)
    r = o.solve(model, stoptol=1e-9)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_obj - fbest) / (1 + abs(fbest)) < 1e-7
    np.testing.assert_allclose(r.x, xbest, atol=1e-5)


def test_random_qp_kkt_in_original_space():
    from hprqp import instances

    o = _oracle()
    model = instances.random_sparse_qp(600, 300, seed=5)
    r = o.solve(model)
    assert r.status == "OPTIMAL" and r.residuals < 1e-6
    ep, ed = _kkt_original(model, r.x, r.y, r.z)
    assert ep < 1e-6 and ed < 1e-6
    # complementarity of z with the box: z>0 only at the lower bound, z<0 at the upper
    act_l = np.isclose(r.x, model.l, atol=1e-6)
    act_u = np.isclose(r.x, model.u, atol=1e-6)
    free = ~(act_l | act_u)
    assert np.max(np.abs(r.z[free]), initial=0) < 1e-4 * (1 + np.abs(model.c).max())


def test_lasso_optimality_conditions():
    """x solves min 1/2||Ax-b||^2 + lam||x||_1 iff |A^T(Ax-b)|_i <= lam with
    equality (and opposite sign) where x_i != 0."""
    from hprqp import instances

    o = _oracle()
    model = instances.lasso(60, 400, mean_nnz=30, seed=4)
    r = o.solve(model, problem_type="LASSO", stoptol=1e-8)
    assert r.status == "OPTIMAL"
    A = model.Q.A
    g = A.T @ (A @ r.x - model.Q.b)
    lam = model.lam
    assert np.max(np.abs(g)) <= lam * (1 + 1e-5)
    nz = np.abs(r.x) > 1e-8
    np.testing.assert_allclose(g[nz], -lam * np.sign(r.x[nz]), rtol=1e-4, atol=1e-6 * lam)
    obj = 0.5 * np.sum((A @ r.x - model.Q.b) ** 2) + lam * np.abs(r.x).sum()
    assert abs(obj - r.primal_obj) / (1 + abs(obj)) < 1e-7


def test_qap_esc64a_kkt():
    from hprqp import build_from_mat

    o = _oracle()
    model = build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False)
    r = o.solve(model, problem_type="QAP")
    assert r.status == "OPTIMAL" and r.residuals < 1e-6
    X = r.x.reshape(64, 64, order="F")
    np.testing.assert_allclose(X.sum(axis=0), 1.0, atol=1e-5)
    np.testing.assert_allclose(X.sum(axis=1), 1.0, atol=1e-5)
    assert X.min() > -1e-6
    ep, ed = _kkt_original(model, r.x, r.y, r.z)
    assert ep < 1e-6 and ed < 1e-6


def test_warm_start_property():
    """test/runtests.jl:206,245: x-only warm start -> iter <= cold; x+y at the
    optimum -> iter == 0."""
    from hprqp import build_from_mps

    o = _oracle()
    model = build_from_mps(fixture_path("model.mps"), verbose=False)
    cold = o.solve(model)
    warm = o.solve(model, initial_x=cold.x, initial_y=cold.y)
    assert warm.iter == 0 and warm.status == "OPTIMAL"
    xonly = o.solve(model, initial_x=cold.x)
    assert xonly.iter <= cold.iter


def test_lp_matches_scipy_linprog():
    from scipy.optimize import linprog

    from hprqp import build_from_QAbc

    o = _oracle()
    rng = np.random.default_rng(7)
    n, m = 30, 12
    A = rng.standard_normal((m, n))
    x0 = rng.uniform(0, 1, n)
    b = A @ x0
    c = rng.uniform(0.1, 1.0, n)
    model = build_from_QAbc(sp.csr_matrix((n, n)), c, A, b, b, np.zeros(n), np.full(n, 3.0), verbose=False)
    r = o.solve(model, stoptol=1e-8)
    ref = linprog(c, A_eq=A, b_eq=b, bounds=[(0, 3)] * n, method="highs")
    assert r.status == "OPTIMAL"
    assert abs(r.primal_obj - ref.fun) / (1 + abs(ref.fun)) < 1e-6


def test_box_qp_without_constraints_matches_lbfgsb():
    """m = 0: no linear constraints, only the box (the reference skips the y
    half-step when ws.m == 0, src/kernels.jl:2246).  The optimum of a strictly
    convex box QP is unique, so L-BFGS-B's answer pins the oracle's."""
    from scipy.optimize import minimize

    from hprqp import build_from_QAbc

    o = _oracle()
    rng = np.random.default_rng(5)
    n = 80
    R = sp.random(40, n, density=0.1, random_state=6, format="csr")
    Q = (R.T @ R + sp.diags(rng.uniform(0.01, 1.0, n))).tocsr()
    c = rng.normal(size=n)
    lo, hi = np.full(n, -1.0), np.full(n, 1.0)
    model = build_from_QAbc(Q, c, sp.csr_matrix((0, n)), np.zeros(0), np.zeros(0), lo, hi, verbose=False)
    assert model.m == 0
    r = o.solve(model, stoptol=1e-8)
    ref = minimize(lambda x: 0.5 * x @ (Q @ x) + c @ x, np.zeros(n), jac=lambda x: Q @ x + c,
                   bounds=list(zip(lo, hi)), method="L-BFGS-B",
                   options=dict(ftol=1e-15, gtol=1e-12, maxiter=10000))
    assert r.status == "OPTIMAL"
    assert abs(r.primal_obj - ref.fun) / (1 + abs(ref.fun)) < 1e-8
    np.testing.assert_allclose(r.x, ref.x, atol=1e-5)
    # some bounds are active, so the box projection is exercised
    assert np.sum(np.abs(np.abs(r.x) - 1.0) < 1e-6) > 0


def test_every_bound_kind_against_active_set_certificate():
    """Free, lower-only, upper-only, boxed and fixed variables with =, <=, >=
    and ranged rows: the oracle's x agrees with the exact minimiser on its
    active set, and that minimiser is certified optimal."""
    from bounds_case import active_set_certificate, every_bound_kind_qp

    from hprqp import build_from_QAbc

    Q, c, A, AL, AU, l, u = every_bound_kind_qp()
    model = build_from_QAbc(Q, c, A, AL, AU, l, u, verbose=False)
    r = _oracle().solve(model, stoptol=1e-9)
    assert r.status == "OPTIMAL"
    xs, ok = active_set_certificate(Q, c, A, AL, AU, l, u, r.x)
    assert ok
    np.testing.assert_allclose(r.x, xs, atol=1e-6)
    fstar = 0.5 * xs @ (Q @ xs) + c @ xs
    assert abs(r.primal_obj - fstar) / (1 + abs(fstar)) < 1e-9


@pytest.mark.parametrize("lam_factor", [1.01, 0.995])
def test_lasso_closed_form_at_the_zero_threshold(lam_factor):
    from lasso_closed_form import closed_form

    from hprqp import instances

    model = instances.lasso(60, 900, mean_nnz=40, seed=7, lam_factor=lam_factor)
    x, f = closed_form(model.Q.A, model.Q.b, model.lam)
    r = _oracle().solve(model, problem_type="LASSO", stoptol=1e-9)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_obj - f) / (1 + abs(f)) < 1e-9
    np.testing.assert_allclose(r.x, x, atol=1e-6)


def test_partially_quadratic_against_active_set_certificate():
    """Q with 200 empty rows (linear variables): the oracle's x agrees with
    the certified exact minimiser on its active set."""
    from bounds_case import active_set_certificate, partially_quadratic_qp

    from hprqp import build_from_QAbc

    Q, c, A, AL, AU, l, u = partially_quadratic_qp()
    assert np.sum(np.diff(Q.indptr) == 0) == 200
    r = _oracle().solve(build_from_QAbc(Q, c, A, AL, AU, l, u, verbose=False), stoptol=1e-9)
    assert r.status == "OPTIMAL"
    xs, ok = active_set_certificate(Q, c, A, AL, AU, l, u, r.x)
    assert ok
    np.testing.assert_allclose(r.x, xs, atol=1e-6)


def test_bound_edges_mps_hand_derived_optimum():
    """tests/golden/bound_edges.mps by hand: x3 is fixed at 2.5, x8 = 0 and
    x4 = x2 - 2.5 (r3 at its lower end), x6 = 4 - x1 - x2 (r1 at its lower
    end), x1 at its upper bound -1; what is left is x2^2 - 3.5 x2 + 1.75,
    minimised at x2 = 1.75 with value -21/16 (x5 and x7 are not unique)."""
    from hprqp import build_from_mps

    model = build_from_mps(fixture_path("bound_edges.mps"), verbose=False)
    r = _oracle().solve(model, stoptol=1e-9)
    assert r.status == "OPTIMAL"
    assert abs(r.primal_obj - (-21 / 16)) < 1e-7
    np.testing.assert_allclose(r.x[[0, 1, 2, 3, 5, 7]], [-1, 1.75, 2.5, -0.75, 3.25, 0], atol=1e-6)
