"""Small QPs with structure the solver must handle (every kind of bound;
a Q with empty rows), and an exact certificate of an optimum.

Variables: free, lower-only, upper-only, boxed and fixed (l = u); rows:
equality, <=, >= and ranged (qp_formulation keeps each kind,
src/utils.jl:78-146).  The certificate reads the active set off a candidate
x, solves the equality-constrained QP on it exactly (dense KKT system), and
checks feasibility and the signs of the multipliers, so it does not depend on
the solver under test."""
import numpy as np
import scipy.sparse as sp


def every_bound_kind_qp():
    rng = np.random.default_rng(21)
    n, m = 60, 30
    R = sp.random(30, n, density=0.15, random_state=22, format="csr")
    Q = (R.T @ R + sp.diags(rng.uniform(0.05, 1.0, n))).tocsr()
    c = rng.normal(size=n)
    x0 = rng.uniform(-0.5, 0.5, n)
    l, u = np.full(n, -np.inf), np.full(n, np.inf)       # 0..11 free
    l[12:24] = -1.0                                      # lower only
    u[24:36] = 1.0                                       # upper only
    l[36:48], u[36:48] = -1.0, 1.0                       # boxed
    l[48:60] = u[48:60] = x0[48:60]                      # fixed
    A = (sp.random(m, n, density=0.15, random_state=23, format="csr") + sp.eye(m, n, format="csr")).tocsr()
    b = A @ x0
    r = rng.uniform(0.1, 1.0, m)
    AL, AU = np.full(m, -np.inf), np.full(m, np.inf)
    AL[:8] = AU[:8] = b[:8]                              # equality
    AU[8:16] = b[8:16] + r[8:16]                         # <=
    AL[16:24] = b[16:24] - r[16:24]                      # >=
    AL[24:], AU[24:] = b[24:] - r[24:], b[24:] + r[24:]  # ranged
    return Q, c, A, AL, AU, l, u


def partially_quadratic_qp():
    """Half the variables enter only linearly: Q's last 200 rows are empty,
    as in many Maros-Meszaros problems; box [0, 2], ranged rows."""
    rng = np.random.default_rng(31)
    n, m = 400, 150
    R = sp.random(80, 200, density=0.05, random_state=32, format="csr")
    Qh = (R.T @ R + sp.diags(rng.uniform(0.05, 1.0, 200))).tocsr()
    Q = sp.block_diag([Qh, sp.csr_matrix((200, 200))], format="csr")
    Q.eliminate_zeros()
    c = rng.normal(size=n)
    A = (sp.random(m, n, density=0.03, random_state=33, format="csr") + sp.eye(m, n, format="csr")).tocsr()
    b = A @ rng.uniform(0, 1, n)
    return Q, c, A, b - 0.5, b + 0.5, np.zeros(n), np.full(n, 2.0)


def active_set_certificate(Q, c, A, AL, AU, l, u, x, tol=1e-6):
    """(x*, ok): x* is the exact minimiser on the active set read off x;
    ok says x* is feasible and its multipliers have the signs optimality
    needs (Q x* + c = A_act^T lam + z)."""
    Q = sp.csr_matrix(Q).toarray()
    A = sp.csr_matrix(A).toarray()
    at_l = np.isfinite(l) & (np.abs(x - l) < tol)
    at_u = np.isfinite(u) & (np.abs(x - u) < tol) & ~at_l
    fixed = at_l | at_u
    xv = np.where(at_l, l, np.where(at_u, u, 0.0))
    Ax = A @ x
    r_l = np.isfinite(AL) & (np.abs(Ax - AL) < tol)
    r_u = np.isfinite(AU) & (np.abs(Ax - AU) < tol) & ~r_l
    act = r_l | r_u
    t = np.where(r_l, AL, AU)[act]
    f = ~fixed
    Aa = A[act]
    na = int(act.sum())
    K = np.block([[Q[np.ix_(f, f)], Aa[:, f].T], [Aa[:, f], np.zeros((na, na))]])
    rhs = np.concatenate([-c[f] - Q[np.ix_(f, fixed)] @ xv[fixed], t - Aa[:, fixed] @ xv[fixed]])
    sol = np.linalg.solve(K, rhs)
    xs = xv.copy()
    xs[f] = sol[:f.sum()]
    lam = -sol[f.sum():]
    z = Q @ xs + c - Aa.T @ lam
    eq = (AL == AU)[act]
    ok_rows = np.all(eq | np.where(r_l[act], lam >= -1e-9, lam <= 1e-9))
    ok_vars = (np.all(np.abs(z[f]) < 1e-8) and np.all(z[at_l & (l != u)] >= -1e-9)
               and np.all(z[at_u & (l != u)] <= 1e-9))
    feas = (np.all(A @ xs >= AL - 1e-9) and np.all(A @ xs <= AU + 1e-9)
            and np.all(xs >= l - 1e-9) and np.all(xs <= u + 1e-9))
    return xs, bool(ok_rows and ok_vars and feas)
