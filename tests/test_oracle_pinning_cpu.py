"""Pins the oracle's scaling! (oracle/hprqp_oracle.c `scaling`, after
src/utils.jl:350-804) and compute_M_norm! (`compute_M_norm`, after
src/algorithm.jl:689-739) by what the mathematics fixes, not by restating
their formulas.

Scaling, on K = [Q A^T; A 0] (symmetric, since the formulated Q is):
- Ruiz passes divide K_kl by s_k s_l with s_k = sqrt(max_l |K_kl|); as
  s_k^2 >= |K_kl| and s_l^2 >= |K_kl|, every entry of the result is <= 1 in
  magnitude.  Each row maximum also rises toward 1: the first pass leaves
  r_k >= sqrt(r0_k / max r0), every later one r_k >= sqrt(r_k) (all entries
  are <= 1 by then), so after the reference's 10 passes
  r_k >= (r0_k / max r0)^(2^-10) for every nonempty row.
- The Pock-Chambolle pass divides by s_k = sqrt(sum_l |K_kl|); the Schur test
  with weights s_k gives ||K~||_2 <= 1.
- The returned row_norm / col_norm are the whole scaling: Q~ = Q / (cn cn^T),
  A~ = A / (rn cn^T), c~ = c / cn, l~ = l cn, AL~ = AL / rn.

M-norm: the reference's unclamped terms add up to a sum of squares,
    M_1 + M_2 + M_3 = (1/s)|dx + s(A^T dy - Q dw)|^2 + s(lA |dy|^2 - |A^T dy|^2)
                      + s(lQ <dw, Q dw> - |Q dw|^2) + s^2/(1 + s lQ) <A^T dy, Q A^T dy>
(diagonal Q: the last two terms become sum_i s^2 d_i (A^T dy)_i^2 / (1 + s d_i)),
each term >= 0 once lA >= lambda_max(A A^T) and lQ >= lambda_max(Q); so
M_1 >= 0, M_3 >= 0, M_1 + M_2 >= 0, the clamps never act and the M-norm is
the square root of that form.  It then obeys the parallelogram law, and for
Q = 0 it is the textbook PDHG metric [dx; dy]^T [[I/s, A^T], [A, s lA I]] [dx; dy].

Each check is also run on copies of the oracle with a plausible slip planted
(compiled here from a textual edit of oracle/hprqp_oracle.c), and must fail
there: that is what makes it a pin."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SRC = ROOT / "oracle" / "hprqp_oracle.c"


# ---------------------------------------------------------------- instances
def _wide(rng, k):
    """Values of both signs spread over six decades."""
    return np.exp(rng.uniform(np.log(1e-3), np.log(1e3), k)) * rng.choice([-1.0, 1.0], k)


def _scaling_model(seed=0, n=60, m=30):
    from hprqp.model import QP_info_cpu

    rng = np.random.default_rng(seed)
    R = sp.random(25, n, density=0.15, random_state=seed + 1, format="csr")
    R.data = _wide(rng, R.nnz)
    d = np.abs(_wide(rng, n))
    d[7] = 0.0
    Q = (R.T @ R + sp.diags(d)).tolil()
    Q[7, :] = 0.0  # one column (and row) of Q empty ...
    Q[:, 7] = 0.0
    Q = sp.csr_matrix(Q)
    Q.eliminate_zeros()
    A = sp.random(m, n, density=0.12, random_state=seed + 2, format="lil")
    A[:, 7] = 0.0  # ... and of A: its statistic is 0, left unscaled
    A[3, :] = 0.0  # an empty constraint row
    A = sp.csr_matrix(A)
    A.data = _wide(rng, A.nnz)
    AL = rng.normal(size=m) - 1.0
    AU = AL + rng.uniform(0.0, 2.0, m)
    AU[:5] = np.inf
    l = rng.uniform(-2.0, 0.0, n)
    u = l + rng.uniform(0.0, 3.0, n)
    l[:4] = -np.inf
    return QP_info_cpu(Q=Q, c=rng.normal(size=n) * 10, A=sp.csc_matrix(A), AL=AL, AU=AU, l=l, u=u)


def _K(Q, A):
    m = A.shape[0]
    return np.block([[np.asarray(Q), np.asarray(A).T], [np.asarray(A), np.zeros((m, m))]])


# ---------------------------------------------------------------- scaling checks
def _ruiz_ok(model, d):
    K0 = np.abs(_K(model.Q.toarray(), model.A.toarray()))
    K = np.abs(_K(d["Q"], d["A"]))
    if K.max() > 1.0 + 1e-14:
        return False
    r0, r = K0.max(axis=1), K.max(axis=1)
    live = r0 > 0
    bound = (r0[live] / r0.max()) ** (2.0 ** -10)
    return bool(np.all(r[live] >= bound * (1.0 - 1e-12)))


def _pc_ok(d):
    K = _K(d["Q"], d["A"])
    return float(np.max(np.abs(np.linalg.eigvalsh(K)))) <= 1.0 + 1e-12


def _bookkeeping_ok(model, d):
    cn, rn = d["col_norm"], d["row_norm"]
    Q0, A0 = model.Q.toarray(), model.A.toarray()
    ok = np.allclose(d["Q"], Q0 / np.outer(cn, cn), rtol=1e-13, atol=0)
    ok &= np.allclose(d["A"], A0 / np.outer(rn, cn), rtol=1e-13, atol=0)
    ok &= np.allclose(d["c"], model.c / cn, rtol=1e-13, atol=0)
    ok &= np.array_equal(np.isinf(d["l"]), np.isinf(model.l))
    fin = np.isfinite(model.l)
    ok &= np.allclose(d["l"][fin], model.l[fin] * cn[fin], rtol=1e-13, atol=0)
    ok &= np.allclose(d["u"], model.u * cn, rtol=1e-13, atol=0)
    ok &= np.allclose(d["AL"], model.AL / rn, rtol=1e-13, atol=0)
    fin = np.isfinite(model.AU)
    ok &= np.allclose(d["AU"][fin], model.AU[fin] / rn[fin], rtol=1e-13, atol=0)
    bt = np.maximum(np.abs(np.where(np.isinf(d["AL"]), 0.0, d["AL"])), np.abs(np.where(np.isinf(d["AU"]), 0.0, d["AU"])))
    ok &= np.isclose(d["norm_b"], np.linalg.norm(bt), rtol=1e-13)
    ok &= np.isclose(d["norm_c"], np.linalg.norm(d["c"]), rtol=1e-13)
    b0 = np.maximum(np.abs(np.where(np.isinf(model.AL), 0.0, model.AL)),
                    np.abs(np.where(np.isinf(model.AU), 0.0, model.AU)))
    ok &= d["norm_b_org"] == np.max(b0) and d["norm_c_org"] == np.max(np.abs(model.c))
    ok &= d["b_scale"] == 1.0 and d["c_scale"] == 1.0
    return bool(ok)


def _scaling_checks(lib=None, seeds=(0, 1, 2)):
    """(Ruiz bound and equilibration with PC off, PC spectral bound with Ruiz
    off and on, bookkeeping in all three settings), over a few instances."""
    ruiz = pc = book = True
    for seed in seeds:
        model = _scaling_model(seed)
        d_r = oracle.scale(model, True, False, lib=lib)
        d_p = oracle.scale(model, False, True, lib=lib)
        d_rp = oracle.scale(model, True, True, lib=lib)
        ruiz &= _ruiz_ok(model, d_r)
        pc &= _pc_ok(d_p) and _pc_ok(d_rp)
        book &= _bookkeeping_ok(model, d_r) and _bookkeeping_ok(model, d_p) and _bookkeeping_ok(model, d_rp)
    return {"ruiz": ruiz, "pc": pc, "bookkeeping": book}


def test_scaling_ruiz_pock_chambolle_and_bookkeeping():
    got = _scaling_checks()
    print(got)
    assert got == {"ruiz": True, "pc": True, "bookkeeping": True}


def test_scaling_of_an_operator_Q_is_the_identity():
    """Operator Q (LASSO, QAP, low rank) skips all scaling
    (src/utils.jl:364-406): norms 1, data unchanged, b_scale = c_scale = 1."""
    from hprqp import instances

    model = instances.lasso(30, 200, mean_nnz=20, seed=0)
    d = oracle.scale(model)
    assert np.all(d["col_norm"] == 1.0) and d["row_norm"].size == 0
    np.testing.assert_array_equal(d["c"], model.c)
    assert np.isclose(d["norm_c"], np.linalg.norm(model.c), rtol=1e-14) and d["b_scale"] == d["c_scale"] == 1.0


# ---------------------------------------------------------------- M-norm checks
def _mnorm_case(seed, diag=False, lp=False, m=25, n=50):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=0.2, random_state=seed + 10, format="csr")
    A.data = rng.normal(size=A.nnz)
    if lp:
        Q = sp.csr_matrix((n, n))
    elif diag:
        Q = sp.diags(rng.uniform(0.0, 3.0, n)).tocsr()
    else:
        R = sp.random(20, n, density=0.2, random_state=seed + 11, format="csr")
        R.data = rng.normal(size=R.nnz)
        Q = (R.T @ R + sp.diags(rng.uniform(0.01, 1.0, n))).tocsr()
    Qd, Ad = Q.toarray(), A.toarray()
    lamA = 1.05 * float(np.max(np.linalg.eigvalsh(Ad @ Ad.T))) if m > 0 else 0.0
    # diagonal Q: lambda_Q = max(diag) without eig_factor (src/unified_operations.jl:712-713)
    lamQ = float(np.max(np.diag(Qd))) if (diag or lp) else 1.05 * float(np.max(np.linalg.eigvalsh(Qd)))
    return Q, A, Qd, Ad, lamA, lamQ, rng


def _sos(Qd, Ad, dx, dy, dw, s, lamA, lamQ, diag):
    u = Ad.T @ dy
    qdw = Qd @ dw
    v = u - qdw
    total = (dx + s * v) @ (dx + s * v) / s + s * (lamA * (dy @ dy) - u @ u)
    if diag:
        dq = np.diag(Qd)
        total += np.sum(s * s * dq * u * u / (1.0 + s * dq))
    else:
        total += s * (lamQ * (dw @ qdw) - qdw @ qdw) + s * s / (1.0 + s * lamQ) * (u @ (Qd @ u))
    return total


def _mnorm_checks(lib=None):
    sos = signs = par = pdhg = True
    for seed in range(4):
        for diag in (False, True):
            Q, A, Qd, Ad, lamA, lamQ, rng = _mnorm_case(seed, diag=diag)
            n, m = Qd.shape[0], Ad.shape[0]
            for s in (0.03, 1.0, 17.0):
                vecs = [rng.normal(size=k) for k in (n, m, n, n, m, n)]
                a, b = vecs[:3], vecs[3:]
                val, t = oracle.mnorm(Q, A, *a, s, lamA, lamQ, diag, lib=lib)
                want = _sos(Qd, Ad, *a, s, lamA, lamQ, diag)
                sos &= want >= 0 and abs(val * val - want) <= 1e-11 * want
                signs &= t[0] >= 0 and t[2] >= 0 and t[0] + t[1] >= 0
                mn = lambda v: oracle.mnorm(Q, A, *v, s, lamA, lamQ, diag, lib=lib)[0] ** 2
                lhs = mn([x + y for x, y in zip(a, b)]) + mn([x - y for x, y in zip(a, b)])
                rhs = 2.0 * mn(a) + 2.0 * mn(b)
                par &= abs(lhs - rhs) <= 1e-11 * rhs
        # Q = 0: the PDHG metric
        Q, A, Qd, Ad, lamA, lamQ, rng = _mnorm_case(seed, lp=True)
        n, m = Qd.shape[0], Ad.shape[0]
        s = 0.7
        dx, dy, dw = rng.normal(size=n), rng.normal(size=m), rng.normal(size=n)
        Mp = np.block([[np.eye(n) / s, Ad.T], [Ad, s * lamA * np.eye(m)]])
        z = np.concatenate([dx, dy])
        for diag in (True, False):  # an empty Q passes the diagonal test; either branch gives the metric
            val, _ = oracle.mnorm(Q, A, dx, dy, dw, s, lamA, 0.0, diag, lib=lib)
            pdhg &= abs(val * val - z @ Mp @ z) <= 1e-11 * (z @ Mp @ z)
    return {"sos": bool(sos), "signs": bool(signs), "parallelogram": bool(par), "pdhg": bool(pdhg)}


def test_mnorm_is_the_sum_of_squares_metric():
    got = _mnorm_checks()
    print(got)
    assert got == {"sos": True, "signs": True, "parallelogram": True, "pdhg": True}


def test_mnorm_without_constraints():
    """m = 0: M^2 = (1/s)|dx - s Q dw|^2 + s(lQ <dw, Q dw> - |Q dw|^2)."""
    Q, A, Qd, Ad, lamA, lamQ, rng = _mnorm_case(5, m=0)
    n = Qd.shape[0]
    dx, dw = rng.normal(size=n), rng.normal(size=n)
    s = 2.5
    val, t = oracle.mnorm(Q, sp.csr_matrix((0, n)), dx, np.zeros(0), dw, s, 0.0, lamQ, False)
    qdw = Qd @ dw
    want = (dx - s * qdw) @ (dx - s * qdw) / s + s * (lamQ * (dw @ qdw) - qdw @ qdw)
    assert abs(val * val - want) <= 1e-12 * want and t[2] == 0.0


# ---------------------------------------------------------------- planted slips
SLIPS = {
    "ruiz_without_sqrt": ("double t = ruiz ? sqrt(fmax_j(sa, sq)) : sqrt(sa + sq);",
                          "double t = ruiz ? fmax_j(sa, sq) : sqrt(sa + sq);", "ruiz"),
    "pc_max_not_sum": ("sa = ruiz ? fmax_j(sa, fabs(A->v[k])) : sa + fabs(A->v[k]);",
                       "sa = fmax_j(sa, fabs(A->v[k]));", "pc"),
    "Q_left_out_of_column_statistic": ("sq = ruiz ? fmax_j(sq, fabs(q->Q.v[k])) : sq + fabs(q->Q.v[k]);",
                                       "(void)sq;", "ruiz"),
    "c_multiplied_not_divided": ("for (int32_t j = 0; j < n; ++j) c[j] /= tc[j];",
                                 "for (int32_t j = 0; j < n; ++j) c[j] *= tc[j];", "bookkeeping"),
    "row_norm_not_accumulated": ("for (int32_t i = 0; i < m; ++i) sc->row_norm[i] *= tr[i];",
                                 "for (int32_t i = 0; i < m; ++i) sc->row_norm[i] = tr[i];", "bookkeeping"),
    "mnorm_cross_term_sign": ("M2 += 2.0 * dot(n, ws->ATdy, ws->dx);", "M2 -= 2.0 * dot(n, ws->ATdy, ws->dx);",
                              "sos"),
    "mnorm_dropped_lambdaQ_term": ("M1 += ws->sigma * ws->lambda_max_Q * dot(n, ws->dw, ws->dQw);", "(void)0;",
                                   "sos"),
    "mnorm_sigma_power_in_sGS_term": ("M3 = (ws->sigma * ws->sigma) / (1.0 + ws->sigma * ws->lambda_max_Q) *",
                                      "M3 = (ws->sigma) / (1.0 + ws->sigma * ws->lambda_max_Q) *", "sos"),
    "mnorm_diag_dropped_dQw_term": ("M1 += ws->sigma * dot(n, ws->dQw, ws->dQw);", "(void)0;", "sos"),
}


def _planted_lib(tmp_path, name):
    old, new, _ = SLIPS[name]
    src = ORACLE_SRC.read_text()
    assert src.count(old) == 1, f"slip {name}: the edited text must occur exactly once"
    out = tmp_path / f"oracle_{name}.c"
    out.write_text(src.replace(old, new))
    so = tmp_path / f"liboracle_{name}.so"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-std=gnu11", f"-I{ORACLE_SRC.parent}", str(out), "-o", str(so),
                    "-lm"], check=True)
    return C.CDLL(str(so))


@pytest.mark.parametrize("name", sorted(SLIPS))
def test_each_check_fails_on_a_planted_slip(tmp_path, name):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lib = _planted_lib(tmp_path, name)
    check = SLIPS[name][2]
    got = _scaling_checks(lib=lib) if check in ("ruiz", "pc", "bookkeeping") else _mnorm_checks(lib=lib)
    print(name, got)
    assert got[check] is False, f"the {check} check does not see the slip {name}"
