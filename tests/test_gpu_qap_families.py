"""The QAP operator's kernel families, each pinned to the FP64 closed form
vec(2(AXB - SX - XT)) of src/Q_operators/QAP_operator.jl:7-30.

From n = 256 the two GEMM passes have more than one form: blocked 32 x 32
tiles read from global memory (family 2), LDS-staged 64 x 64 blocks (3;
diagonal S and T) and, for the split-BF16 passes, LDS-staged 64 x 128
blocks (4; a grid of at least 512 blocks).  The solver times the families
that can run at an operator's first solve and keeps the fastest
(params.qap_family = 0), or takes the caller's.  Each family is set on the
operator (hprqp_qop_set_qap_family), the family reported must be the one
asked for where it can run and the size rule's elsewhere, and every FP64
apply must match the closed form to 1e-13 and every split-BF16 apply must
meet the bound of the split (1e-5 of |A||X||B| + ... for BF16-exact
coefficients, 2e-4 for Gaussian ones).  Solves of a 300 x 300 relaxation
under each family, and autotuned, must reach 1e-6 and satisfy the KKT
conditions in the original space by the CPU operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense(Q, X):
    return 2.0 * (Q.A @ X @ Q.B - Q.S @ X - X @ Q.T)


def _cases():
    from hprqp import instances
    from hprqp.model import QAP_Q_operator

    rng = np.random.default_rng(31)
    out = []
    for nq, full in ((300, False), (400, True), (1100, False)):
        A = rng.standard_normal((nq, nq))
        B = rng.standard_normal((nq, nq))
        if full:
            S, T = rng.standard_normal((nq, nq)), rng.standard_normal((nq, nq))
        else:
            S, T = np.diag(rng.standard_normal(nq)), np.diag(rng.standard_normal(nq))
        out.append((f"gauss{nq}{'full' if full else 'diag'}", QAP_Q_operator(A, B, S, T), False, not full))
    for nq in (100, 512, 2048):
        out.append((f"flow{nq}", instances.qap(nq, seed=nq + 1)[0].Q, True, True))
    return out


CASES = _cases()


def _expected(n, diag, fam, bf16):
    """The family a pass runs (include/hprqp.h, params.qap_family)."""
    if n < 256:
        return 1
    np_ = (n + 127) // 128 * 128
    wide_ok = (np_ // 64) * (np_ // 128) >= 512
    if fam == 2:
        return 2
    if fam == 3 and diag:
        return 3
    if fam == 4 and diag and bf16 and wide_ok:
        return 4
    if n >= 1024 and diag:
        return 4 if bf16 and wide_ok else 3
    return 2


@pytest.mark.parametrize("name,Q,exact,diag", CASES, ids=[c[0] for c in CASES])
def test_every_family_matches_closed_form(name, Q, exact, diag):
    from hprqp import to_gpu

    nq = Q.A.shape[0]
    rng = np.random.default_rng(nq + 7)
    Xg = rng.standard_normal((nq, nq))
    Xu = rng.uniform(0.0, 1.0, (nq, nq)) / nq
    want_g = _dense(Q, Xg).ravel(order="F")
    want_u = _dense(Q, Xu).ravel(order="F")
    aX = np.abs(Xu)
    M = (2.0 * (np.abs(Q.A) @ aX @ np.abs(Q.B) + np.abs(Q.S) @ aX + aX @ np.abs(Q.T))).ravel(order="F")
    dev = to_gpu(Q, nq * nq)
    try:
        for fam in (0, 2, 3, 4):
            f64, f16 = dev.set_qap_family(fam)
            assert (f64, f16) == (_expected(nq, diag, fam, False), _expected(nq, diag, fam, True)), (fam, f64, f16)
            got = dev.Qmap(Xg.ravel(order="F"))
            err = np.max(np.abs(got - want_g)) / np.max(np.abs(want_g))
            got16 = dev.Qmap(Xu.ravel(order="F"), precision="mixed")
            err16 = np.max(np.abs(got16 - want_u) / M)
            print(f"{name}: family {fam} -> fp64 {f64} (max rel err {err:.2e}), split-bf16 {f16} "
                  f"(max err/M {err16:.2e})")
            assert err < 1e-13
            assert err16 < (1e-5 if exact else 2e-4)
    finally:
        dev.close()


def test_family_refuses_other_operators_and_codes():
    import scipy.sparse as sp

    from hprqp import HPRQP_parameters, instances, to_gpu

    dev = to_gpu(sp.eye(5, format="csr"), 5)
    with pytest.raises(RuntimeError, match="not a QAP operator"):
        dev.set_qap_family(2)
    dev.close()
    dev = to_gpu(instances.qap(40, seed=1)[0].Q, 1600)
    with pytest.raises(RuntimeError, match="qap_family"):
        dev.set_qap_family(5)
    dev.close()
    with pytest.raises(ValueError, match="qap_family"):
        HPRQP_parameters(qap_family=1).to_c()


def _kkt_original(model, x, y, z):
    Qx = model.Q.Qmap(x)
    Ax = model.A @ x
    rp = np.clip(Ax, model.AL, model.AU) - Ax
    aty = model.A.T @ y
    rd = Qx + model.c - aty - z
    box = np.maximum(model.l - x, 0) + np.maximum(x - model.u, 0)
    bnorm = np.max(np.abs(np.concatenate([model.AL, model.AU])))
    ep = max(np.max(np.abs(rp)), np.max(box)) / (1 + max(bnorm, np.max(np.abs(Ax))))
    ed = np.max(np.abs(rd)) / (1 + max(np.max(np.abs(model.c)), np.max(np.abs(aty)), np.max(np.abs(Qx))))
    return ep, ed


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_under_each_family(precision):
    from hprqp import HPRQP_parameters, instances
    from hprqp.solver import DeviceModel

    model = instances.qap(300, seed=11)[0]
    dm = DeviceModel(model)
    ref = None
    # 0 twice: the second autotuned solve takes the family the first timed
    for fam in (2, 3, 0, 0):
        res = dm.solve(HPRQP_parameters(verbose=False, problem_type="QAP", precision=precision, qap_family=fam))
        ep, ed = _kkt_original(model, res.x, res.y, res.z)
        print(f"{precision} family {fam}: ran {res.qap_family}, {res.status} iter={res.iter} "
              f"kkt={res.residuals:.2e} obj={res.primal_obj:.10e} time={res.time:.3f}s ep={ep:.2e} ed={ed:.2e}")
        assert res.status == "OPTIMAL" and res.residuals < 1e-6
        assert ep < 1e-5 and ed < 1e-5
        if fam:
            assert res.qap_family == fam
        else:
            assert res.qap_family in (2, 3)
        if ref is None:
            ref = res
        else:
            assert abs(res.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj)) < 1e-6
            assert abs(res.iter - ref.iter) <= 0.1 * ref.iter + 10
    dm.close()
