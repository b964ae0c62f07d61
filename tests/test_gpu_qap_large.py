"""QAP operator at sizes where the blocked MFMA passes run, against the FP64
closed form vec(2(AXB - SX - XT)), plus a QAP-relaxation solve at n = 400
whose returned (x, y, z) is checked in the original space with the CPU
operator.

Sizes cover every large-n path: 256 <= n < 1024 (register-blocked passes,
four waves splitting K), n >= 1024 with diagonal S and T (LDS-tiled
split-BF16 passes: 64 x 64 blocks at n = 1100, 64 x 128 blocks with swizzled
LDS rows at n = 2048 and 2100), and a full
S, T at n = 1100 (which stays on the register-blocked passes).  Sizes are
ragged against the block and MFMA tiles (300, 400, 517, 1100, 2100) as well as
aligned (256, 448, 512, 2048); coefficients are Gaussian (two-piece BF16 split of both
operands) or integer flows and distances (BF16-exact, three-piece data
split)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense(Q, X):
    return 2.0 * (Q.A @ X @ Q.B - Q.S @ X - X @ Q.T)


def _cases():
    from hprqp import instances
    from hprqp.model import QAP_Q_operator

    rng = np.random.default_rng(21)
    out = []
    for nq, full in ((300, False), (400, True), (517, False), (1100, False), (1100, True), (2100, False)):
        A = rng.standard_normal((nq, nq))
        B = rng.standard_normal((nq, nq))
        if full:
            S, T = rng.standard_normal((nq, nq)), rng.standard_normal((nq, nq))
        else:
            S, T = np.diag(rng.standard_normal(nq)), np.diag(rng.standard_normal(nq))
        out.append((f"gauss{nq}{'full' if full else 'diag'}", QAP_Q_operator(A, B, S, T), False))
    for nq in (256, 448, 512, 2048):
        out.append((f"flow{nq}", instances.qap(nq, seed=nq)[0].Q, True))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,Q,exact", CASES, ids=[c[0] for c in CASES])
def test_fp64_blocked_matches_closed_form(name, Q, exact):
    from hprqp import to_gpu

    nq = Q.A.shape[0]
    rng = np.random.default_rng(nq)
    X = rng.standard_normal((nq, nq))
    dev = to_gpu(Q, nq * nq)
    got = dev.Qmap(X.ravel(order="F"))
    want = _dense(Q, X).ravel(order="F")
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"{name}: fp64 max rel err {err:.2e}")
    assert err < 1e-13
    dev.close()


@pytest.mark.parametrize("name,Q,exact", CASES, ids=[c[0] for c in CASES])
def test_bf16x3_blocked_accuracy(name, Q, exact):
    from hprqp import to_gpu

    nq = Q.A.shape[0]
    rng = np.random.default_rng(nq + 1)
    X = rng.uniform(0.0, 1.0, (nq, nq)) / nq
    dev = to_gpu(Q, nq * nq)
    got = dev.Qmap(X.ravel(order="F"), precision="mixed")
    want = _dense(Q, X).ravel(order="F")
    aX = np.abs(X)
    M = (2.0 * (np.abs(Q.A) @ aX @ np.abs(Q.B) + np.abs(Q.S) @ aX + aX @ np.abs(Q.T))).ravel(order="F")
    err = np.max(np.abs(got - want) / M)
    print(f"{name}: split-bf16x3 max err/M {err:.2e} ({'exact' if exact else 'general'} coefficients)")
    assert err < (1e-5 if exact else 2e-4)
    dev.close()


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


def test_qap400_solve_fp64_and_mixed():
    """n = 400 (160,000 variables): the FP64 solve (blocked FP64 passes with
    the fused half-step epilogues) reaches 1e-6 and its (x, y, z) satisfies
    the KKT conditions in the original space by the CPU operator; the mixed
    solve (split-BF16 x3 passes) agrees on the objective within 1e-6."""
    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.qap(400, seed=5)[0]
    res = optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP"))
    ep, ed = _kkt_original(model, res.x, res.y, res.z)
    print(f"fp64: {res.status} iter={res.iter} kkt={res.residuals:.2e} obj={res.primal_obj:.10e} "
          f"original-space ep={ep:.2e} ed={ed:.2e}")
    assert res.status == "OPTIMAL" and res.residuals < 1e-6
    assert ep < 1e-5 and ed < 1e-5
    mix = optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP", precision="mixed"))
    rel = abs(mix.primal_obj - res.primal_obj) / (1.0 + abs(res.primal_obj))
    print(f"mixed: {mix.status} iter={mix.iter} polish={mix.polish_iter} kkt={mix.residuals:.2e} rel obj {rel:.2e}")
    assert mix.status == "OPTIMAL" and mix.residuals < 1e-6
    assert rel < 1e-6


def test_qap1024_solve_mixed_lds_passes():
    """n = 1024 (1,048,576 variables): the mixed solve runs the LDS-tiled
    split-BF16 x3 passes with the solver's fused half-step epilogues; it must
    reach 1e-6, satisfy the KKT conditions in the original space by the CPU
    operator, and agree with the FP64 solve's objective within 1e-6."""
    from hprqp import HPRQP_parameters, instances, optimize

    model = instances.qap(1024, seed=9)[0]
    ref = optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP"))
    mix = optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP", precision="mixed"))
    ep, ed = _kkt_original(model, mix.x, mix.y, mix.z)
    ep0, ed0 = _kkt_original(model, ref.x, ref.y, ref.z)
    rel = abs(mix.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj))
    print(f"fp64: {ref.status} iter={ref.iter} kkt={ref.residuals:.2e} time={ref.time:.2f}s original-space "
          f"ep={ep0:.2e} ed={ed0:.2e} | mixed: {mix.status} iter={mix.iter} polish={mix.polish_iter} "
          f"kkt={mix.residuals:.2e} time={mix.time:.2f}s rel obj {rel:.2e} original-space ep={ep:.2e} ed={ed:.2e}")
    assert ref.status == "OPTIMAL" and ref.residuals < 1e-6
    assert mix.status == "OPTIMAL" and mix.residuals < 1e-6
    assert ep0 < 1e-5 and ed0 < 1e-5
    assert ep < 1e-5 and ed < 1e-5
    assert rel < 1e-6
