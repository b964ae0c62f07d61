"""The QAP relaxation's one-apply iteration against the reference's two-apply
iteration.

With the relaxation's doubly stochastic constraints (build_from_ABST,
src/utils.jl:1133-1180) the library carries Q w and Q ATy by their linear
recurrences and forms Q A^T y_bar from y_bar in O(nq^2), so each HPR step
applies the operator once (to w_bar) instead of twice
(src/kernels.jl:1078-1154 recomputes Q w every step).  In exact arithmetic
the two iterations are the same; here they differ by rounding only: the same
status, iteration count and restarts, objectives to 1e-9 and, stopped
part-way, iterates to 1e-7 (as the other mid-solve trajectory tests; the
dual of the 128 x 128 relaxation drifts by 1.5e-9 in 200 iterations) --
against the oracle too.  Cases: esc64a (full S
and T, the full-matrix products of the constants), the 128 x 128
flow/distance relaxation (diagonal shifts) and a tile-ragged nq = 45; warm
starts (the products formed from the given iterates) and the reduced
precisions (split-BF16 x3 applies)."""
import numpy as np
import pytest

from conftest import fixture_path
from test_gpu_parity import _check, _solve_both

pytestmark = pytest.mark.gpu


def _model(case):
    from hprqp import build_from_mat, instances

    if case == "esc64a":
        return build_from_mat(fixture_path("esc64a.mat"), "QAP", verbose=False)
    if case == "flow128":
        return instances.qap(128, seed=0)[0]
    return instances.qap(45, seed=2)[0]


def _solve(model, **kw):
    from hprqp import HPRQP_parameters, optimize

    return optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP", **kw))


@pytest.mark.parametrize("case", ["esc64a", "flow128", "ragged45"])
def test_one_apply_form_matches_two_apply_form_and_oracle(case):
    model = _model(case)
    res, ref = _solve_both(model, problem_type="QAP")  # one-apply form (the default)
    _check(res, ref)
    two = _solve(model, qap_recompute_Qw=True)
    rel = abs(res.primal_obj - two.primal_obj) / (1.0 + abs(two.primal_obj))
    dx = np.max(np.abs(res.x - two.x)) / (1.0 + np.max(np.abs(two.x)))
    print(f"{case}: one-apply iter={res.iter} restarts={res.restarts} time={res.time:.4f}s; two-apply "
          f"iter={two.iter} restarts={two.restarts} time={two.time:.4f}s; oracle iter={ref.iter}; "
          f"rel obj {rel:.2e}, x diff {dx:.2e}")
    assert two.status == "OPTIMAL"
    assert res.iter == two.iter and res.restarts == two.restarts
    assert rel < 1e-9 and dx < 1e-7


@pytest.mark.parametrize("case", ["esc64a", "flow128"])
def test_one_apply_trajectory_mid_solve(case):
    """Stopped at 200 iterations: the one-apply iterates equal the two-apply
    ones and the oracle's to rounding."""
    from oracle import oracle

    model = _model(case)
    its = 200
    one = _solve(model, max_iter=its)
    two = _solve(model, max_iter=its, qap_recompute_Qw=True)
    ref = oracle.solve(model, max_iter=its, problem_type="QAP")
    assert one.iter == two.iter == ref.iter == its and one.restarts == two.restarts == ref.restarts
    for other, name in ((two, "two-apply"), (ref, "oracle")):
        dx = np.max(np.abs(one.x - other.x)) / (1.0 + np.max(np.abs(other.x)))
        dy = np.max(np.abs(one.y - other.y)) / (1.0 + np.max(np.abs(other.y)))
        print(f"{case}: vs {name}: x diff {dx:.2e}, y diff {dy:.2e}")
        assert dx < 1e-7 and dy < 1e-7


def test_one_apply_warm_start():
    """initial_x / initial_y: the carried products are formed from the given
    iterates; the same iterations as the two-apply form and the oracle."""
    from hprqp import HPRQP_parameters, optimize
    from oracle import oracle

    model = _model("ragged45")
    sol = _solve(model)
    rng = np.random.default_rng(7)
    x0 = sol.x + 1e-3 * rng.standard_normal(model.n)
    y0 = sol.y + 1e-3 * rng.standard_normal(model.m)
    one = optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP", initial_x=x0, initial_y=y0))
    two = optimize(model, HPRQP_parameters(verbose=False, problem_type="QAP", initial_x=x0, initial_y=y0,
                                           qap_recompute_Qw=True))
    ref = oracle.solve(model, problem_type="QAP", initial_x=x0, initial_y=y0)
    print(f"warm: one-apply iter={one.iter}, two-apply iter={two.iter}, oracle iter={ref.iter}, cold {sol.iter}")
    assert one.status == two.status == "OPTIMAL"
    assert one.iter == two.iter
    assert abs(one.iter - ref.iter) <= 0.1 * ref.iter + 10
    assert abs(one.primal_obj - ref.primal_obj) / (1.0 + abs(ref.primal_obj)) < 1e-6


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_one_apply_reduced_precision(precision):
    """Split-BF16 x3 applies of w_bar with the carried products: 1e-6 KKT and
    the two-apply form's objective to 1e-6."""
    for case in ("esc64a", "flow128"):
        model = _model(case)
        one = _solve(model, precision=precision)
        two = _solve(model, precision=precision, qap_recompute_Qw=True)
        fp64 = _solve(model)
        rel = abs(one.primal_obj - fp64.primal_obj) / (1.0 + abs(fp64.primal_obj))
        print(f"{precision} {case}: one-apply iter={one.iter} polish={one.polish_iter} time={one.time:.4f}s, "
              f"two-apply iter={two.iter} time={two.time:.4f}s, rel obj vs fp64 {rel:.2e}")
        assert one.status == two.status == "OPTIMAL" and one.residuals < 1e-6
        assert rel < 1e-6
        assert one.iter <= 1.1 * two.iter + 10
