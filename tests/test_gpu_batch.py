"""Batched solves: every instance must meet the same bar as a single solve
against the CPU oracle (1e-6 relative KKT, objective within 1e-6 relative,
iteration count within 10%)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(count, seed, nmax=400):
    from hprqp import instances

    return instances.portfolio_batch(count, seed=seed, n_range=(20, nmax), k_range=(2, 12))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle(precision):
    from hprqp import HPRQP_parameters
    from hprqp.batch import BatchModel
    from oracle import oracle

    insts = _batch(24, seed=5)
    bm = BatchModel([m for m, _ in insts])
    res = bm.solve(HPRQP_parameters(verbose=False, precision=precision))
    print(f"{precision}: solved {res.solved}/{bm.count} in {res.time:.3f}s, intervals={res.intervals}")
    assert res.solved == bm.count
    for i, (model, _) in enumerate(insts):
        ref = oracle.solve(model)
        rel = abs(res.primal_obj[i] - ref.primal_obj) / (1 + abs(ref.primal_obj))
        step = 10 if ref.iter < 1000 else 100
        print(f"  inst {i}: n={model.n} iter={res.iter[i]} (oracle {ref.iter}) kkt={res.residuals[i]:.2e} rel={rel:.1e}")
        assert res.status[i] == "OPTIMAL" and res.residuals[i] < 1e-6
        assert rel < 1e-6
        assert abs(res.iter[i] - ref.iter) <= max(0.1 * ref.iter, step)
        # feasibility of the returned x
        x = res.x[i]
        assert abs(x.sum() - 1.0) < 1e-5 and x.min() > -1e-6 and x.max() < 0.1 + 1e-6


def test_batch_matches_single_instance_solves():
    from hprqp import HPRQP_parameters, optimize
    from hprqp.batch import BatchModel

    insts = _batch(8, seed=9)
    bm = BatchModel([m for m, _ in insts])
    res = bm.solve(HPRQP_parameters(verbose=False))
    for i, (model, _) in enumerate(insts):
        single = optimize(model, HPRQP_parameters(verbose=False))
        assert single.status == "OPTIMAL"
        assert abs(single.iter - res.iter[i]) <= max(0.1 * single.iter, 10)
        assert abs(single.primal_obj - res.primal_obj[i]) / (1 + abs(single.primal_obj)) < 1e-6


def test_batch_warm_start_at_solution_retires_at_once():
    from hprqp import HPRQP_parameters
    from hprqp.batch import BatchModel

    insts = _batch(6, seed=13)
    bm = BatchModel([m for m, _ in insts])
    cold = bm.solve(HPRQP_parameters(verbose=False))
    warm = bm.solve(HPRQP_parameters(verbose=False), x0=cold.x, y0=cold.y)
    assert warm.solved == bm.count
    assert np.all(warm.iter == 0)


@pytest.mark.parametrize("sizes", [[(11, 1), (5000, 50), (12, 64), (300, 7)], [(700, 20)]],
                         ids=["extremes", "single"])
def test_batch_extreme_sizes_match_oracle(sizes):
    """Instances at the edges of what a batch accepts, side by side: n = 11
    with the 0.1 cap (eleven variables must share a budget of one), n = 5000
    with k = 50 (the top of the portfolio workload's range), rank 64 (the
    largest the batch accepts) on n = 12, and a batch of one instance."""
    from hprqp import HPRQP_parameters, instances
    from hprqp.batch import BatchModel
    from oracle import oracle

    models = [instances.portfolio(n, k, seed=40 + i, explicit=False)[0] for i, (n, k) in enumerate(sizes)]
    bm = BatchModel(models)
    res = bm.solve(HPRQP_parameters(verbose=False))
    assert res.solved == bm.count
    for i, model in enumerate(models):
        ref = oracle.solve(model)
        rel = abs(res.primal_obj[i] - ref.primal_obj) / (1 + abs(ref.primal_obj))
        print(f"  inst {i}: n={model.n} k={model.Q.U.shape[1]} iter={res.iter[i]} (oracle {ref.iter}) "
              f"kkt={res.residuals[i]:.2e} rel={rel:.1e}")
        assert res.status[i] == "OPTIMAL" and res.residuals[i] < 1e-6
        assert rel < 1e-6
        step = 10 if ref.iter < 1000 else 100
        assert abs(res.iter[i] - ref.iter) <= max(0.1 * ref.iter, step)
        x = res.x[i]
        assert abs(x.sum() - 1.0) < 1e-5 and x.min() > -1e-6 and x.max() < 0.1 + 1e-6
