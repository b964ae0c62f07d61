"""Batched instances with a sparse Q (hprqp_batch_create_csr): each instance is
scaled as the single solve scales it (Ruiz and Pock-Chambolle, then
(Q + Q')/2) and iterated by its own workgroup with CSR products of Q
(Qmap!, src/Q_operators/sparse_matrix_operator.jl:27-51).  Every instance
must meet the single solve's bar against the CPU oracle: 1e-6 relative KKT,
objective within 1e-6 relative, iteration count within 10% (or one residual
step).  The instances are the random sparse QP generator's (Q = R'R + diag,
5 nonzeros per row of A, 30% equality rows) at uneven sizes n in
[200, 2000]."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

SIZES = [200, 1999, 350, 1200, 640, 2000, 275, 900, 1500, 420]


def _models(sizes, seed0=100):
    from hprqp import instances

    return [instances.random_sparse_qp(n, n // 2, seed=seed0 + i) for i, n in enumerate(sizes)]


def _iter_ok(got, want):
    return abs(got - want) <= max(0.1 * want, 10 if want < 1000 else 100)


def test_csr_batch_matches_oracle():
    from hprqp import HPRQP_parameters
    from hprqp.batch import BatchModel
    from oracle import oracle

    models = _models(SIZES)
    bm = BatchModel(models)
    assert bm.kind == "csr"
    res = bm.solve(HPRQP_parameters(verbose=False))
    print(f"solved {res.solved}/{bm.count} in {res.time:.3f} s, {res.intervals} intervals, "
          f"interval kernels {res.interval_time:.3f} s")
    assert res.solved == bm.count
    for i, model in enumerate(models):
        ref = oracle.solve(model)
        rel = abs(res.primal_obj[i] - ref.primal_obj) / (1 + abs(ref.primal_obj))
        print(f"  inst {i}: n={model.n} iter={res.iter[i]} (oracle {ref.iter}) kkt={res.residuals[i]:.2e} "
              f"rel obj {rel:.1e}")
        assert res.status[i] == "OPTIMAL" and res.residuals[i] < 1e-6
        assert rel < 1e-6
        assert _iter_ok(res.iter[i], ref.iter)
        # the returned (x, y) are in the original space: x within the box,
        # A x within [AL, AU] to the tolerance's scale
        x = res.x[i]
        assert np.all(x >= model.l - 1e-9) and np.all(x <= model.u + 1e-9)
        ax = model.A @ x
        viol = np.max(np.maximum(model.AL - ax, 0) + np.maximum(ax - model.AU, 0))
        assert viol <= 1e-5 * (1 + np.max(np.abs(model.AL)))


def test_csr_batch_matches_single_instance_solves():
    """The same instances solved singly (scaled by build_scaled, power-method
    lambda): the batch's iterations and objectives agree."""
    from hprqp import HPRQP_parameters, optimize
    from hprqp.batch import BatchModel

    models = _models([300, 1100, 700, 1800], seed0=300)
    bm = BatchModel(models)
    res = bm.solve(HPRQP_parameters(verbose=False))
    for i, model in enumerate(models):
        single = optimize(model, HPRQP_parameters(verbose=False))
        rel = abs(single.primal_obj - res.primal_obj[i]) / (1 + abs(single.primal_obj))
        dy = np.max(np.abs(single.y - res.y[i])) / (1 + np.max(np.abs(single.y)))
        print(f"  inst {i}: n={model.n} batch iter={res.iter[i]} single iter={single.iter} rel obj {rel:.1e} "
              f"y diff {dy:.1e}")
        assert single.status == res.status[i] == "OPTIMAL"
        assert _iter_ok(res.iter[i], single.iter)
        assert rel < 1e-6 and dy < 1e-3


def test_csr_batch_warm_start_at_solution_retires_at_once():
    """Warm starts go in, and results come out, in the original space: the
    solution handed back as (x0, y0) is optimal at iteration 0."""
    from hprqp import HPRQP_parameters
    from hprqp.batch import BatchModel

    models = _models([250, 800, 1300], seed0=500)
    bm = BatchModel(models)
    cold = bm.solve(HPRQP_parameters(verbose=False))
    warm = bm.solve(HPRQP_parameters(verbose=False), x0=cold.x, y0=cold.y)
    print(f"cold iters {list(cold.iter)}, warm iters {list(warm.iter)}")
    assert warm.solved == bm.count and np.all(warm.iter == 0)


def test_csr_batch_with_a_box_only_instance_and_scaling_off():
    """An instance with no linear constraints (m = 0) beside constrained ones,
    and the same batch with scaling switched off: each at the oracle's bar
    under the same parameters."""
    from hprqp import HPRQP_parameters, build_from_QAbc
    from hprqp.batch import BatchModel
    from oracle import oracle

    rng = np.random.default_rng(5)
    n = 300
    R = sp.random(60, n, density=0.05, random_state=6, format="csr")
    Q = (R.T @ R + sp.diags(rng.uniform(0.01, 1.0, n))).tocsr()
    box = build_from_QAbc(Q, rng.normal(size=n), sp.csr_matrix((0, n)), np.zeros(0), np.zeros(0),
                          np.full(n, -1.0), np.full(n, 1.0), verbose=False)
    models = [box] + _models([400, 900], seed0=700)
    for scaling in (True, False):
        kw = dict(use_Ruiz_scaling=scaling, use_Pock_Chambolle_scaling=scaling)
        res = BatchModel(models).solve(HPRQP_parameters(verbose=False, **kw))
        for i, model in enumerate(models):
            ref = oracle.solve(model, **kw)
            rel = abs(res.primal_obj[i] - ref.primal_obj) / (1 + abs(ref.primal_obj))
            print(f"scaling {scaling} inst {i}: m={model.m} iter={res.iter[i]} (oracle {ref.iter}) rel {rel:.1e}")
            assert res.status[i] == "OPTIMAL" and rel < 1e-6 and _iter_ok(res.iter[i], ref.iter)


def test_csr_batch_refuses_other_paths():
    """A diagonal Q (diagonal-Q path) or mostly free variables (noC path) is
    refused with HPRQP_E_UNSUP and a message; a batch mixing Q kinds is
    refused by the Python host."""
    from hprqp import build_from_QAbc, instances
    from hprqp.batch import BatchModel

    n = 50
    rng = np.random.default_rng(1)
    A = sp.random(20, n, density=0.2, random_state=2, format="csr") + sp.eye(20, n)
    b = A @ rng.uniform(0, 1, n)
    diag = build_from_QAbc(sp.diags(rng.uniform(0.1, 1, n)).tocsr(), rng.normal(size=n), A, b, b,
                           np.zeros(n), np.full(n, 5.0), verbose=False)
    with pytest.raises(RuntimeError, match="off-diagonal"):
        BatchModel([instances.random_sparse_qp(200, 100, seed=1), diag])
    R = sp.random(10, n, density=0.3, random_state=3, format="csr")
    free = build_from_QAbc((R.T @ R + sp.eye(n)).tocsr(), rng.normal(size=n), A, b, b, np.full(n, -np.inf),
                           np.full(n, np.inf), verbose=False)
    with pytest.raises(RuntimeError, match="noC"):
        BatchModel([free])
    with pytest.raises(TypeError):
        BatchModel([instances.random_sparse_qp(200, 100, seed=1), instances.portfolio(50, 5, seed=1, explicit=False)[0]])
