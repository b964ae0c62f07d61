"""Learned warm start on the GPU: HIP forward/backward against the FP64
numpy restatement, training progress, and the gain over cold start with the
tolerance unchanged (every instance still certified by its KKT residual)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_hip_mlp_matches_numpy_restatement():
    from hprqp.warmstart import MLP
    from oracle import mlp_oracle

    rng = np.random.default_rng(0)
    net = MLP(n_in=12, hidden=128, seed=3)
    p = net.get_params()
    group, G = 16, 40
    X = rng.standard_normal((G * group, 12)).astype(np.float32)
    tx = rng.uniform(0, 1, G * group).astype(np.float32)
    ty = rng.standard_normal(G).astype(np.float32)
    out = net.forward(X, normalize=False)
    ref = mlp_oracle.forward(p, X, 12, 128)[-1]
    np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-5)
    L, g = net.loss_grad(X, tx, ty, group, wy=0.5)
    Lr, gr = mlp_oracle.loss_grad(p.astype(np.float64), X, tx, ty, group, 12, 128, wy=0.5)
    assert abs(L - Lr) <= 1e-5 * (1 + abs(Lr))
    scale = np.abs(gr).max()
    assert np.max(np.abs(g - gr)) <= 2e-4 * scale, np.max(np.abs(g - gr)) / scale


def test_training_reduces_loss():
    from hprqp.warmstart import MLP

    rng = np.random.default_rng(1)
    group, G = 8, 512
    X = rng.standard_normal((G * group, 12)).astype(np.float32)
    tx = (np.tanh(X[:, 0]) * 0.5 + 0.5).astype(np.float32)
    ty = X[:, 1].reshape(G, group).mean(axis=1).astype(np.float32)
    net = MLP(seed=2)
    hist = net.fit(X, tx, ty, group, epochs=15, batch_groups=64, lr=2e-3)
    print("loss history", hist[0], hist[-1])
    assert hist[-1] < 0.2 * hist[0]


def test_learned_warm_start_gain_on_held_out_batch():
    """Train on solver-generated solutions of one set of instances, then
    warm-start a disjoint set; every warm instance must still reach 1e-6 KKT,
    and the network must cut the total iteration count."""
    from hprqp import HPRQP_parameters, instances
    from hprqp.batch import BatchModel
    from hprqp.warmstart import MLP, training_set

    params = HPRQP_parameters(verbose=False)
    train = instances.portfolio_batch(768, seed=1000, n_range=(50, 400), k_range=(3, 12))
    bt = BatchModel([m for m, _ in train])
    sol = bt.solve(params)
    assert sol.solved == len(train)
    X, TX, TY = training_set([m for m, _ in train], sol.x, sol.y, group=32)
    net = MLP(seed=7)
    hist = net.fit(X, TX, TY, 32, epochs=40, batch_groups=64, lr=1e-3)
    test = instances.portfolio_batch(128, seed=5000, n_range=(50, 400), k_range=(3, 12))
    bm = BatchModel([m for m, _ in test])
    cold = bm.solve(params)
    x0, y0 = net.predict_batch([m for m, _ in test])
    warm = bm.solve(params, x0=x0, y0=y0)
    print(f"train loss {hist[0]:.4f} -> {hist[-1]:.4f}; cold iters {cold.iter.sum()} time {cold.time:.3f}s; "
          f"warm iters {warm.iter.sum()} time {warm.time:.3f}s")
    assert cold.solved == len(test) and warm.solved == len(test)
    assert np.all(warm.residuals < 1e-6)
    for i in range(len(test)):
        assert abs(warm.primal_obj[i] - cold.primal_obj[i]) / (1 + abs(cold.primal_obj[i])) < 1e-5
    assert warm.iter.sum() < cold.iter.sum()


def _trained():
    from pathlib import Path

    from hprqp.warmstart import MLP

    return MLP.load(str(Path(__file__).resolve().parents[1] / "hpr-qp_amd" / "hprqp" / "data" / "warm_mlp.npz"))


def test_device_prediction_matches_host_features_path():
    """hprqp_batch_predict (features, MLP and output map on the device) against
    the host path (numpy features, then the HIP MLP)."""
    from hprqp import instances
    from hprqp.batch import BatchModel

    net = _trained()
    insts = instances.portfolio_batch(12, seed=321, n_range=(30, 900), k_range=(2, 40))
    models = [m for m, _ in insts]
    bm = BatchModel(models)
    hx, hy = net.predict_batch(models)
    dx, dy = bm.predict(net)
    for i in range(len(models)):
        np.testing.assert_allclose(dx[i], hx[i], rtol=0, atol=2e-6 * max(1.0, np.abs(hx[i]).max()))
        np.testing.assert_allclose(dy[i], hy[i], rtol=1e-5, atol=1e-7)


def test_solve_warm_on_device_matches_host_warm_start():
    from hprqp import HPRQP_parameters, instances
    from hprqp.batch import BatchModel

    net = _trained()
    insts = instances.portfolio_batch(48, seed=654, n_range=(200, 1500), k_range=(10, 50))
    models = [m for m, _ in insts]
    bm = BatchModel(models)
    params = HPRQP_parameters(verbose=False, precision="mixed")
    cold = bm.solve(params)
    x0, y0 = net.predict_batch(models)
    host = bm.solve(params, x0=x0, y0=y0)
    dev = bm.solve(params, warm=net)
    print(f"cold iters {cold.iter.sum()} host-warm {host.iter.sum()} device-warm {dev.iter.sum()} "
          f"predict {dev.predict_time:.3f}s")
    assert dev.solved == len(models) and np.all(dev.residuals < 1e-6)
    assert dev.predict_time > 0
    assert abs(int(dev.iter.sum()) - int(host.iter.sum())) <= 0.02 * host.iter.sum()
    assert dev.iter.sum() < cold.iter.sum()
    for i in range(len(models)):
        assert abs(dev.primal_obj[i] - cold.primal_obj[i]) / (1 + abs(cold.primal_obj[i])) < 1e-5
