"""CPU checks of the warm-start pieces: the numpy MLP restatement against
finite differences, and the features' permutation equivariance."""
import numpy as np


def test_mlp_oracle_gradient_matches_finite_differences():
    from oracle import mlp_oracle

    rng = np.random.default_rng(0)
    n_in, hid, group, G = 5, 7, 3, 4
    npar = n_in * hid + hid + 2 * (hid * hid + hid) + hid * 2 + 2
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((G * group, n_in))
    tx = rng.standard_normal(G * group)
    ty = rng.standard_normal(G)
    L, g = mlp_oracle.loss_grad(p, X, tx, ty, group, n_in, hid, wy=0.7)
    eps = 1e-6
    for idx in rng.choice(npar, size=40, replace=False):
        e = np.zeros(npar)
        e[idx] = eps
        Lp, _ = mlp_oracle.loss_grad(p + e, X, tx, ty, group, n_in, hid, wy=0.7)
        Lm, _ = mlp_oracle.loss_grad(p - e, X, tx, ty, group, n_in, hid, wy=0.7)
        assert abs((Lp - Lm) / (2 * eps) - g[idx]) < 1e-6 * (1 + abs(g[idx]))


def test_features_are_permutation_equivariant():
    from hprqp import instances
    from hprqp.model import DiagonalPlusLowRank_Q_operator, QP_info_cpu
    from hprqp.warmstart import features

    model, f = instances.portfolio(120, 7, seed=4, explicit=False)
    perm = np.random.default_rng(1).permutation(model.n)
    Q = model.Q
    pm = QP_info_cpu(DiagonalPlusLowRank_Q_operator(Q.diag[perm], Q.U[perm]), model.c[perm], model.A[:, perm],
                     model.AL, model.AU, model.l[perm], model.u[perm])
    np.testing.assert_allclose(features(pm), features(model)[perm], rtol=1e-5, atol=1e-6)
