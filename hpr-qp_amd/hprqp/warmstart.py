"""Learned warm start for batches of diagonal-plus-low-rank QPs.

A per-variable MLP (hprqp_mlp_*, HIP kernels) maps permutation-equivariant
per-variable features to the variable's scaled primal value and a dual
contribution; the instance's dual estimate is the mean of the contributions.
Predictions are handed to the solver as initial_x / initial_y, so the KKT
residual still decides termination; cold start stays the fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

N_FEATURES = 12
Y_SCALE = 10.0


def features(model) -> np.ndarray:
    """Per-variable features of one instance (rows = variables).

    Every feature of variable i depends on variable i's own data and on
    permutation-invariant summaries of the instance, so permuting the
    variables permutes the rows."""
    Q = model.Q
    n = model.n
    d = Q.diag
    U = Q.U
    c = model.c
    u = model.u
    qii = d + np.einsum("ij,ij->i", U, U)
    ubar = U.mean(axis=0)
    expo = U @ ubar
    cm, cs = c.mean(), c.std() + 1e-12
    ratio = -c / qii
    rank_c = np.argsort(np.argsort(c)) / max(n - 1, 1)
    rank_r = np.argsort(np.argsort(-ratio)) / max(n - 1, 1)
    capn = np.where(np.isfinite(u), u, 1.0) * n
    f = np.stack([
        (c - cm) / cs,
        c * 20.0,
        d * 10.0,
        (qii - d) * 3.0,
        expo * 30.0,
        ratio * 0.5,
        rank_c,
        rank_r,
        np.full(n, 100.0 / n),
        np.full(n, U.shape[1] / 50.0),
        np.minimum(capn, 100.0) / 100.0,
        np.full(n, cs * 50.0),
    ], axis=1)
    return f.astype(np.float32)


def targets(model, x, y):
    """Scaled training targets: x_i / u_i per variable, Y_SCALE * y for the
    instance's first constraint."""
    u = np.where(np.isfinite(model.u), model.u, 1.0)
    return (x / u).astype(np.float32), np.float32(Y_SCALE * (y[0] if len(y) else 0.0))


class MLP:
    """The HIP MLP (forward, gradient, Adam training on the GPU)."""

    def __init__(self, n_in=N_FEATURES, hidden=128, seed=0, device=0):
        self.ctx = _lib.context(device)
        h = C.c_void_p()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_mlp_create(self.ctx.h, n_in, hidden, seed, C.byref(h)))
        self.h = h
        self.n_in = n_in
        self.hidden = hidden
        self.feat_mean = np.zeros(n_in, dtype=np.float32)
        self.feat_std = np.ones(n_in, dtype=np.float32)

    @property
    def n_params(self):
        return int(self.ctx.lib.hprqp_mlp_param_count(self.h))

    def get_params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_mlp_get_params(self.ctx.h, self.h, _lib.ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_mlp_set_params(self.ctx.h, self.h, _lib.ptr(p)))

    def _norm(self, X):
        return np.ascontiguousarray((X - self.feat_mean) / self.feat_std, dtype=np.float32)

    def forward(self, X, normalize=True):
        X = self._norm(X) if normalize else np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], 2), dtype=np.float32)
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_mlp_forward(self.ctx.h, self.h, X.shape[0], _lib.ptr(X),
                                                              _lib.ptr(out)))
        return out

    def loss_grad(self, X, tx, ty, group, wy=1.0, normalize=False):
        X = self._norm(X) if normalize else np.ascontiguousarray(X, dtype=np.float32)
        tx = np.ascontiguousarray(tx, dtype=np.float32)
        ty = np.ascontiguousarray(ty, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double()
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_mlp_loss_grad(self.ctx.h, self.h, ty.size, group, _lib.ptr(X),
                                                                _lib.ptr(tx), _lib.ptr(ty), wy, C.byref(loss),
                                                                _lib.ptr(g)))
        return loss.value, g

    def fit(self, X, tx, ty, group, epochs=20, batch_groups=256, lr=1e-3, wy=1.0, seed=0):
        self.feat_mean = X.mean(axis=0).astype(np.float32)
        self.feat_std = (X.std(axis=0) + 1e-6).astype(np.float32)
        Xn = self._norm(X)
        tx = np.ascontiguousarray(tx, dtype=np.float32)
        ty = np.ascontiguousarray(ty, dtype=np.float32)
        hist = np.zeros(epochs)
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_mlp_train(self.ctx.h, self.h, ty.size, group, _lib.ptr(Xn),
                                                            _lib.ptr(tx), _lib.ptr(ty), wy, batch_groups, epochs,
                                                            lr, seed, _lib.ptr(hist)))
        return hist

    def predict(self, model):
        """Warm start (x0, y0) for one instance."""
        o = self.forward(features(model))
        u = np.where(np.isfinite(model.u), model.u, 1.0)
        x0 = np.clip(o[:, 0], 0.0, 1.0).astype(np.float64) * u
        y0 = np.full(model.m, float(o[:, 1].mean()) / Y_SCALE)
        return x0, y0

    def predict_batch(self, models):
        feats = [features(m) for m in models]
        sizes = [f.shape[0] for f in feats]
        o = self.forward(np.concatenate(feats, axis=0))
        xs, ys, r0 = [], [], 0
        for m, s in zip(models, sizes):
            oo = o[r0:r0 + s]
            r0 += s
            u = np.where(np.isfinite(m.u), m.u, 1.0)
            xs.append(np.clip(oo[:, 0], 0.0, 1.0).astype(np.float64) * u)
            ys.append(np.full(m.m, float(oo[:, 1].mean()) / Y_SCALE))
        return xs, ys

    def save(self, path):
        np.savez(path, params=self.get_params(), feat_mean=self.feat_mean, feat_std=self.feat_std,
                 n_in=self.n_in, hidden=self.hidden)

    @classmethod
    def load(cls, path, device=0):
        z = np.load(path, allow_pickle=False)
        net = cls(int(z["n_in"]), int(z["hidden"]), device=device)
        net.set_params(z["params"])
        net.feat_mean = z["feat_mean"]
        net.feat_std = z["feat_std"]
        return net

    def close(self):
        if self.h:
            self.ctx.lib.hprqp_mlp_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def training_set(models, xs, ys, group=64, seed=0):
    """Sample `group` variables per solved instance: rows (X, tx) and one
    dual target per instance."""
    rng = np.random.default_rng(seed)
    X, TX, TY = [], [], []
    for m, x, y in zip(models, xs, ys):
        f = features(m)
        tx, ty = targets(m, x, y)
        idx = rng.choice(m.n, size=group, replace=m.n < group)
        X.append(f[idx])
        TX.append(tx[idx])
        TY.append(ty)
    return np.concatenate(X), np.concatenate(TX), np.array(TY, dtype=np.float32)
