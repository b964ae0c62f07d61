"""FP64 numpy restatement of the warm-start MLP's loss and gradient.

TEST INFRASTRUCTURE ONLY (checks the HIP kernels of hprqp_mlp.hip).  Layout
of the flat parameter vector: W1 (n_in x h, row-major), b1, W2, b2, W3, b3,
W4 (h x 2), b4; ReLU after layers 1-3; loss
  mean_r (o_r0 - tx_r)^2 + wy * mean_g (mean_{r in g} o_r1 - ty_g)^2.
The reference has no learned component; this restatement is pinned by a
finite-difference test (tests/test_mlp_cpu.py).
"""
import numpy as np


def unpack(p, n_in, hidden):
    w = [n_in, hidden, hidden, hidden, 2]
    out, o = [], 0
    for l in range(4):
        W = p[o:o + w[l] * w[l + 1]].reshape(w[l], w[l + 1])
        o += w[l] * w[l + 1]
        b = p[o:o + w[l + 1]]
        o += w[l + 1]
        out.append((W, b))
    assert o == p.size
    return out


def forward(p, X, n_in, hidden):
    layers = unpack(np.asarray(p, dtype=np.float64), n_in, hidden)
    acts = [np.asarray(X, dtype=np.float64)]
    for l, (W, b) in enumerate(layers):
        h = acts[-1] @ W + b
        acts.append(np.maximum(h, 0.0) if l < 3 else h)
    return acts


def loss_grad(p, X, tx, ty, group, n_in, hidden, wy=1.0):
    p = np.asarray(p, dtype=np.float64)
    layers = unpack(p, n_in, hidden)
    acts = forward(p, X, n_in, hidden)
    O = acts[-1]
    R = O.shape[0]
    G = R // group
    ex = O[:, 0] - tx
    ym = O[:, 1].reshape(G, group).mean(axis=1)
    ey = ym - ty
    loss = np.mean(ex * ex) + wy * np.mean(ey * ey)
    dO = np.zeros_like(O)
    dO[:, 0] = 2.0 * ex / R
    dO[:, 1] = np.repeat(wy * 2.0 * ey / (G * group), group)
    grads = []
    d = dO
    for l in range(3, -1, -1):
        W, b = layers[l]
        grads.append((acts[l].T @ d, d.sum(axis=0)))
        if l > 0:
            d = (d @ W.T) * (acts[l] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
