"""Train the warm-start MLP on solutions this solver generates (GPU).

    cd hpr-qp_amd && python -m hprqp.train_warmstart --solutions 50000 \
        --out hprqp/data/warm_mlp.npz

Instances are long-only factor-model portfolio QPs drawn like the batched
workload (n ~ U{200..5000}, k ~ U{10..50}) from seeds disjoint from the
evaluation batch (which uses instance seeds 1..4096); they are solved cold by
the batched mixed-precision solver, 64 variables per solution are sampled as
training rows, and the network is fitted with the HIP Adam kernels.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from . import HPRQP_parameters, instances
from .batch import BatchModel
from .warmstart import MLP, training_set


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--solutions", type=int, default=50_000)
    ap.add_argument("--chunk", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=100_000)
    ap.add_argument("--group", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="hprqp/data/warm_mlp.npz")
    args = ap.parse_args(argv)
    params = HPRQP_parameters(verbose=False, precision="mixed")
    t0 = time.time()
    Xs, TXs, TYs = [], [], []
    done = 0
    c = 0
    while done < args.solutions:
        cnt = min(args.chunk, args.solutions - done)
        insts = instances.portfolio_batch(cnt, seed=args.seed + c * 100_000)
        models = [m for m, _ in insts]
        bm = BatchModel(models)
        r = bm.solve(params)
        ok = [i for i, s in enumerate(r.status) if s == "OPTIMAL"]
        X, TX, TY = training_set([models[i] for i in ok], [r.x[i] for i in ok], [r.y[i] for i in ok],
                                 group=args.group, seed=c)
        Xs.append(X)
        TXs.append(TX)
        TYs.append(TY)
        done += cnt
        c += 1
        print(f"[train] chunk {c}: {len(ok)}/{cnt} solved in {r.time:.2f}s, {done} solutions, "
              f"{time.time() - t0:.0f}s elapsed", flush=True)
        bm.close()
    X, TX, TY = np.concatenate(Xs), np.concatenate(TXs), np.concatenate(TYs)
    net = MLP(seed=0)
    t = time.time()
    hist = net.fit(X, TX, TY, args.group, epochs=args.epochs, batch_groups=256, lr=args.lr)
    print(f"[train] {X.shape[0]} rows from {TY.size} solutions, {args.epochs} epochs in {time.time() - t:.1f}s, "
          f"loss {hist[0]:.5f} -> {hist[-1]:.5f}", flush=True)
    net.save(args.out)
    print(f"[train] saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
