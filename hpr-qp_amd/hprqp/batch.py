"""Batched solves: thousands of instances in one launch per check interval.

Each instance is a QP_info_cpu whose A is sparse and whose Q is either a
DiagonalPlusLowRank_Q_operator (diag + U U^T; hprqp_batch_create) or a sparse
matrix (hprqp_batch_create_csr: every instance scaled as the single solve
scales it, then the same iteration).  One batch holds one kind.  The batch is
packed once on the GPU; solve() runs every instance with its own restart and
sigma state until each one reaches the tolerance (hprqp_batch_solve).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib
from .model import DiagonalPlusLowRank_Q_operator
from .solver import STATUS, HPRQP_parameters


@dataclass
class BatchResults:
    status: list
    iter: np.ndarray
    primal_obj: np.ndarray
    residuals: np.ndarray
    x: list
    y: list
    time: float
    power_time: float
    solved: int
    intervals: int
    total_iter: int
    predict_time: float = 0.0
    interval_time: float = 0.0  # seconds of the check-interval launches (HIP events)


class BatchModel:
    def __init__(self, models, device: int = 0):
        self.ctx = _lib.context(device)
        self.models = list(models)
        cnt = len(self.models)
        n = np.array([mdl.n for mdl in self.models], dtype=np.int32)
        m = np.array([mdl.m for mdl in self.models], dtype=np.int32)
        lowrank = [isinstance(mdl.Q, DiagonalPlusLowRank_Q_operator) for mdl in self.models]
        sparse = [sp.issparse(mdl.Q) for mdl in self.models]
        if not (all(lowrank) or all(sparse)):
            raise TypeError("batched instances need a DiagonalPlusLowRank_Q_operator Q or a sparse-matrix Q, "
                            "the same kind for every instance")
        self.kind = "lowrank" if all(lowrank) else "csr"
        k = np.zeros(cnt, dtype=np.int32)
        U, d, rp, ci, av, qrp, qci, qv = [], [], [], [], [], [], [], []
        for i, mdl in enumerate(self.models):
            Q = mdl.Q
            if self.kind == "lowrank":
                k[i] = Q.U.shape[1]
                U.append(np.asarray(Q.U, dtype=np.float64).ravel(order="F"))
                d.append(Q.diag)
            else:
                Qc = sp.csr_matrix(Q, dtype=np.float64)
                Qc.sort_indices()
                qrp.append(Qc.indptr.astype(np.int32))
                qci.append(Qc.indices.astype(np.int32))
                qv.append(Qc.data)
            A = sp.csr_matrix(mdl.A, dtype=np.float64)
            A.sort_indices()
            rp.append(A.indptr.astype(np.int32))
            ci.append(A.indices.astype(np.int32))
            av.append(A.data)
        cat = lambda lst, dt: np.ascontiguousarray(np.concatenate(lst) if lst else np.zeros(0), dtype=dt)
        self.arrays = dict(
            n=n, m=m, k=k, U=cat(U, np.float64), d=cat(d, np.float64), rp=cat(rp, np.int32), ci=cat(ci, np.int32),
            av=cat(av, np.float64), qrp=cat(qrp, np.int32), qci=cat(qci, np.int32), qv=cat(qv, np.float64),
            c=cat([mdl.c for mdl in self.models], np.float64),
            l=cat([mdl.l for mdl in self.models], np.float64), u=cat([mdl.u for mdl in self.models], np.float64),
            AL=cat([mdl.AL for mdl in self.models], np.float64), AU=cat([mdl.AU for mdl in self.models], np.float64),
            c0=np.array([mdl.obj_constant for mdl in self.models], dtype=np.float64))
        a = self.arrays
        h = C.c_void_p()
        P = _lib.ptr
        if self.kind == "lowrank":
            _lib.check(self.ctx.h, self.ctx.lib.hprqp_batch_create(
                self.ctx.h, cnt, P(a["n"]), P(a["m"]), P(a["k"]), P(a["U"]), P(a["d"]), a["ci"].size, P(a["rp"]),
                P(a["ci"]), P(a["av"]), P(a["c"]), P(a["l"]), P(a["u"]), P(a["AL"]), P(a["AU"]), P(a["c0"]),
                C.byref(h)))
        else:
            _lib.check(self.ctx.h, self.ctx.lib.hprqp_batch_create_csr(
                self.ctx.h, cnt, P(a["n"]), P(a["m"]), a["qci"].size, P(a["qrp"]), P(a["qci"]), P(a["qv"]),
                a["ci"].size, P(a["rp"]), P(a["ci"]), P(a["av"]), P(a["c"]), P(a["l"]), P(a["u"]), P(a["AL"]),
                P(a["AU"]), P(a["c0"]), C.byref(h)))
        self.h = h
        self.count = cnt
        self.noff = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
        self.moff = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)

    def pack_x(self, xs):
        return np.ascontiguousarray(np.concatenate(xs), dtype=np.float64)

    def pack_y(self, ys):
        return np.ascontiguousarray(np.concatenate(ys) if len(ys) else np.zeros(0), dtype=np.float64)

    def predict(self, net):
        """Learned warm start of every instance, computed on the device by
        hprqp_batch_predict; returns lists (x0_i, y0_i) like
        MLP.predict_batch does on the host."""
        from .warmstart import Y_SCALE

        N, M = int(self.noff[-1]), int(self.moff[-1])
        x = np.zeros(max(N, 1))
        y = np.zeros(max(M, 1))
        mean = np.ascontiguousarray(net.feat_mean, dtype=np.float32)
        std = np.ascontiguousarray(net.feat_std, dtype=np.float32)
        P = _lib.ptr
        _lib.check(self.ctx.h, self.ctx.lib.hprqp_batch_predict(self.ctx.h, self.h, net.h, P(mean), P(std),
                                                                Y_SCALE, P(x), P(y)))
        return ([x[self.noff[i]:self.noff[i + 1]].copy() for i in range(self.count)],
                [y[self.moff[i]:self.moff[i + 1]].copy() for i in range(self.count)])

    def solve(self, params: HPRQP_parameters = None, x0=None, y0=None, warm=None) -> BatchResults:
        """Solve every instance.  x0 / y0: optional warm starts (lists per
        instance or packed arrays).  warm: an MLP whose prediction, computed
        on the device, is the warm start (hprqp_batch_solve_warm); the
        prediction time is in BatchResults.predict_time."""
        params = params or HPRQP_parameters(verbose=False)
        p = params.to_c()
        cnt = self.count
        N, M = int(self.noff[-1]), int(self.moff[-1])
        if warm is not None and (x0 is not None or y0 is not None):
            raise ValueError("give either x0 / y0 or warm, not both")
        x0p = None if x0 is None else (self.pack_x(x0) if isinstance(x0, (list, tuple)) else np.ascontiguousarray(x0))
        y0p = None if y0 is None else (self.pack_y(y0) if isinstance(y0, (list, tuple)) else np.ascontiguousarray(y0))
        if x0p is not None and x0p.size != N:
            raise ValueError("packed x0 has the wrong length")
        if y0p is not None and y0p.size != M:
            raise ValueError("packed y0 has the wrong length")
        x = np.zeros(N)
        y = np.zeros(max(M, 1))
        st = np.zeros(cnt, dtype=np.int32)
        it = np.zeros(cnt, dtype=np.int64)
        obj = np.zeros(cnt)
        kkt = np.zeros(cnt)
        r = _lib.hprqp_batch_result()
        P = _lib.ptr
        if warm is None:
            _lib.check(self.ctx.h, self.ctx.lib.hprqp_batch_solve(
                self.ctx.h, self.h, C.byref(p), P(x0p), P(y0p), P(x), P(y), P(st), P(it), P(obj), P(kkt),
                C.byref(r)))
        else:
            from .warmstart import Y_SCALE

            mean = np.ascontiguousarray(warm.feat_mean, dtype=np.float32)
            std = np.ascontiguousarray(warm.feat_std, dtype=np.float32)
            _lib.check(self.ctx.h, self.ctx.lib.hprqp_batch_solve_warm(
                self.ctx.h, self.h, warm.h, P(mean), P(std), Y_SCALE, C.byref(p), P(x), P(y), P(st), P(it), P(obj),
                P(kkt), C.byref(r)))
        xs = [x[self.noff[i]:self.noff[i + 1]].copy() for i in range(cnt)]
        ys = [y[self.moff[i]:self.moff[i + 1]].copy() for i in range(cnt)]
        return BatchResults([STATUS.get(int(s), str(int(s))) for s in st], it, obj, kkt, xs, ys, r.time,
                            r.power_time, r.solved, r.intervals, r.total_iter, r.predict_time,
                            r.interval_time)

    def close(self):
        if self.h:
            self.ctx.lib.hprqp_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
