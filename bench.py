#!/usr/bin/env python3
"""bench.py -- HPR-QP on one MI355X: solves/sec to 1e-6 relative KKT.

Workload (one "step" = one complete solve, inputs resident in HBM): the
structured grid QP, 1,414 x 1,414 cells (n = 1,999,396, m = 999,698),
Q = 5-point Laplacian + 1e-3 I, A = horizontal-edge divergence operator,
seed 0 -- the largest configuration that is quoted single-instance.

Also reported: `roofline` for the dominant kernel (the fused z/x/w step
whose row product is Q w), measured live with HIP events inside libhprqp on
the context stream; `cpu_baseline` -- the CPU oracle (a restatement of the
reference CPU path, 1 thread) timed on a bounded sample of the same
workload and extrapolated to one solve; and `batched` -- 4,096 portfolio QPs
solved by the batched mixed-precision solver, cold and warm-started by the
learned MLP, against the plain FP64 form solving them one per call.

--workload random|lasso|qap runs the same single-instance measurement on
the other workloads (random sparse QP n=10,000; LASSO 3,308 x 150,360; QAP
128 x 128), --workload batch only the batched leg, --workload qap-apply the
QAP operator apply alone at several sizes; these are for profiling.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hpr-qp_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def zxw1_bytes(nnz_q, n, prec, bound_streams=2, dia_nd=None):
    """Algorithmic HBM bytes of one z/x/w1 launch (partial write set).
    fp64 / mixed: Q values+indices, row pointers, gathered w, streamed ATy c x,
    the bound vectors l u that are not a single value (0 or +-inf for every
    variable is passed as a scalar), the FP64 anchor last_x; written Qw x x_hat
    w_bar (mixed: FP32 Q values, c, l, u; FP64 iterates).  fp32 (the offset form): FP32 Q values + indices, row pointers, the
    gathered FP32 offset dW, the row's dX dT dW (FP32), the FP64 anchor
    vectors G0 X0, the FP32 anchor difference D0 and the FP64 bounds that are
    not one value; written Q dW, dX, dX_hat, dW_bar (FP32).  With Q in the
    diagonal form (dia_nd diagonals) the values are the dia_nd x n diagonal
    arrays, padding included, and there are no indices or row pointers."""
    if prec == "fp32":
        qb = dia_nd * n * 4 if dia_nd else nnz_q * 8 + 4 * (n + 1)
        return qb + 4 * n + 12 * n + 20 * n + 8 * bound_streams * n + 16 * n
    v = 8 if prec == "fp64" else 4  # matrix values and the data c, l, u
    qb = dia_nd * n * v if dia_nd else nnz_q * (v + 4) + 4 * (n + 1)
    # w, ATy, x, last_x and the four written vectors are FP64 in both forms
    return qb + 8 * n + 8 * n + v * n + 8 * n + bound_streams * v * n + 8 * n + 4 * 8 * n


def q_diagonals(Q):
    """The number of diagonals (distinct j - i) of Q's symmetric pattern, the
    dia_nd of its diagonal form."""
    import numpy as np
    import scipy.sparse as sp

    P = sp.csr_matrix(Q, copy=True)
    P = (abs(P) + abs(P.T)).tocsr()
    rows = np.repeat(np.arange(P.shape[0]), np.diff(P.indptr))
    return int(np.unique(P.indices - rows).size)


def bound_streams(model):
    """How many of l, u the z/x/w1 kernel streams (the rule of the solver:
    one value for every variable, 0 or +-inf, is passed as a scalar)."""
    import numpy as np

    def uniform(v):
        v = np.asarray(v)
        return v.size > 0 and (v[0] == 0 or np.isinf(v[0])) and bool(np.all(v == v[0]))

    return 2 - int(uniform(model.l)) - int(uniform(model.u))


def lasso_step_bytes(nnz, m, n, prec):
    """Algorithmic HBM bytes of the timed LASSO step (Q w = A^T (A w), two
    launches): A and A^T values+indices and row pointers, gathered w and
    temp, temp written; epilogue reads ATy c x w and the FP64 anchors
    last_x last_w, writes Qw x x_hat w_bar w."""
    v = 8 if prec == "fp64" else 4
    return 2 * nnz * (v + 4) + 4 * (m + n + 2) + v * n + 2 * v * m + 9 * v * n + 16 * n


def batch_iter_bytes(n, k, m, nnz_a):
    """Algorithmic HBM bytes of one partial iteration of one batched instance
    (mixed: FP32 U, d, c, l, u, A values; FP64 iterates; 4-byte indices):
    two passes over U (Q w = d.*w + U (U'w) with U'w kept in k-space, then
    U' x_hat) = 8 n k; the z/x/w1 row pass reads w x ATy last_x (FP64), d c l
    u (FP32) and writes x x_hat w_bar v = 80 n; U' x_hat reads x_hat = 8 n;
    the y step reads A's values and indices and gathers v = 8 nnz + 8 n; the
    w2 step reads A^T's row pointers, indices and values, w w_bar ATy last_w
    last_ATy and writes w ATy = 4 n + 8 nnz + 56 n; the m-vectors and A U
    (m x k) are negligible beside these."""
    return 8.0 * n * k + 156.0 * n + 16.0 * nnz_a


def measured_batch_traffic():
    """HBM bytes of the batched leg's k_batch_interval launches over their
    algorithmic bytes, from the FETCH_SIZE / WRITE_SIZE counter runs
    summarised under profiles/ (None when not measured)."""
    f = ROOT / "profiles" / "portfolio_batch4096_interval_traffic.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_over_algorithmic:"):
            return float(line.split(":")[1])
    return None


def measured_traffic(workload_name, grid, prec):
    """HBM bytes per launch of the dominant kernel, from the FETCH_SIZE /
    WRITE_SIZE counter runs summarised under profiles/ (None when that
    workload has not been measured)."""
    f = ROOT / "profiles" / f"grid{grid}_{prec}_zxw1_traffic.md"
    if workload_name != "grid" or not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def workload(name, grid):
    """(model, problem_type, description) of a single-instance workload."""
    from hprqp import instances

    if name == "grid":
        m = instances.grid_qp(grid, seed=0)
        return m, "QP", f"grid QP {grid}x{grid} (n={m.n}, m={m.m})"
    if name == "random":
        m = instances.random_sparse_qp(10_000, 5_000, seed=0)
        return m, "QP", f"random sparse QP (n={m.n}, m={m.m}, nnz(Q)={m.Q.nnz})"
    if name == "lasso":
        m = instances.lasso(seed=0)
        return m, "LASSO", f"sparse LASSO, A {m.Q.A.shape[0]}x{m.Q.A.shape[1]} with {m.Q.A.nnz} nnz"
    if name == "qap":
        m = instances.qap(128, seed=0)[0]
        return m, "QAP", "QAP relaxation 128x128 (n=16384, m=256), flow/distance"
    raise ValueError(name)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def batched_leg(count, plain_sample):
    """The batched workload: `count` long-only factor-model portfolio QPs
    (n ~ U{200..5000}, k ~ U{10..50}, instance seeds 1..count).  Reports
    solves/s of (a) the plain FP64 form (one instance per call, cold,
    sequential, model resident in HBM; timed on the first `plain_sample`
    instances), (b) batched mixed cold, (c) batched mixed warm-started by the
    learned MLP, features and inference on the device (their time included).  Every solve must reach 1e-6
    relative KKT."""
    import numpy as np

    from hprqp import DeviceModel, HPRQP_parameters, instances
    from hprqp.batch import BatchModel
    from hprqp.warmstart import MLP

    t = time.time()
    insts = instances.portfolio_batch(count, seed=0)
    models = [m for m, _ in insts]
    log(f"batched: {count} instances generated in {time.time() - t:.1f}s")
    plain = HPRQP_parameters(verbose=False)
    dms = [DeviceModel(m) for m in models[:plain_sample]]
    dms[0].solve(plain)
    t = time.perf_counter()
    ok = 0
    for dm in dms:
        r = dm.solve(plain)
        ok += r.status == "OPTIMAL" and r.residuals < 1e-6
    plain_sps = ok / (time.perf_counter() - t)
    for dm in dms:
        dm.close()
    bm = BatchModel(models)
    mixed = HPRQP_parameters(verbose=False, precision="mixed")
    bm.solve(mixed)  # warm-up (code objects, graph)
    t = time.perf_counter()
    cold = bm.solve(mixed)
    cold_s = time.perf_counter() - t
    net = MLP.load(str(ROOT / "hpr-qp_amd" / "hprqp" / "data" / "warm_mlp.npz"))
    t = time.perf_counter()
    warm = bm.solve(mixed, warm=net)  # prediction on the device, then the solve
    warm_s = time.perf_counter() - t
    t_pred = warm.predict_time
    per_it = np.array([batch_iter_bytes(m.n, m.Q.U.shape[1], m.m, m.A.nnz) for m in models])
    cold_bytes = float(np.sum(per_it * cold.iter))
    warm_bytes = float(np.sum(per_it * warm.iter))
    ratio = measured_batch_traffic()
    cold_ok = int(np.sum((np.array(cold.status) == "OPTIMAL") & (cold.residuals < 1e-6)))
    warm_ok = int(np.sum((np.array(warm.status) == "OPTIMAL") & (warm.residuals < 1e-6)))
    out = {"workload": f"{count} portfolio QPs, n~U{{200..5000}}, k~U{{10..50}}, 1e-6 relative KKT",
           "plain_fp64_solves_per_s": plain_sps, "plain_sample": plain_sample,
           "batched_mixed_cold_solves_per_s": cold_ok / cold_s,
           "batched_mixed_warm_solves_per_s": warm_ok / warm_s,
           "warm_vs_plain": (warm_ok / warm_s) / plain_sps,
           "cold_mean_iter": float(cold.iter.mean()), "warm_mean_iter": float(warm.iter.mean()),
           "cold_max_iter": int(cold.iter.max()), "warm_max_iter": int(warm.iter.max()),
           "mlp_inference_s": t_pred, "all_solved": cold_ok == count and warm_ok == count,
           "cold_iteration_hbm_gbs": cold_bytes / (cold.time - cold.power_time) / 1e9,
           "cold_iteration_bytes": cold_bytes, "warm_iteration_bytes": warm_bytes,
           "cold_lambda_max_s": cold.power_time, "cold_interval_kernel_s": cold.interval_time,
           "intervals": cold.intervals,
           "roofline": {"bound": "hbm", "kernel": "k_batch_interval<float> (one workgroup per instance, "
                                                  "one launch per check interval)",
                        "achieved": cold_bytes / cold.interval_time / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": cold_bytes / cold.interval_time / 1e9 / HBM_PEAK_GBS,
                        "traffic": None if ratio is None else ratio * cold_bytes,
                        "traffic_over_algorithmic": ratio,
                        "bytes_per_solve": cold_bytes,
                        "traffic_unit": "HBM bytes per cold solve (over all its interval launches), from the "
                                        "counter ratio measured on the same workload",
                        "timing": "HIP events around each check-interval replay of the cold solve"}}
    log("batched: " + json.dumps(out))
    bm.close()
    net.close()
    return out


def batch_csr_iter_bytes(n, m, nnzq, nnza):
    """Algorithmic bytes of one HPR iteration of a sparse-Q instance in the
    batch (FP64): Q w and Q w_bar (12 B per nonzero: value, index, gathered
    entry ~ 8 + 4, plus the row pointers), A tempv and A^T y_bar likewise,
    and the elementwise updates' n- and m-vector streams (z/x/w1 reads 7
    writes 4, tempv 3 + 1, w2 6 + 2; y 4 + 2 per row)."""
    return 2 * (12 * nnzq + 4 * n) + 12 * nnza + 4 * m + 12 * nnza + 4 * n + 8 * (23 * n + 6 * m)


def batched_csr_leg(count, plain_sample):
    """`count` random sparse QPs of the reference-runnable generator at uneven
    sizes (n ~ U{200..2000}, m = n/2, Q = R'R + diag with ~160 nonzeros per
    row, instance seeds 1..count), FP64: plain single-instance GPU solves
    (sequential, models resident; timed on the first `plain_sample`) against
    one batched solve of all of them (hprqp_batch_create_csr: every instance
    scaled and iterated by its own workgroup).  Every solve must reach 1e-6
    relative KKT."""
    import numpy as np

    from hprqp import DeviceModel, HPRQP_parameters, instances
    from hprqp.batch import BatchModel

    rng = np.random.default_rng(2024)
    sizes = rng.integers(200, 2001, count)
    t = time.time()
    models = [instances.random_sparse_qp(int(n), int(n) // 2, seed=1 + i) for i, n in enumerate(sizes)]
    log(f"batched csr: {count} instances generated in {time.time() - t:.1f}s")
    plain = HPRQP_parameters(verbose=False)
    dms = [DeviceModel(m) for m in models[:plain_sample]]
    dms[0].solve(plain)
    t = time.perf_counter()
    ok = 0
    for dm in dms:
        r = dm.solve(plain)
        ok += r.status == "OPTIMAL" and r.residuals < 1e-6
    plain_sps = ok / (time.perf_counter() - t)
    for dm in dms:
        dm.close()
    bm = BatchModel(models)
    bm.solve(plain)  # warm-up
    t = time.perf_counter()
    res = bm.solve(plain)
    wall = time.perf_counter() - t
    solved = int(np.sum((np.array(res.status) == "OPTIMAL") & (res.residuals < 1e-6)))
    per_it = np.array([batch_csr_iter_bytes(m.n, m.m, m.Q.nnz, m.A.nnz) for m in models])
    nbytes = float(np.sum(per_it * res.iter))
    out = {"workload": f"{count} random sparse QPs, n~U{{200..2000}}, m=n/2, FP64, 1e-6 relative KKT",
           "plain_fp64_solves_per_s": plain_sps, "plain_sample": plain_sample,
           "batched_fp64_solves_per_s": solved / wall, "batched_vs_plain": (solved / wall) / plain_sps,
           "all_solved": solved == count, "mean_iter": float(res.iter.mean()), "max_iter": int(res.iter.max()),
           "lambda_max_s": res.power_time, "interval_kernel_s": res.interval_time, "intervals": res.intervals,
           "roofline": {"bound": "hbm", "kernel": "k_batch_interval<double, 1> (one workgroup per instance, "
                                                  "one launch per check interval)",
                        "achieved": nbytes / res.interval_time / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": nbytes / res.interval_time / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "bytes_per_solve": nbytes,
                        "timing": "HIP events around each check-interval replay"}}
    log("batched csr: " + json.dumps(out))
    bm.close()
    return out


def qap_apply_leg(sizes, reps):
    """The QAP operator apply Q(X) = 2(A X B - S X - X T) alone, buffers
    resident in HBM, at several sizes: FP64 MFMA and split-BF16 x3 MFMA,
    diagonal S and T and integer A, B (the QAP-relaxation form: 4 n^3 flops
    per apply).  Times back-to-back applies per size and precision (at least
    reps, and at least ~2 GFLOP of them) between two stream synchronisations;
    at small n the figure includes the host's per-call cost."""
    import numpy as np

    from hprqp import _lib
    from hprqp.model import QAP_Q_operator
    from hprqp.solver import DeviceQOperator

    ctx = _lib.Context(0)
    lib = ctx.lib
    out = {"metric": "QAP operator applies/sec and MFMA TFLOP/s", "unit": "TFLOP/s", "sizes": []}
    for nq in sizes:
        rng = np.random.default_rng(nq)
        F = np.triu(rng.integers(0, 11, (nq, nq)) * (rng.random((nq, nq)) < 0.1), 1).astype(np.float64)
        F = F + F.T
        pts = rng.integers(0, 32, (nq, 2))
        D = (np.abs(pts[:, None, 0] - pts[None, :, 0]) + np.abs(pts[:, None, 1] - pts[None, :, 1])).astype(float)
        I = np.eye(nq)
        op = DeviceQOperator(ctx, QAP_Q_operator(F / 2.0, D, -1.0 * I, -1.0 * I), nq * nq)
        xb = _lib.DeviceBuffer.from_array(ctx, rng.random(nq * nq))
        qb = _lib.DeviceBuffer(ctx, nq * nq)
        fl = 4.0 * nq ** 3
        row = {"n": nq, "flops_per_apply": fl}
        for prec, code, peak in (("fp64", 0, 78.6), ("bf16x3", 1, 2500.0 / 3.0)):
            def run(k):
                for _ in range(k):
                    _lib.check(ctx.h, lib.hprqp_qop_apply_prec(ctx.h, op.h, xb.h, qb.h, code))
                _lib.check(ctx.h, lib.hprqp_ctx_sync(ctx.h))

            k = max(reps, int(2e9 / fl) + 1)  # at least ~2 GFLOP of applies per size
            run(3)
            t = time.perf_counter()
            run(k)
            dt = (time.perf_counter() - t) / k
            tf = fl / dt / 1e12
            row[prec] = {"us_per_apply": dt * 1e6, "tflops": tf, "peak_tflops": peak, "frac": tf / peak}
            log(f"QAP apply n={nq} ({nq * nq} variables) {prec}: {dt * 1e6:.1f} us, {tf:.2f} TFLOP/s "
                f"({100 * tf / peak:.1f} % of {peak:.1f})")
        out["sizes"].append(row)
        op.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=1414)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "fp32"])
    ap.add_argument("--profile-reps", type=int, default=50)
    ap.add_argument("--spmv-lanes", type=int, default=0,
                    help="row-product shape of the single-instance solves (0 = autotuned; -2000 pins the diagonal "
                         "form, as counter runs do, whose timings under the profiler can flip the autotuner)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=300,
                    help="HPR iterations of the CPU oracle sample (about 12 s on the 1414^2 grid)")
    ap.add_argument("--cpu-power-iters", type=int, default=100)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-csr-count", type=int, default=512,
                    help="--workload batch-csr: number of random sparse QPs in the batch")
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="--workload copy: bytes per copy (MiB) for calibrating the HBM byte counters")
    ap.add_argument("--qap-sizes", default="128,256,512,1024,2048,4096",
                    help="--workload qap-apply: QAP sizes n (X is n x n)")
    ap.add_argument("--workload", default="grid",
                    choices=["grid", "random", "lasso", "qap", "batch", "batch-csr", "copy", "qap-apply"],
                    help="grid (default): the headline single-instance workload; random / lasso / qap: the "
                         "other single-instance workloads (for profiling); batch: only the batched leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    from hprqp import DeviceModel, HPRQP_parameters

    if args.workload == "copy":
        # a plain device-to-device copy of known size: what FETCH_SIZE /
        # WRITE_SIZE read for it calibrates the counters of the other runs
        import torch

        nb = args.copy_mib << 20
        a = torch.empty(nb // 4, dtype=torch.float32, device="cuda:0").uniform_()
        b = torch.empty_like(a)
        for _ in range(args.warmup):
            b.copy_(a)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps

        # torch's own read-only (sum) and write-only (fill) kernels over the
        # same buffers beside it, for reference
        def rate(op):
            op()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                op()
            torch.cuda.synchronize()
            return nb / ((time.perf_counter() - t) / args.steps) / 1e9

        read_gbs = rate(lambda: a.sum())
        write_gbs = rate(lambda: b.fill_(1.0))
        print(json.dumps({"metric": "device copy GB/s (read + write)", "value": 2 * nb / dt / 1e9, "unit": "GB/s",
                          "bytes_read_per_copy": nb, "bytes_written_per_copy": nb, "steps": args.steps,
                          "torch_sum_read_gbs": read_gbs, "torch_fill_write_gbs": write_gbs}), flush=True)
        return
    if args.workload == "qap-apply":
        print(json.dumps(qap_apply_leg([int(v) for v in args.qap_sizes.split(",")], args.steps)), flush=True)
        return
    if args.workload == "batch-csr":
        print(json.dumps(batched_csr_leg(args.batch_csr_count, args.plain_sample)), flush=True)
        return
    if args.workload == "batch":
        out = batched_leg(args.batch_count, args.plain_sample)
        print(json.dumps({"metric": "batched solves/sec to 1e-6 rel KKT", "value":
                          out["batched_mixed_warm_solves_per_s"], "unit": "solves/s", "n_gpus": 1,
                          "higher_is_better": True, "data": "synthetic (hprqp.instances.portfolio_batch)",
                          "batched": out}), flush=True)
        return
    t0 = time.time()
    model, ptype, desc = workload(args.workload, args.grid)
    log(f"{desc}: n={model.n} m={model.m} built in {time.time() - t0:.1f}s")
    dm = DeviceModel(model, 0)
    params = HPRQP_parameters(verbose=False, precision=args.precision, time_kernels=True, problem_type=ptype,
                              spmv_lanes=args.spmv_lanes)
    for w in range(args.warmup):
        r = dm.solve(params, hist_cap=1000)
        log(f"warmup {w}: {r.status} iter={r.iter} kkt={r.residuals:.2e} time={r.time:.3f}s "
            f"power={r.power_time:.3f}s ({r.power_iter_A}+{r.power_iter_Q} it) setup={r.setup_time:.3f}s")
    results = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        r = dm.solve(params, hist_cap=1000)
        results.append(r)
        log(f"step {s}: {r.status} iter={r.iter} kkt={r.residuals:.2e} obj={r.primal_obj:.10e} "
            f"time={r.time:.3f}s")
    elapsed = time.perf_counter() - t_start
    solved = sum(1 for r in results if r.status == "OPTIMAL" and r.residuals < 1e-6)
    value = solved / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps
    r0 = results[-1]

    # in-solve average of the dominant kernel over the timed steps (HIP events)
    calls = sum(r.kernel_calls for r in results)
    kms = sum(r.kernel_ms * r.kernel_calls for r in results) / max(1, calls)
    prof = dm.profile(params, reps=args.profile_reps)
    log(f"in-solve z/x/w1 kernel: {kms * 1e3:.2f} us avg over {calls} launches; isolated kernel ms: "
        + ", ".join(f"{k}={v:.4f}" for k, v in prof.items()))
    if ptype == "QAP":
        nq = int(round(model.n ** 0.5))
        fl = 4.0 * nq ** 3  # two n^3 GEMM pairs per apply (diagonal S, T)
        peak = 78.6 if args.precision == "fp64" else 2500.0 / 3.0
        achieved = fl / (kms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "kernel": "the step's one operator apply: MFMA passes 1 and 2 of Q d, d = w - "
                    "x_hat, with the tempv epilogue and its tile sums (timed together)",
                    "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": None, "flops_per_launch": fl, "avg_ms": kms, "launches": calls,
                    "launch_sample": "every 8th partial-step launch of the timed solves carries the event pair",
                    "isolated_kernels_ms": prof}
    else:
        if ptype == "LASSO":
            A = model.Q.A
            nb = lasso_step_bytes(A.nnz, A.shape[0], A.shape[1], args.precision)
            kname = "LASSO step: k_csr (A w) + k_csr<EpiLasso> (A^T temp), timed together"
        else:
            dia_nd = q_diagonals(model.Q) if r0.lanes[0] == -2000 else None
            nb = zxw1_bytes(model.Q.nnz, model.n, args.precision, bound_streams(model), dia_nd)
            if dia_nd:
                kname = (f"z/x/w1 step: Q w in the diagonal form ({dia_nd} diagonals, k_dia) with the fused "
                         "EpiZxw1 epilogue (form autotuned)")
            else:
                kname = ("z/x/w1 step: CSR row product Q w with the fused EpiZxw1 epilogue (k_csr / k_csr_stream, "
                         "lanes autotuned)")
        achieved = nb / (kms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": kname,
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(args.workload, args.grid, args.precision),
                    "bytes_per_launch": nb, "avg_ms": kms, "launches": calls,
                    "launch_sample": "every 8th partial-step launch of the timed solves carries the event pair",
                    "isolated_kernels_ms": prof}

    # the same workload in the reduced-precision forms, with solves/s and the
    # error against the FP64 solves above: fp32 (FP32 matrices, FP32 offsets
    # of the iterates from the FP64 Halpern anchors, FP32 accumulation, FP64
    # anchor products, residuals and KKT) and mixed (FP32 matrices and data on
    # FP64 iterates with FP64 accumulation, FP64 polish near the tolerance)
    reduced = {}
    if args.precision == "fp64" and not args.no_mixed:
        for prec in ("fp32", "mixed"):
            mp = HPRQP_parameters(verbose=False, precision=prec, problem_type=ptype, time_kernels=True,
                                  spmv_lanes=args.spmv_lanes)
            dm.solve(mp, hist_cap=16)
            t = time.perf_counter()
            mr = [dm.solve(mp, hist_cap=16) for _ in range(max(1, args.steps))]
            mt = time.perf_counter() - t
            ok = sum(1 for r in mr if r.status == "OPTIMAL" and r.residuals < 1e-6)
            m0 = mr[-1]
            kc = sum(r.kernel_calls for r in mr)
            km = sum(r.kernel_ms * r.kernel_calls for r in mr) / max(1, kc)
            reduced[prec] = {"precision": prec, "solves_per_s": ok / mt, "vs_fp64": (ok / mt) / value,
                             "iterations": m0.iter, "iterations_vs_fp64": m0.iter / r0.iter,
                             "fp64_polish_iterations": m0.polish_iter, "kkt": m0.residuals,
                             "rel_obj_err_vs_fp64": abs(m0.primal_obj - r0.primal_obj) / (1.0 + abs(r0.primal_obj)),
                             "rel_x_err_vs_fp64": float(np.max(np.abs(m0.x - r0.x)) / (1.0 + np.max(np.abs(r0.x)))),
                             "zxw1_kernel_ms": km}
            if ptype == "QP":
                zb = zxw1_bytes(model.Q.nnz, model.n, prec, bound_streams(model),
                                q_diagonals(model.Q) if m0.lanes[0] == -2000 else None)
                reduced[prec]["zxw1_bytes_per_launch"] = zb
                reduced[prec]["zxw1_hbm_frac"] = zb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS if km > 0 else None
            log(f"{prec}: {ok}/{len(mr)} solved, {reduced[prec]['solves_per_s']:.3f} solves/s "
                f"({reduced[prec]['vs_fp64']:.2f}x fp64), iter={m0.iter} (polish {m0.polish_iter}), "
                f"rel obj err vs fp64 {reduced[prec]['rel_obj_err_vs_fp64']:.2e}, z/x/w1 {km * 1e3:.1f} us")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle

        ncores = 1
        t = time.time()
        ref = oracle.solve(model, problem_type=ptype, max_iter=args.cpu_iters, max_power_iter=args.cpu_power_iters)
        wall = time.time() - t
        t_pa = ref.power_time_A / max(1, ref.power_iter_A)
        t_pq = (ref.power_time - ref.power_time_A) / max(1, ref.power_iter_Q)
        t_it = (ref.time - ref.power_time) / max(1, ref.iter)
        est = ref.setup_time + t_pa * r0.power_iter_A + t_pq * r0.power_iter_Q + t_it * r0.iter
        cpu = {"value": 1.0 / est, "unit": "solves/s", "cores": ncores, "kind": "port",
               "sample": (f"oracle (C restatement of the reference CPU path, 1 thread) on the same workload: "
                          f"scaling + {ref.power_iter_A}+{ref.power_iter_Q} power iterations + {ref.iter} HPR "
                          f"iterations in {wall:.1f}s wall; extrapolated to the GPU run's "
                          f"{r0.power_iter_A}+{r0.power_iter_Q} power and {r0.iter} HPR iterations = "
                          f"{est:.1f}s per solve"),
               "sec_per_iter": t_it, "sec_per_power_iter_A": t_pa, "sec_per_power_iter_Q": t_pq,
               "setup_s": ref.setup_time, "nproc": os.cpu_count()}
        log(f"cpu baseline: {cpu['sample']}")

    line = {
        "metric": "solves/sec to 1e-6 rel KKT + SpMV HBM GB/s; mixed+batched+warm vs plain FP64",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": f"synthetic (seeded generator, hprqp.instances, workload {args.workload})",
        "config": {"workload": f"{desc}, {args.precision}, cold start, 1e-6 relative KKT",
                   "n": model.n, "m": model.m, "nnz_A": int(model.A.nnz),
                   "nnz_Q": int(model.Q.nnz) if hasattr(model.Q, "nnz") else None,
                   "precision": args.precision, "iterations": r0.iter, "kkt": r0.residuals,
                   "solve_time_s": r0.time, "power_time_s": r0.power_time, "setup_time_s": r0.setup_time,
                   "primal_obj": r0.primal_obj, "row_product_shapes_Q_A_AT": list(r0.lanes),
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32": reduced.get("fp32"),
        "mixed": reduced.get("mixed"),
        "batched": (None if args.no_batch or args.workload != "grid"
                    else batched_leg(args.batch_count, args.plain_sample)),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
