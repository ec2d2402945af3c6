#!/usr/bin/env python3
"""Benchmark of the MI355X SMO solver.

Workload (the metric's "n=100k RBF" configuration): C-SVC, RBF gamma=1/256,
C=1, eps=1e-3 on gmm_dense n=100,000 x d=256 float32 rows (8-component
Gaussian mixture, label = component parity, seed 1003).  One step = one full
training solve of that problem on the device with the rows already resident
in HBM: solver state reset (labels, alpha, G = p, empty kernel-row cache),
the whole SMO loop, and rho / obj on the device.  value = SMO iterations per
second over the timed steps.

Also reported on the same JSON line:
  roofline     the persistent SMO kernel (k_smo): algorithmic bytes per launch
               (iters * 44*l + rows_computed * (4*l*d + 8*l)) / launch time
               measured with HIP events on the solver's stream, vs 8 TB/s
  cpu_baseline the FP64 CPU oracle (LIBSVM 3.18 restatement) on one host core,
               on the first iterations of the same solve (bounded sample)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "libsvm-cuda_amd"))

import numpy as np  # noqa: E402

import libsvm_hip  # noqa: E402
from libsvm_hip import datasets, ffi  # noqa: E402

METRIC = "SMO iters/s & train time (n=100k RBF) + p50/p99 solve latency at fixed QPS"
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(prob, prm, seconds):
    """Oracle on one core over the first iterations of the same solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline leg only
    p = ffi.make_param(gamma=prm.gamma, C_=prm.C, eps=prm.eps, cache_size=100.0, shrinking=1,
                       cuda_flag=0)
    iters = 50
    t0 = time.perf_counter()
    oracle.train_one(prob, p, max_iter=iters)
    dt = time.perf_counter() - t0
    if dt < seconds / 4:
        iters = max(iters, int(iters * seconds / max(dt, 1e-3)))
        t0 = time.perf_counter()
        oracle.train_one(prob, p, max_iter=iters)
        dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "SMO iters/s", "cores": 1, "kind": "port",
            "sample": f"first {iters} iterations of the same solve (LIBSVM 3.18 restatement, "
                      f"-m 100 MB row cache, shrinking on), {dt:.1f} s"}


def copy_probe(gib):
    """A plain device-to-device copy of known size through torch's copy
    kernel: its FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    import torch
    n = int(gib * 2 ** 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: {4 * n} bytes read, {4 * n} bytes written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--d", type=int, default=256)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--cache-gb", type=float, default=64.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="instead of the solve, run a plain 1-pass device copy of this many GiB "
                         "(reference for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_probe_gib > 0:
        copy_probe(args.copy_probe_gib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this solver build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    libsvm_hip.set_quiet(True)
    n, d = args.n, args.d
    t0 = time.perf_counter()
    X, y = datasets.gmm_dense_matrix(n=n, d=d, seed=1003)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
    log(f"[bench] data n={n} d={d} built in {time.perf_counter() - t0:.1f}s")
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, cache_size=100.0, shrinking=0)
    ys, p, a0, Cp, Cn, kind = libsvm_hip.formulate(prob, prm)
    st0 = libsvm_hip.initial_status(ys, a0, Cp, Cn)
    s = libsvm_hip.DeviceSolver(prob, prm, kind=kind, grid_blocks=args.grid,
                                cache_bytes=args.cache_gb * 2 ** 30)
    s.setup_rbf_variables()
    log("[bench] problem resident on the device")

    def one_step():
        t = time.perf_counter()
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        st = s.solve()
        return time.perf_counter() - t, st

    for w in range(args.warmup):
        dt, st = one_step()
        log(f"[bench] warmup {w}: {st['iters']} iters in {dt:.2f}s")
    walls, stats = [], []
    for k in range(args.steps):
        dt, st = one_step()
        walls.append(dt)
        stats.append(st)
        log(f"[bench] step {k}: {st['iters']} iters, {dt:.3f}s wall, kernel {st['kernel_ms']:.1f} ms, "
            f"rows computed {st['row_misses']}, obj {st['obj']:.10g}, rho {st['rho']:.10g}")
    iters = sum(st["iters"] for st in stats)
    wall = sum(walls)
    kern_s = sum(st["kernel_ms"] for st in stats) / 1e3
    launches = sum(st["launches"] for st in stats)
    misses = sum(st["row_misses"] for st in stats)
    b_iter = 44.0 * n
    b_row = 4.0 * n * d + 8.0 * n
    alg_bytes = (iters * b_iter + misses * b_row) / launches
    launch_s = kern_s / launches
    achieved = alg_bytes / launch_s
    last = stats[-1]
    out = {
        "metric": METRIC,
        "value": iters / wall,
        "unit": "SMO iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gmm_dense n={n} d={d} C-SVC RBF gamma=1/{d} C=1 eps=1e-3",
                   "train_time_s": wall / args.steps, "iters_per_solve": last["iters"],
                   "rows_computed_per_solve": last["row_misses"], "cache_rows": last["cache_slots"],
                   "grid_blocks": last["grid_blocks"], "obj": last["obj"], "rho": last["rho"]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_smo", "avg_launch_ms": 1e3 * launch_s,
                     "bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, prm, args.cpu_seconds)
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
