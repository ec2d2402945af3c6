#!/usr/bin/env python3
"""Offline autotuner of the persistent solver's launch geometry.

Times whole device solves (svmhip_solve, HIP-event kernel time plus the
Gram-tile time) over a sweep of
    grid sizes        workgroups of the persistent kernel (one owns a range
                      of 64-row slices), and
    exchange copies   copies of the per-iteration exchange records (more
                      copies spread the readers over more memory channels,
                      but cost the writer a store per word per copy)
for a set of problem shapes, cache regimes (the whole kernel matrix cached,
or an LRU cache of part of it) and both row modes (exact rows on demand,
and the Gram-tile mode on dense rows), and writes the fastest of each shape to
launch_table.txt, which svmhip_create reads (rows cover the problem sizes
between the measured ones, split at geometric midpoints).  Results do not
depend on either parameter, so the table only changes speed.

Run on the device, from the repository root:
    python libsvm-cuda_amd/tune/autotune.py [--quick] [--reps N] [--out PATH] [--log PATH]
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import libsvm_hip  # noqa: E402
from libsvm_hip import datasets, ffi  # noqa: E402

GRIDS = [32, 64, 128, 256]
# dense rows with the whole kernel matrix cached: also grids between the
# powers of two (how many 64-row slices each workgroup owns changes how
# evenly the row passes end: n=100k d=256 exact rows, 256 / 240 / 224 / 208 /
# 196 workgroups: 178.6 / 177.4 / 176.5 / 176.0 / 177.8 ms)
GRIDS_DENSE_FULL = [32, 64, 96, 128, 160, 192, 208, 224, 240, 256]
COPIES = [1, 2, 4]


def shapes(quick):
    """(kind, rows, features, cache bytes): full caches, and the 60k sparse
    configuration's 4 GB LRU cache."""
    dense_l = [10000, 30000, 100000] if quick else [10000, 20000, 50000, 100000]
    full = 64 * 2 ** 30
    out = [("dense", l, d, full) for d in (32, 256) for l in dense_l]
    out += [("sparse", l, 784, full) for l in ([20000, 60000] if not quick else [60000])]
    out += [("sparse", 60000, 784, 4 * 2 ** 30)]
    return out


def problem(kind, l, d):
    if kind == "dense":
        X, y = datasets.gmm_dense_matrix(n=l, d=d, seed=1003)
        return ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True), 1.0 / d, 1.0
    ip, ix, v, y, dd = datasets.blob_mask_csr(n=l)
    return ffi.Problem(ip, ix, v, y, dd), 0.02, 10.0


def time_solve(prob, prm, grid, copies, mfma, reps, cache):
    ys, p, a0, Cp, Cn, kind = libsvm_hip.formulate(prob, prm)
    st0 = libsvm_hip.initial_status(ys, a0, Cp, Cn)
    s = libsvm_hip.DeviceSolver(prob, prm, kind=kind, grid_blocks=grid, cache_bytes=cache, mfma_rows=mfma,
                                exchange_copies=copies)
    s.setup_rbf_variables()
    best, iters = float("inf"), 0
    for _ in range(reps + 1):  # the first solve warms up
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        st = s.solve()
        best = min(best, st["kernel_ms"] + st["gram_ms"])
        iters = st["iters"]
    s.close()
    return best, iters


def ranges(ls):
    """Row ranges around the measured sizes, split at geometric midpoints."""
    ls = sorted(ls)
    cuts = [int(math.sqrt(a * b)) for a, b in zip(ls, ls[1:])]
    lo = [1] + [c + 1 for c in cuts]
    hi = cuts + [2 ** 31 - 1]
    return list(zip(ls, lo, hi))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    # timed solves per point, the best kept (box-to-box and run-to-run noise
    # is ~1%: the committed table was written with --reps 3)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "launch_table.txt"))
    ap.add_argument("--log", default=None, help="also write every timed point to this JSON file")
    args = ap.parse_args()
    libsvm_hip.set_quiet(True)
    best = {}  # (mode, full, kind, d) -> {l: (grid, copies, ms)}
    log = []
    cached = {}
    for kind, l, d, cache in shapes(args.quick):
        if (kind, l, d) not in cached:
            cached = {(kind, l, d): problem(kind, l, d)}
        prob, gamma, C = cached[(kind, l, d)]
        prm = ffi.make_param(gamma=gamma, C_=C, eps=1e-3, shrinking=0)
        full = 1 if cache >= 4.0 * l * ((l + 63) // 64 * 64) else 0
        grids, copies = (GRIDS, COPIES) if full else ([128, 256], [2, 4])  # LRU solves are long: fewer points
        if full and kind == "dense":
            grids = GRIDS_DENSE_FULL
        for mode in (("exact", "gram") if kind == "dense" and full else ("exact",)):
            results = []
            for g in grids:
                if g > (l + 63) // 64:
                    continue
                for c in copies:
                    ms, iters = time_solve(prob, prm, g, c, 1 if mode == "gram" else 0, args.reps, cache)
                    results.append((ms, g, c))
                    log.append(dict(mode=mode, full=full, kind=kind, l=l, d=d, grid=g, copies=c, ms=ms, iters=iters))
                    print(f"[tune] {mode} full={full} {kind} l={l} d={d} grid={g} copies={c}: {ms:.1f} ms "
                          f"({iters} iters)", flush=True)
            ms, g, c = min(results)
            best.setdefault((mode, full, kind, d), {})[l] = (g, c, ms)
            print(f"[tune] best {mode} full={full} {kind} l={l} d={d}: grid {g}, copies {c}, {ms:.1f} ms", flush=True)
    lines = ["# Launch geometry of the persistent SMO kernel, written by tune/autotune.py on "
             f"{time.strftime('%Y-%m-%d')} from timed whole solves on one MI355X.",
             "# mode full dense l_min l_max d_min d_max grid xrep   (first matching row wins; full = 1: the",
             "# whole kernel matrix fits the row cache)"]
    for (mode, full, kind, d), per_l in sorted(best.items()):
        dense = 1 if kind == "dense" else 0
        if kind == "dense":
            dmin, dmax = (1, 64) if d == 32 else (65, 2 ** 31 - 1)
        else:
            dmin, dmax = 1, 2 ** 31 - 1
        for l, lo, hi in ranges(per_l.keys()):
            g, c, ms = per_l[l]
            lines.append(f"{mode} {full} {dense} {lo} {hi} {dmin} {dmax} {g} {c}   # measured l={l} d={d}: {ms:.1f} ms")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    if args.log:  # every timed point, as JSON
        with open(args.log, "w") as f:
            json.dump(log, f, indent=1)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
