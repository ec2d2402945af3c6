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
  cpu_baseline the FP64 CPU oracle (LIBSVM 3.18 restatement) on the host: one
               full training solve of the same problem (large row cache,
               missing kernel rows filled by all the cores it states)
  config.train_time_with_upload_s  svm_train's wall clock for the same solve
               from host rows: upload and SELL packing, cache allocation, the
               solve and the copy-out
  legs         one measurement of each other configuration in the same run:
               zipf_2000 (the CPU reference path's configuration, GPU and
               one-core oracle full solves), blob_csr_60k (the plain sparse
               configuration with a 4 GB row cache), ovo_cv_grid (900 solves of
               one-vs-one x 5-fold CV x 4 gammas in batched launches) and
               serving (a Poisson stream at 60 req/s: p50 / p99 latency,
               solves/s, results identical alone and in the mix)

Other workloads (--workload) run one configuration alone with its own
steps: blob_csr, ovo_cv_grid, serving (cpu_baseline_parallel: the oracle as
min(16, host CPUs) independent processes over the same solves) and predict
(10,000 rows against a serving-size model).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-legs]
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


def _noop(x):
    return x


def _oracle_solve(task):
    """One oracle solve in a worker process of the parallel CPU baseline:
    (row format, arrays, make_param keywords) -> SMO iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline leg only
    form, a, kw = task
    if form == "dense":
        prob = ffi.Problem.from_dense(a["X"].astype(np.float64), a["y"], keep_zeros=True)
    else:
        prob = ffi.Problem(a["indptr"], a["indices"], a["data"], a["y"], a["d"])
    return oracle.train_one(prob, ffi.make_param(**kw))["iters"]


def cpu_parallel(make_task, count, seconds):
    """The oracle as P = min(16, host CPUs) independent worker processes
    (LIBSVM has no threads; independent solves are how a CPU host runs such
    a workload in parallel).  Tasks 0, 1, ... are handed out in order, one
    per free worker, until `seconds` have passed; solves still running then
    are not counted.  Returns (P, solves done, their iterations, seconds)."""
    import multiprocessing as mp
    P = min(16, os.cpu_count() or 1)
    pool = mp.get_context("spawn").Pool(P)  # spawn: the parent has a device context
    try:
        pool.map(_noop, range(4 * P))  # workers started before the clock starts
        t0 = time.perf_counter()
        pending, nxt, done, iters = [], 0, 0, 0
        while time.perf_counter() - t0 < seconds:
            while len(pending) < P and nxt < count:
                pending.append(pool.apply_async(_oracle_solve, (make_task(nxt),)))
                nxt += 1
            if not pending:
                break
            still = []
            for r in pending:
                if r.ready():
                    iters += r.get()
                    done += 1
                else:
                    still.append(r)
            if len(still) == len(pending):
                time.sleep(0.01)
            pending = still
        dt = time.perf_counter() - t0
    finally:
        pool.terminate()
        pool.join()
    return P, done, iters, dt


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline legs only
    return oracle


def cpu_threads():
    """Host threads for the CPU baseline: the box's CPU share (16 at most)."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_full_solve(prob, prm, cache_mb, threads, shrinking=0):
    """One full oracle training solve (LIBSVM 3.18 restatement): wall seconds
    and the result.  threads > 1 fills each missing kernel row with that many
    threads (the elements are computed as before, so the result is the same)."""
    oracle = _oracle()
    p = ffi.make_param(svm_type=prm.svm_type, gamma=prm.gamma, C_=prm.C, nu=prm.nu, eps=prm.eps,
                       cache_size=cache_mb, shrinking=shrinking, cuda_flag=0)
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    r = oracle.train_one(prob, p)
    dt = time.perf_counter() - t0
    oracle.set_threads(1)
    return dt, r


def cpu_baseline(prob, prm):
    """The oracle's full training solve of the benchmark problem on the host:
    a 16 GB row cache (every row it computes stays cached, as on the device)
    and the box's cores filling missing rows."""
    P = cpu_threads()
    dt, r = cpu_full_solve(prob, prm, 16000.0, P)
    return {"value": r["iters"] / dt, "unit": "SMO iters/s", "cores": P, "kind": "port",
            "train_time_s": dt, "iters": r["iters"], "obj": r["obj"], "rho": r["rho"],
            "sample": f"one full training solve of the same problem ({r['iters']} iterations, {dt:.1f} s): "
                      f"LIBSVM 3.18 restatement, 16 GB row cache (labelled large-cache run), shrinking off, "
                      f"missing kernel rows filled by {P} threads"}


def measured_traffic(n, d, grid):
    """HBM bytes per k_smo launch from the calibrated FETCH_SIZE / WRITE_SIZE
    runs kept in profiles/ (valid for the configuration they were taken on:
    gmm_dense n=100000 d=256 with the default grid); None otherwise."""
    if (n, d, grid) != (100000, 256, 0):
        return None
    path = os.path.join(ROOT, "profiles", "k_smo_gmm_dense_100k_counters.md")
    try:
        for line in open(path):
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


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


def ovo_cv_grid_jobs(X, y, fold, gammas, C):
    """Subproblems of 5-fold CV x gamma grid x one-vs-one pairs, formed as
    svm_cross_validation / svm_train form them (classes in order of first
    appearance in each training subset)."""
    jobs = []
    nf = int(fold.max()) + 1
    for f in range(nf):
        train = np.nonzero(fold != f)[0]
        labels = list(dict.fromkeys(y[train].tolist()))
        by = {c: train[y[train] == c] for c in labels}
        for a in range(len(labels)):
            for b in range(a + 1, len(labels)):
                rows = np.concatenate([by[labels[a]], by[labels[b]]])
                yy = np.concatenate([np.ones(len(by[labels[a]])), -np.ones(len(by[labels[b]]))]).astype(np.int8)
                jobs.append((f, a, b, rows, yy))
    return jobs


def bench_ovo_cv_grid(args, leg=False):
    """10 classes, ~40k x 128 dense, 5-fold CV x 4 gammas x 45 pairs = 900
    solves per step through the batched solver (one workgroup or a team of
    workgroups per solve, all solves in flight together).  leg: one warm-up
    and one timed step, without the plain-path and CPU samples; returns the
    result instead of printing it."""
    steps, warmup = (1, 1) if leg else (args.steps, args.warmup)
    no_plain, no_cpu = (True, True) if leg else (args.no_plain, args.no_cpu)
    gammas = [2.0 ** -9, 2.0 ** -7, 2.0 ** -5, 2.0 ** -3]
    C = 4.0
    X, y, fold = datasets.clusters_10class()
    prob = ffi.Problem.from_dense(X.astype(np.float64), y.astype(np.float64), keep_zeros=True)
    prm = ffi.make_param(gamma=gammas[0], C_=C, eps=1e-3, shrinking=0)
    jobs = ovo_cv_grid_jobs(X, y, fold, gammas, C)
    b = libsvm_hip.BatchSolver(prob, prm, cache_bytes=args.batch_cache_mb * 2 ** 20)
    t0 = time.perf_counter()
    rids = [b.add_rowset(rows) for (_, _, _, rows, _) in jobs]
    pack_s = time.perf_counter() - t0
    log(f"[bench] {len(jobs)} row sets packed in {pack_s:.1f}s (n={prob.l})")
    subs = []
    for rid, (_, _, _, rows, yy) in zip(rids, jobs):
        for g in gammas:
            n = len(rows)
            subs.append(dict(rowset=rid, kind=libsvm_hip.KIND_C, y=yy, p=-np.ones(n), alpha0=np.zeros(n),
                             Cp=C, Cn=C, gamma=g, eps=1e-3))
    first = {}  # the handle's first solve also builds the shared dot products
    for w in range(warmup):
        t = time.perf_counter()
        _, st = b.solve(subs)
        first = first or dict(wall_s=time.perf_counter() - t, shared_dots_ms=st["gram_ms"])
        log(f"[bench] warmup {w}: {st['iters']} iters, device {st['solve_ms']:.0f} ms")
    walls, stats = [], []
    for k in range(steps):
        t = time.perf_counter()
        res, st = b.solve(subs)
        walls.append(time.perf_counter() - t)
        first = first or dict(wall_s=walls[-1], shared_dots_ms=st["gram_ms"])
        stats.append(st)
        log(f"[bench] step {k}: {len(subs)} solves, {st['iters']} iters, {walls[-1]:.2f}s wall, "
            f"device {st['solve_ms']:.0f} ms in {st['launches']} launches, rows computed {st['row_misses']}")
    iters = sum(st["iters"] for st in stats)
    wall = sum(walls)
    # plain path (one persistent solve at a time, the whole grid per solve)
    # on fold 0: its 45 pairs at all 4 gammas, a fifth of the workload
    sample = [(k, g) for k, jb in enumerate(jobs) if jb[0] == 0 for g in gammas]
    p_iters, p_time = 0, 0.0
    if not no_plain:
        for n_done, (k, g) in enumerate(sample):
            _, _, _, rows, yy = jobs[k]
            sub = prob.subset(rows, y=yy.astype(np.float64))
            sp = ffi.make_param(gamma=g, C_=C, eps=1e-3, shrinking=0, cache_size=4096.0)
            t = time.perf_counter()
            r = libsvm_hip.solve_one(sub, sp)
            p_time += time.perf_counter() - t
            p_iters += r["iters"]
            if n_done % 45 == 44:
                log(f"[bench] plain path: {n_done + 1} solves, {p_iters} iters, {p_time:.1f}s")
        log(f"[bench] plain persistent path, {len(sample)} solves one at a time: {p_iters} iters in {p_time:.2f}s")
    last = stats[-1]
    out = {
        "metric": METRIC,
        "value": iters / wall,
        "unit": "SMO iters/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"clusters_10class n={prob.l} d=128: 5-fold CV x 4 gammas x 45 OvO pairs, "
                               f"C-SVC C=4 eps=1e-3, {len(subs)} solves per step in one batched launch",
                   "solves_per_s": len(subs) * steps / wall, "train_time_s": wall / steps,
                   "iters_per_step": last["iters"], "rows_computed_per_step": last["row_misses"],
                   "stages_per_step": last["launches"],
                   "handle_setup": {"rowsets_packed_s": pack_s, "shared_dot_products": bool(last["shared_dots"]),
                                    "shared_dots_ms": first.get("shared_dots_ms"),
                                    "first_solve_wall_s": first.get("wall_s")},
                   "plain_path_sample": (f"fold 0, {len(sample)} solves one at a time: {p_iters} iters "
                                         f"in {p_time:.2f}s") if p_time else None,
                   "plain_path_iters_per_s": (p_iters / p_time) if p_time else None,
                   "plain_path_train_time_s_extrapolated": (p_time * last["iters"] / p_iters) if p_time else None,
                   "speedup_vs_plain_path": ((iters / wall) / (p_iters / p_time)) if p_time else None},
    }
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline leg only
        t = time.perf_counter()
        c_iters = 0
        done = 0
        for k, g in sorted(sample, key=lambda kg: (len(jobs[kg[0]][3]), kg[1])):
            _, _, _, rows, yy = jobs[k]
            sub = prob.subset(rows, y=yy.astype(np.float64))
            sp = ffi.make_param(gamma=g, C_=C, eps=1e-3, cache_size=100.0, shrinking=1, cuda_flag=0)
            c_iters += oracle.train_one(sub, sp)["iters"]
            done += 1
            if time.perf_counter() - t > args.cpu_seconds:
                break
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": c_iters / dt, "unit": "SMO iters/s", "cores": 1, "kind": "port",
                               "sample": f"{done} fold-0 solves (smallest pairs first, all 4 gammas), "
                                         f"oracle with shrinking and a 100 MB cache, {dt:.1f}s"}
        order = sorted(sample, key=lambda kg: (len(jobs[kg[0]][3]), kg[1]))

        def ovo_task(i):
            k, g = order[i]
            _, _, _, rows, yy = jobs[k]
            return ("dense", {"X": X[rows], "y": yy.astype(np.float64)},
                    dict(gamma=g, C_=C, eps=1e-3, cache_size=100.0, shrinking=1, cuda_flag=0))

        P, pdone, piters, pdt = cpu_parallel(ovo_task, len(order), args.cpu_seconds)
        out["cpu_baseline_parallel"] = {
            "value": piters / pdt, "unit": "SMO iters/s", "cores": P, "kind": "port",
            "sample": f"{P} oracle processes over the same fold-0 solves in the same order, one solve "
                      f"per process at a time: {pdone} solves done in {pdt:.1f}s"}
    b.close()
    return out


def _gen_request(k):
    r = datasets.serving_request(k)
    return {key: r[key] for key in ("kind", "n", "d", "density", "C", "gamma", "nu", "indptr", "indices", "data", "y")}


def _serving_problem(r):
    return ffi.Problem(r["indptr"], r["indices"], r["data"], r["y"], r["d"])


def _serving_param(r):
    t = ffi.NU_SVC if r["kind"] == "nu_svc" else ffi.C_SVC
    return ffi.make_param(svm_type=t, gamma=r["gamma"], C_=r["C"], nu=r["nu"], eps=1e-3, shrinking=0)


def bench_serving(args, leg=False):
    """Poisson stream at --qps for --window seconds: 90% training requests
    (80% C-SVC / 20% nu-SVC, n log-uniform in [500, 20000], d in
    {32, 128, 512}, density in {0.05, 0.3, 1.0}), 10% predictions of 10,000
    rows against the newest resident model.  Latency = arrival to result on
    the host.  leg: 60 req/s for --leg-window seconds without the CPU
    samples; returns the result instead of printing it."""
    import multiprocessing as mp
    qps, window = (60.0, args.leg_window) if leg else (args.qps, args.window)
    no_cpu = True if leg else args.no_cpu
    rng = np.random.Generator(np.random.PCG64(1005))
    gaps = rng.exponential(1.0 / qps, size=int(qps * window * 2) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < window]
    is_pred = rng.random(len(arrivals)) < 0.10
    n_train = int((~is_pred).sum())
    t = time.perf_counter()
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        raw = pool.map(_gen_request, range(n_train), chunksize=4)
    log(f"[bench] {len(arrivals)} requests ({n_train} train) over {window:.0f}s at {qps} req/s; "
        f"data generated in {time.perf_counter() - t:.1f}s")
    probs = [_serving_problem(r) for r in raw]
    prms = [_serving_param(r) for r in raw]
    tests = {}
    for dd in (32, 128, 512):
        ip, ix, v = datasets.serving_rows(1005 + dd, 10000, dd, 0.3)
        tests[dd] = ffi.Problem(ip, ix, v, np.zeros(10000), dd)
    srv = libsvm_hip.Server()
    # warm-up: one small request end to end (module load, first launches)
    warm = srv.submit_train(probs[0], prms[0])
    srv.run_until_idle()
    srv.release(warm)
    lat_train, lat_pred, info, slow = [], [], {}, []
    models = []          # (id, d) of finished, still resident models
    pending_model = {}   # predict id -> model id
    ntr = 0
    nxt = 0
    t0 = time.perf_counter()
    last_log = t0
    done_train = done_pred = 0
    dev_ms = 0.0
    iters = 0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(arrivals) and arrivals[nxt] <= now:
            if is_pred[nxt] and models:
                mid, dd = models[-1]
                rid = srv.submit_predict(mid, tests[dd])
                pending_model[rid] = mid
                info[rid] = ("predict", arrivals[nxt])
            else:
                k = ntr % n_train
                rid = srv.submit_train(probs[k], prms[k])
                info[rid] = ("train", arrivals[nxt], k)
                ntr += 1
            nxt += 1
        a, q = srv.step(args.stage_ms)
        tnow = time.perf_counter() - t0
        for c in srv.poll():
            kind = info[c["id"]][0]
            lat = tnow - info[c["id"]][1]
            if c["status"] not in (0, 6):  # OK or SVMHIP_ERR_MAXITER
                raise RuntimeError(f"request {c['id']} failed with status {c['status']}")
            if kind == "train":
                lat_train.append(lat)
                rr = raw[info[c["id"]][2]]
                slow.append((lat, rr["kind"], rr["n"], rr["d"], rr["density"], rr["C"], c["iters"]))
                done_train += 1
                iters += c["iters"]
                dev_ms += c["device_ms"]
                models.append((c["id"], raw[info[c["id"]][2]]["d"]))
                # keep the newest 8 models resident
                while len(models) > 8:
                    old = models[0][0]
                    if old in pending_model.values():
                        break
                    srv.release(old)
                    models.pop(0)
            else:
                lat_pred.append(lat)
                done_pred += 1
                pending_model.pop(c["id"], None)
        if time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            log(f"[bench] t={tnow:.0f}s submitted {nxt} done {done_train}+{done_pred} active {a} queued {q}")
        if nxt >= len(arrivals) and a == 0 and q == 0:
            break
        if a == 0 and q == 0 and nxt < len(arrivals):
            time.sleep(max(0.0, min(0.005, arrivals[nxt] - (time.perf_counter() - t0))))
    makespan = time.perf_counter() - t0
    sst = srv.stats()
    for row in sorted(slow, reverse=True)[:5]:
        log("[bench] slow: %.2fs %s n=%d d=%d density=%.2f C=%g iters=%d" % row)
    lt = np.array(lat_train)
    lp = np.array(lat_pred) if lat_pred else np.zeros(1)
    # each result must be the same alone: re-run a sample of requests alone
    sample = [0, n_train // 3, (2 * n_train) // 3]
    alone_ok = True
    for k in sample:
        s1 = libsvm_hip.Server()
        a1 = s1.submit_train(probs[k], prms[k])
        r1 = s1.run_until_idle()[0]
        c1, _ = s1.train_result(a1, probs[k].l)
        s1.close()
        s2 = libsvm_hip.Server()
        ids2 = [s2.submit_train(probs[j], prms[j]) for j in range(k, min(k + 24, n_train))]
        r2 = {c["id"]: c for c in s2.run_until_idle()}[ids2[0]]
        c2, _ = s2.train_result(ids2[0], probs[k].l)
        s2.close()
        alone_ok &= bool(np.array_equal(c1, c2) and r1["iters"] == r2["iters"] and r1["rho"] == r2["rho"])
    out = {
        "metric": METRIC,
        "value": done_train / makespan,
        "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * makespan,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving stream: Poisson {qps} req/s for {window:.0f}s, 90% train "
                               f"(80% C-SVC / 20% nu-SVC, n in [500, 20000], d in {{32,128,512}}), "
                               f"10% predict of 10,000 rows against a resident model",
                   "train_requests": done_train, "predict_requests": done_pred,
                   "train_latency_p50_s": float(np.percentile(lt, 50)),
                   "train_latency_p99_s": float(np.percentile(lt, 99)),
                   "predict_latency_p50_s": float(np.percentile(lp, 50)),
                   "predict_latency_p99_s": float(np.percentile(lp, 99)),
                   "smo_iters_per_s": iters / makespan, "stage_ms": args.stage_ms,
                   "device_busy_frac": sst["solve_ms"] / 1e3 / makespan, "stages": sst["launches"],
                   "results_same_alone_and_in_mix": alone_ok},
    }
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline leg only
        t = time.perf_counter()
        done, c_it = 0, 0
        for k in range(n_train):
            y1 = np.where(probs[k].y == probs[k].y[0], 1.0, -1.0)
            pk = ffi.Problem(probs[k].indptr, probs[k].indices, probs[k].data, y1, probs[k].n_features)
            c_it += oracle.train_one(pk, prms[k])["iters"]
            done += 1
            if time.perf_counter() - t > args.cpu_seconds:
                break
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": done / dt, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"first {done} training requests of the stream, one after another, "
                                         f"{c_it} iterations in {dt:.1f}s"}

        def serve_task(k):
            r = raw[k]
            y1 = np.where(r["y"] == r["y"][0], 1.0, -1.0)
            t = ffi.NU_SVC if r["kind"] == "nu_svc" else ffi.C_SVC
            return ("csr", {"indptr": r["indptr"], "indices": r["indices"], "data": r["data"], "y": y1,
                            "d": r["d"]},
                    dict(svm_type=t, gamma=r["gamma"], C_=r["C"], nu=r["nu"], eps=1e-3, shrinking=0))

        P, pdone, piters, pdt = cpu_parallel(serve_task, n_train, args.cpu_seconds)
        out["cpu_baseline_parallel"] = {
            "value": pdone / pdt, "unit": "solves/s", "cores": P, "kind": "port",
            "sample": f"{P} oracle processes over the stream's training requests in arrival order, one "
                      f"request per process at a time: {pdone} done in {pdt:.1f}s ({piters} iterations)"}
    srv.close()
    return out


def bench_single(args, workload, leg=False):
    """One persistent-kernel solve per step (gmm_dense: the metric's n=100k
    configuration; blob_csr: 60,000 x 784 CSR with a 4 GB row cache).  leg:
    no warm-up, one timed step, no step-wise and CPU samples."""
    steps, warmup = (1, 0) if leg else (args.steps, args.warmup)
    t0 = time.perf_counter()
    if workload == "blob_csr":
        # 60,000 x 784 CSR rows (~19% nonzero), RBF gamma=0.02, C=10, a 4 GB
        # device row cache: the plain sparse-dot configuration
        n = args.n if (args.n != 100000 and not leg) else 60000
        ip, ix, v, y, d = datasets.blob_mask_csr(n=n)
        prob = ffi.Problem(ip, ix, v, y, d)
        nnz = int(ip[-1])
        prm = ffi.make_param(gamma=0.02, C_=10.0, eps=1e-3, cache_size=4096.0, shrinking=0)
        cache_gb = args.cache_gb if (args.cache_gb != 64.0 and not leg) else 4.0
        b_row = 8.0 * nnz + 12.0 * n
        wl = (f"blob_mask_csr n={n} d={d} nnz={nnz} C-SVC RBF gamma=0.02 C=10 eps=1e-3, "
              f"{cache_gb:g} GB row cache")
        row_bytes = "rows_computed * (8*nnz + 12*l)"
    else:
        n, d = args.n, args.d
        X, y = datasets.gmm_dense_matrix(n=n, d=d, seed=1003)
        prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
        prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, cache_size=100.0, shrinking=0)
        cache_gb = args.cache_gb
        b_row = 4.0 * n * d + 8.0 * n
        wl = f"gmm_dense n={n} d={d} C-SVC RBF gamma=1/{d} C=1 eps=1e-3"
        row_bytes = "rows_computed * (4*l*d + 8*l)"
    log(f"[bench] {workload}: data n={n} d={d} built in {time.perf_counter() - t0:.1f}s")
    ys, p, a0, Cp, Cn, kind = libsvm_hip.formulate(prob, prm)
    st0 = libsvm_hip.initial_status(ys, a0, Cp, Cn)
    mfma = 1 if (args.mfma and workload == "gmm_dense") else 0
    s = libsvm_hip.DeviceSolver(prob, prm, kind=kind, grid_blocks=args.grid,
                                cache_bytes=cache_gb * 2 ** 30, mfma_rows=mfma)
    s.setup_rbf_variables()
    log("[bench] problem resident on the device")

    def one_step():
        t = time.perf_counter()
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        st = s.solve()
        return time.perf_counter() - t, st

    for w in range(warmup):
        dt, st = one_step()
        log(f"[bench] warmup {w}: {st['iters']} iters in {dt:.3f}s")
    walls, stats = [], []
    for k in range(steps):
        dt, st = one_step()
        walls.append(dt)
        stats.append(st)
        log(f"[bench] step {k}: {st['iters']} iters, {dt:.3f}s wall, kernel {st['kernel_ms']:.1f} ms, "
            f"gram {st['gram_ms']:.1f} ms, rows computed {st['row_misses']}, obj {st['obj']:.12g}, "
            f"rho {st['rho']:.12g}, device allocations {st['device_allocs']}")
    _, alpha_dev, _ = s.fetch_vectors()  # the last timed solve's alpha
    iters = sum(st["iters"] for st in stats)
    wall = sum(walls)
    kern_s = sum(st["kernel_ms"] for st in stats) / 1e3
    launches = sum(st["launches"] for st in stats)
    misses = sum(st["row_misses"] for st in stats)
    b_iter = 44.0 * n
    alg_bytes = (iters * b_iter + misses * b_row) / launches
    launch_s = kern_s / launches
    achieved = alg_bytes / launch_s
    traffic = measured_traffic(n, d, args.grid) if workload == "gmm_dense" else None
    last = stats[-1]
    out = {
        "metric": METRIC,
        "value": iters / wall,
        "unit": "SMO iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl,
                   "train_time_s": wall / steps, "iters_per_solve": last["iters"],
                   "rows_computed_per_solve": last["row_misses"], "cache_rows": last["cache_slots"],
                   "grid_blocks": last["grid_blocks"], "obj": last["obj"], "rho": last["rho"],
                   "device_allocs_per_step_after_first": (stats[-1]["device_allocs"] - stats[0]["device_allocs"])
                   / max(1, steps - 1)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_smo", "avg_launch_ms": 1e3 * launch_s,
                     "bytes_per_launch": alg_bytes,
                     "bytes_formula": f"(iters * 44*l + {row_bytes}) / launches"},
    }
    if not args.no_plain and not leg:
        # the reference's shape of the loop on the same kernels: one host call
        # per phase (select, alpha, gradient, status), each a launch, with the
        # blocking copy of the stopping test, for a bounded number of
        # iterations from the same start
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        k_plain = min(args.plain_iters, last["iters"]) if args.plain_iters > 0 else last["iters"]
        t_start = time.perf_counter()
        done = 0
        for _ in range(k_plain):
            opt, _, _ = s.select_working_set()
            if opt:
                break
            s.compute_alpha()
            s.update_gradient()
            s.update_alpha_status()
            done += 1
        t_plain = time.perf_counter() - t_start
        # the fused kernel over exactly those iterations
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        t = time.perf_counter()
        sk = s.solve(max_iter=done)
        t_fused = time.perf_counter() - t
        out["config"]["stepwise_path_iters_per_s"] = done / t_plain
        whole = done == last["iters"]
        out["config"]["stepwise_path_sample"] = (f"{'the whole solve' if whole else f'first {done} iterations'} "
                                                 f"({done} iterations) through "
                                                 f"svmhip_select_working_set/compute_alpha/update_gradient/"
                                                 f"update_alpha_status, {t_plain:.2f}s; the persistent kernel "
                                                 f"over the same {sk['iters']} iterations: {t_fused:.3f}s")
        out["config"]["speedup_vs_stepwise_path"] = t_plain / t_fused
        if whole:
            out["config"]["stepwise_path_train_time_s"] = t_plain
    if workload == "gmm_dense" and not leg:
        out["config"]["error_vs_oracle"] = oracle_errors(last["obj"], last["rho"], alpha_dev, n, d)
    s.close()
    if workload == "gmm_dense" and not leg and not args.mfma:
        out["config"]["precision_variant"] = bench_precision_variant(args, prob, prm, ys, p, a0, st0, Cp, Cn, kind,
                                                                     cache_gb, steps, warmup)
        t_step = out["config"].get("stepwise_path_train_time_s")
        if t_step:
            pv = out["config"]["precision_variant"]
            pv["speedup_vs_stepwise_path"] = t_step / pv["train_time_s"]
    if not leg:
        # svm_train's own wall clock from the host rows (upload, SELL packing,
        # cache allocation, solve, copy-out, model), the row cache sized by
        # svm_parameter.cache_size
        prm_t = ffi.make_param(gamma=prm.gamma, C_=prm.C, eps=prm.eps, shrinking=0,
                               cache_size=cache_gb * 1024.0)
        t = time.perf_counter()
        m = libsvm_hip.svm_train(prob, prm_t)
        out["config"]["train_time_with_upload_s"] = time.perf_counter() - t
        del m
        log(f"[bench] svm_train from host rows: {out['config']['train_time_with_upload_s']:.3f}s")
    if not args.no_cpu and not leg:
        t = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(prob, prm)
        log(f"[bench] CPU oracle full solve: {time.perf_counter() - t:.1f}s")
    return out


def bench_predict(args):
    """Decision values of 10,000 test rows (d=512, CSR density 0.3, the
    serving stream's prediction shape) against the model of serving request
    333 (C-SVC, n=19,189 dense d=512 rows, C=8), through
    svmhip_predict_batch: one step = one call, host packing and upload
    included.  roofline: k_predict's FP64 kernel evaluations,
    3 * d flops per (test row, support vector) pair (subtract, multiply,
    add), against 78.6 TFLOP/s (FP64)."""
    r = datasets.serving_request(333)
    prob = _serving_problem(r)
    prm = _serving_param(r)
    m = libsvm_hip.svm_train(prob, prm)
    nsv = m.arrays()["l"]
    ip, ix, v = datasets.serving_rows(1005 + r["d"], 10000, r["d"], 0.3)
    test = ffi.Problem(ip, ix, v, np.zeros(10000), r["d"])
    for w in range(args.warmup):
        m.predict_batch(test)
    walls = []
    for k in range(args.steps):
        t = time.perf_counter()
        m.predict_batch(test)
        walls.append(time.perf_counter() - t)
        log(f"[bench] predict step {k}: {walls[-1] * 1e3:.1f} ms")
    wall = sum(walls) / len(walls)
    flops = 3.0 * r["d"] * nsv * 10000
    # the opt-in Gram-tile path (mfma_rows = 1) on the same call
    _, dec_e = m.predict_batch(test)
    _, dec_m = m.predict_batch(test, mfma=True)
    mwalls = []
    for k in range(args.steps):
        t = time.perf_counter()
        m.predict_batch(test, mfma=True)
        mwalls.append(time.perf_counter() - t)
        log(f"[bench] predict (Gram tiles) step {k}: {mwalls[-1] * 1e3:.1f} ms")
    mwall = sum(mwalls) / len(mwalls)
    variant = {"workload": "the same call with mfma_rows = 1: FP16 hi/lo split dots on the matrix cores "
                           "(FP32 accumulation), kernel values and sums in FP64",
               "value": 10000 / mwall, "unit": "decision values/s", "ms_per_step": 1e3 * mwall,
               "max_abs_err_vs_fp64": float(np.max(np.abs(dec_m - dec_e))),
               "decision_value_scale": float(np.max(np.abs(dec_e))),
               "mfma_flops_issued_per_call": 3 * 2.0 * r["d"] * nsv * 10000}
    return {"metric": METRIC, "value": 10000 / wall, "unit": "decision values/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"predict 10,000 rows d={r['d']} density 0.3 against {nsv} support vectors "
                                   f"(serving request 333: C-SVC n={r['n']} dense d={r['d']} C={r['C']})",
                       "nSV": nsv, "kernel_flops_per_call": flops,
                       "flops_formula": "3 * d * nSV * rows (RBF: subtract, multiply, add per feature)",
                       "gram_tile_variant": variant}}


MFMA_PEAK_F16 = 2.5e15  # MI355X dense FP16 matrix peak, flop/s


def oracle_errors(dev_obj, dev_rho, dev_alpha, n, d):
    """Errors of a device solve of the n=100k dense problem against the
    oracle's full-size results (tests/golden/fullsize_gmm_dense_100k.npz):
    objective (relative), rho (absolute), support-vector set difference."""
    path = os.path.join(ROOT, "tests", "golden", "fullsize_gmm_dense_100k.npz")
    if (n, d) != (100000, 256) or not os.path.exists(path):
        return None
    fx = np.load(path)
    sv = np.nonzero(dev_alpha > 0)[0]
    out = {}
    for v, name in (("float", "reference_exact"), ("double", "pure_fp64")):
        out[name] = {"obj_rel_err": abs(dev_obj - float(fx[f"{v}_obj"])) / abs(float(fx[f"{v}_obj"])),
                     "rho_abs_err": abs(dev_rho - float(fx[f"{v}_rho"])),
                     "sv_set_symmetric_difference": int(len(np.setxor1d(sv, fx[f"{v}_sv"]))),
                     "oracle_nSV": int(fx[f"{v}_nSV"]), "oracle_iters": int(fx[f"{v}_iters"])}
    return out


def bench_precision_variant(args, prob, prm, ys, p, a0, st0, Cp, Cn, kind, cache_gb, steps, warmup):
    """The configuration's precision variant on the same problem: the whole
    kernel matrix as FP16-split MFMA Gram tiles (mfma_rows = 1), then the
    SMO loop on cache hits; G, alpha, the gain, rho and the objective stay
    FP64.  Same steps as the headline, same problem resident in HBM."""
    s = libsvm_hip.DeviceSolver(prob, prm, kind=kind, grid_blocks=args.grid, cache_bytes=cache_gb * 2 ** 30,
                                mfma_rows=1)
    s.setup_rbf_variables()
    for w in range(warmup):
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        s.solve()
    walls, stats = [], []
    for k in range(steps):
        t = time.perf_counter()
        s.setup_solver(ys, p, a0, st0, Cp, Cn)
        st = s.solve()
        walls.append(time.perf_counter() - t)
        stats.append(st)
        log(f"[bench] precision variant step {k}: {st['iters']} iters, {walls[-1]:.3f}s wall, gram {st['gram_ms']:.1f} "
            f"ms, loop {st['kernel_ms']:.1f} ms, grid {st['grid_blocks']}")
    _, alpha, _ = s.fetch_vectors()
    s.close()
    last = stats[-1]
    n, d = prob.l, prob.n_features
    wall = sum(walls)
    gram_s = sum(x["gram_ms"] for x in stats) / 1e3 / steps
    loop_s = sum(x["kernel_ms"] for x in stats) / 1e3 / steps
    flops = 2.0 * n * n * d
    loop_bytes = last["iters"] * 44.0 * n
    return {"workload": f"gmm_dense n={n} d={d}, kernel matrix as FP16 hi/lo-split MFMA Gram tiles (three "
                        f"products, FP32 accumulation), SMO loop in FP64 on cache hits",
            "value": sum(x["iters"] for x in stats) / wall, "unit": "SMO iters/s", "ms_per_step": 1e3 * wall / steps,
            "train_time_s": wall / steps, "iters_per_solve": last["iters"], "obj": last["obj"], "rho": last["rho"],
            "grid_blocks": last["grid_blocks"], "gram_ms": 1e3 * gram_s, "k_smo_ms": 1e3 * loop_s,
            "dtype": "f16x3 MFMA / f32 accumulation (kernel rows), f64 (G, alpha, gain, rho)",
            "error_vs_oracle": oracle_errors(last["obj"], last["rho"], alpha, n, d),
            "roofline_gram": {"bound": "mfma", "kernel": "k_gram_f16", "achieved": flops / gram_s / 1e12,
                              "peak": MFMA_PEAK_F16 / 1e12, "unit": "TFLOP/s", "frac": flops / gram_s / MFMA_PEAK_F16,
                              "flops_formula": "2 * l^2 * d (every kernel value's dot product; the hi/lo products "
                                               "and the symmetric half are not counted)"},
            "roofline_loop": {"bound": "hbm", "kernel": "k_smo", "achieved": loop_bytes / loop_s / 1e9,
                              "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": loop_bytes / loop_s / HBM_PEAK,
                              "bytes_formula": "iters * 44*l (cache hits only)"}}


def leg_zipf(args):
    """The CPU reference path's configuration (zipf_binary n=2000, d=123,
    RBF gamma=1/123, C=1, eps=1e-3, shrinking on, 100 MB cache): svm_train
    on the device from host rows, and the oracle's full solve on one core."""
    ip, ix, v, y, d = datasets.zipf_binary(n=2000)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / 123, C_=1.0, eps=1e-3, cache_size=100.0, shrinking=1)
    libsvm_hip.svm_train(prob, prm)  # first call: module load, first launches
    t = time.perf_counter()
    m = libsvm_hip.svm_train(prob, prm)
    gpu_s = time.perf_counter() - t
    nsv = m.arrays()["l"]
    dev = libsvm_hip.solve_one(prob, prm)
    out = {"workload": "zipf_binary n=2000 d=123 C-SVC RBF gamma=1/123 C=1 eps=1e-3, 100 MB cache",
           "gpu_svm_train_s": gpu_s, "gpu_iters": dev["iters"], "gpu_obj": dev["obj"], "gpu_rho": dev["rho"],
           "nSV": nsv}
    if not args.no_cpu:
        for shr in (1, 0):
            dt, r = cpu_full_solve(prob, prm, 100.0, 1, shrinking=shr)
            key = "cpu_shrinking" if shr else "cpu_no_shrinking"
            out[key] = {"train_time_s": dt, "iters": r["iters"], "obj": r["obj"], "rho": r["rho"], "cores": 1,
                        "obj_rel_err_vs_gpu": abs(r["obj"] - dev["obj"]) / abs(r["obj"]),
                        "rho_abs_err_vs_gpu": abs(r["rho"] - dev["rho"])}
    log(f"[bench] zipf_2000: {out}")
    return out


def leg_blob(args):
    o = bench_single(args, "blob_csr", leg=True)
    c, r = o["config"], o["roofline"]
    out = {"workload": c["workload"], "iters": c["iters_per_solve"], "train_time_s": c["train_time_s"],
           "iters_per_s": o["value"], "rows_computed": c["rows_computed_per_solve"], "obj": c["obj"],
           "rho": c["rho"], "k_smo_ms": r["avg_launch_ms"], "roofline_frac": r["frac"],
           "roofline_achieved_GBs": r["achieved"], "bytes_formula": r["bytes_formula"]}
    log(f"[bench] blob_csr_60k: {out}")
    return out


def leg_ovo(args):
    o = bench_ovo_cv_grid(args, leg=True)
    c = o["config"]
    out = {"workload": c["workload"], "solves_per_s": c["solves_per_s"], "train_time_s": c["train_time_s"],
           "iters_per_s": o["value"], "iters": c["iters_per_step"], "rows_computed": c["rows_computed_per_step"],
           "stages": c["stages_per_step"]}
    log(f"[bench] ovo_cv_grid: {out}")
    return out


def leg_serving(args):
    o = bench_serving(args, leg=True)
    out = dict(o["config"])
    out["solves_per_s"] = o["value"]
    log(f"[bench] serving: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--d", type=int, default=256)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--cache-gb", type=float, default=64.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="the n=100k solve only")
    ap.add_argument("--mfma", action="store_true",
                    help="gmm_dense in Gram-tile mode (whole kernel matrix on the matrix cores)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="instead of the solve, run a plain 1-pass device copy of this many GiB "
                         "(reference for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--workload", default="gmm_dense",
                    choices=["gmm_dense", "blob_csr", "ovo_cv_grid", "serving", "predict"])
    ap.add_argument("--qps", type=float, default=20.0)
    ap.add_argument("--window", type=float, default=60.0)
    ap.add_argument("--leg-window", type=float, default=15.0, help="window of the serving leg (60 req/s)")
    ap.add_argument("--stage-ms", type=float, default=10.0)
    ap.add_argument("--batch-cache-mb", type=float, default=400.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-iters", type=int, default=0,
                    help="iterations of the step-wise comparison (0 = the whole solve)")
    args = ap.parse_args()
    if args.copy_probe_gib > 0:
        copy_probe(args.copy_probe_gib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this solver build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    libsvm_hip.set_quiet(True)
    if args.workload == "ovo_cv_grid":
        print(json.dumps(bench_ovo_cv_grid(args)), flush=True)
        return
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)), flush=True)
        return
    if args.workload == "predict":
        print(json.dumps(bench_predict(args)), flush=True)
        return
    out = bench_single(args, args.workload)
    if args.workload == "gmm_dense" and not args.no_legs:
        out["legs"] = {"zipf_2000": leg_zipf(args), "blob_csr_60k": leg_blob(args),
                       "ovo_cv_grid": leg_ovo(args), "serving": leg_serving(args)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
