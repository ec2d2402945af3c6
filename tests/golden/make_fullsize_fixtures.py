"""Writes the oracle's results on the full-size workloads as fixtures.

Runs the FP64 CPU oracle (oracle/, test infrastructure) in this container on
the benchmark configurations at their stated sizes and stores what the GPU
tests compare against, for both oracle variants: "float" caches Q rows as
float (LIBSVM's Qfloat, libsvm/svm.cpp:16, the reference-exact variant) and
"double" caches them in FP64 (the pure-FP64 variant).  shrinking=0, as the
device solver runs (the reference's device path requires it,
svm-train/svm-train.c:260-262); the cache size changes no iterate.

    python tests/golden/make_fullsize_fixtures.py NAME [--threads N] [--procs P]

NAME is one of
  gmm_dense_100k   n=100,000 x d=256 dense, C-SVC RBF gamma=1/256, C=1
  blob_csr_60k     n=60,000 x d=784 CSR (~19%), C-SVC RBF gamma=0.02, C=10
  ovo_cv_grid      10 classes (~38k x 128 dense), 5-fold CV x 4 gammas x 45
                   one-vs-one pairs = 900 C-SVC solves, C=4
  serving_slice    a fixed slice of the serving stream at its real sizes
                   (see SERVING_SLICE) plus one 10,000-row prediction

Single solves store, per variant: iters, obj, rho, nSV, the support-vector
indices and their alpha (raw solver variables).  The 900 solves store
per-solve iters, obj, rho, nSV and a 64-bit digest of the support-vector
index set.  Output: tests/golden/fullsize_NAME.npz.
"""
import argparse
import hashlib
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "libsvm-cuda_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

from libsvm_hip import datasets, ffi  # noqa: E402

# Serving requests (ids of datasets.serving_request) solved in the slice, and
# the request whose model predicts the 10,000 test rows.  Chosen to span the
# stream's sizes: the largest dense d=512 request of the first 400 ids, the
# largest request overall, nu-SVC requests and a few of the mid-size ones.
SERVING_SLICE_N = 400


def sv_digest(idx):
    return np.frombuffer(hashlib.sha256(np.asarray(idx, dtype=np.int32).tobytes()).digest()[:8], dtype=np.uint64)[0]


def single(prob, prm_kw, cache_mb, out, tag):
    import oracle
    for variant, qd in (("float", False), ("double", True)):
        prm = ffi.make_param(cache_size=cache_mb * (2 if qd else 1), shrinking=0, cuda_flag=0, **prm_kw)
        t = time.time()
        r = oracle.train_one(prob, prm, q_double=qd)
        a = r["alpha_raw"]
        sv = np.nonzero(a > 0)[0].astype(np.int32)
        out[f"{variant}_iters"] = r["iters"]
        out[f"{variant}_obj"] = r["obj"]
        out[f"{variant}_rho"] = r["rho"]
        out[f"{variant}_nSV"] = len(sv)
        out[f"{variant}_sv"] = sv
        out[f"{variant}_alpha_sv"] = a[sv]
        print(f"[{tag}] {variant}: iters={r['iters']} obj={r['obj']:.15g} rho={r['rho']:.15g} nSV={len(sv)} "
              f"{time.time() - t:.1f}s", flush=True)


def gmm_dense_100k(args, out):
    X, y = datasets.gmm_dense_matrix(n=100000, d=256, seed=1003)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
    single(prob, dict(gamma=1.0 / 256, C_=1.0, eps=1e-3), 16000.0, out, "gmm_dense_100k")


def blob_csr_60k(args, out):
    ip, ix, v, y, d = datasets.blob_mask_csr(n=60000)
    prob = ffi.Problem(ip, ix, v, y, d)
    single(prob, dict(gamma=0.02, C_=10.0, eps=1e-3), 15000.0, out, "blob_csr_60k")


def _ovo_worker(task):
    import oracle
    X, rows, yy, g, qd = task
    prob = ffi.Problem.from_dense(X[rows].astype(np.float64), yy.astype(np.float64), keep_zeros=True)
    prm = ffi.make_param(gamma=g, C_=4.0, eps=1e-3, cache_size=2000.0, shrinking=0, cuda_flag=0)
    r = oracle.train_one(prob, prm, q_double=qd)
    sv = np.nonzero(r["alpha_raw"] > 0)[0]
    return r["iters"], r["obj"], r["rho"], len(sv), sv_digest(sv)


def ovo_jobs():
    """The 900 subproblems in bench.py's order: for each fold, the pairs of
    classes in order of first appearance in the training subset, rows of
    class a (+1) then class b (-1), each at the 4 gammas."""
    X, y, fold = datasets.clusters_10class()
    gammas = [2.0 ** -9, 2.0 ** -7, 2.0 ** -5, 2.0 ** -3]
    jobs = []
    for f in range(int(fold.max()) + 1):
        train = np.nonzero(fold != f)[0]
        labels = list(dict.fromkeys(y[train].tolist()))
        by = {c: train[y[train] == c] for c in labels}
        for a in range(len(labels)):
            for b in range(a + 1, len(labels)):
                rows = np.concatenate([by[labels[a]], by[labels[b]]])
                yy = np.concatenate([np.ones(len(by[labels[a]])), -np.ones(len(by[labels[b]]))]).astype(np.int8)
                for g in gammas:
                    jobs.append((rows, yy, g))
    return X, jobs


def ovo_cv_grid(args, out):
    import multiprocessing as mp
    X, jobs = ovo_jobs()
    for variant, qd in (("float", False), ("double", True)):
        t = time.time()
        tasks = [(X, rows, yy, g, qd) for rows, yy, g in jobs]
        order = sorted(range(len(tasks)), key=lambda k: -len(jobs[k][0]))  # largest first
        with mp.get_context("fork").Pool(args.procs) as pool:
            res = pool.map(_ovo_worker, [tasks[k] for k in order], chunksize=1)
        back = [None] * len(tasks)
        for k, r in zip(order, res):
            back[k] = r
        out[f"{variant}_iters"] = np.array([r[0] for r in back], dtype=np.int64)
        out[f"{variant}_obj"] = np.array([r[1] for r in back])
        out[f"{variant}_rho"] = np.array([r[2] for r in back])
        out[f"{variant}_nSV"] = np.array([r[3] for r in back], dtype=np.int64)
        out[f"{variant}_sv_digest"] = np.array([r[4] for r in back], dtype=np.uint64)
        print(f"[ovo_cv_grid] {variant}: {len(back)} solves, {out[f'{variant}_iters'].sum()} iters, "
              f"{time.time() - t:.1f}s", flush=True)


def serving_slice_ids():
    """Deterministic choice of request ids from the first SERVING_SLICE_N."""
    info = []
    for k in range(SERVING_SLICE_N):
        rng = np.random.Generator(np.random.PCG64(1005 * 10 ** 6 + k))
        kind = "nu_svc" if rng.random() < 0.2 else "c_svc"
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(20000)))))
        d = int(rng.choice([32, 128, 512]))
        dens = float(rng.choice([0.05, 0.3, 1.0]))
        info.append((k, kind, n, d, dens))
    dense512 = max((r for r in info if r[3] == 512 and r[4] == 1.0), key=lambda r: r[2])
    largest = max(info, key=lambda r: r[2])
    nus = sorted((r for r in info if r[1] == "nu_svc"), key=lambda r: -r[2])[:2]
    mids = [r for r in info if 4000 <= r[2] <= 9000][:3]
    smalls = [r for r in info if r[2] < 1500][:2]
    ids = []
    for r in [dense512, largest] + nus + mids + smalls:
        if r[0] not in ids:
            ids.append(r[0])
    return ids, dense512[0]


def serving_slice(args, out):
    import oracle
    ids, pred_model = serving_slice_ids()
    out["ids"] = np.array(ids, dtype=np.int64)
    out["pred_model"] = pred_model
    for variant, qd in (("float", False), ("double", True)):
        for k in ids:
            r = datasets.serving_request(k)
            y1 = np.where(r["y"] == r["y"][0], 1.0, -1.0)
            prob = ffi.Problem(r["indptr"], r["indices"], r["data"], y1, r["d"])
            t = ffi.NU_SVC if r["kind"] == "nu_svc" else ffi.C_SVC
            prm = ffi.make_param(svm_type=t, gamma=r["gamma"], C_=r["C"], nu=r["nu"], eps=1e-3, shrinking=0,
                                 cache_size=4000.0 * (2 if qd else 1), cuda_flag=0)
            t0 = time.time()
            res = oracle.train_one(prob, prm, q_double=qd)
            sv = np.nonzero(res["alpha"] != 0)[0].astype(np.int32)
            out[f"{variant}_{k}_iters"] = res["iters"]
            out[f"{variant}_{k}_obj"] = res["obj"]
            out[f"{variant}_{k}_rho"] = res["rho"]
            out[f"{variant}_{k}_sv"] = sv
            out[f"{variant}_{k}_coef_sv"] = res["alpha"][sv]
            print(f"[serving_slice] {variant} request {k}: {r['kind']} n={r['n']} d={r['d']} dens={r['density']} "
                  f"C={r['C']} iters={res['iters']} nSV={len(sv)} {time.time() - t0:.1f}s", flush=True)
            if k == pred_model and variant == "float":
                # decision values of 10,000 test rows against this model, in
                # FP64 from the oracle's coefficients (svm_predict_values'
                # sum, libsvm/svm.cpp:2569-2643, RBF through k_function)
                ip, ix, v = datasets.serving_rows(1005 + r["d"], 10000, r["d"], 0.3)
                test = ffi.Problem(ip, ix, v, np.zeros(10000), r["d"])
                m = oracle.svm_train(prob, prm)
                _, dec = m.predict_values(test)
                out["pred_dec"] = dec[:, 0]
                out["pred_labels"] = np.array([r["y"][0], -r["y"][0]])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=["gmm_dense_100k", "blob_csr_60k", "ovo_cv_grid", "serving_slice"])
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--procs", type=int, default=8)
    args = ap.parse_args()
    import oracle
    oracle.set_threads(args.threads)
    out = {}
    t = time.time()
    globals()[args.name](args, out)
    path = os.path.join(HERE, f"fullsize_{args.name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes) in {time.time() - t:.0f}s", flush=True)


if __name__ == "__main__":
    main()
