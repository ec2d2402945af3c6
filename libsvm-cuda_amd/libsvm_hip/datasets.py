"""Deterministic synthetic training sets for the benchmark workloads.

Every generator uses numpy's PCG64 seeded explicitly and rounds all feature
values to float32 (then stores them as float64), so the device solver, which
keeps features in float32, and the FP64 CPU oracle see identical inputs.
Each function returns (indptr, indices, data, y, n_features) in CSR form
with 0-based column indices.

Workloads (sizes are the defaults):
  zipf_binary      n=2,000   d=123  sparse binary features, Zipf(1.1) activity
  blob_mask_csr    n=60,000  d=784  ~19% dense CSR in a 28x28 centred disc
  gmm_dense        n=100,000 d=256  dense, 8-component Gaussian mixture
  clusters_10class ~40k rows d=128  10 Gaussian clusters, log-uniform sizes
  serving_request  one request of the mixed serving stream
"""
from __future__ import annotations

import numpy as np


def _csr_from_mask(mask, values):
    """CSR arrays from a boolean activity mask and a dense value matrix."""
    rows, cols = np.nonzero(mask)
    indptr = np.zeros(mask.shape[0] + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=mask.shape[0]), out=indptr[1:])
    data = values[rows, cols].astype(np.float32).astype(np.float64)
    return indptr, cols.astype(np.int32), data


def zipf_binary(n=2000, d=123, nnz_per_row=14.0, pos_frac=0.24, flips=None, seed=1001):
    """Binary sparse features; feature j (1-based) is active with probability
    min(1, c j^-1.1), c set so the expected row length is nnz_per_row.
    Labels sign(w.x + b), b at the (1-pos_frac) quantile, then `flips`
    labels (default 5% of n) flipped, chosen without replacement."""
    rng = np.random.Generator(np.random.PCG64(seed))
    j = np.arange(1, d + 1, dtype=np.float64)
    base = j ** -1.1
    lo, hi = 0.0, 1e6
    for _ in range(200):  # bisection for c with sum min(1, c*base) = nnz_per_row
        c = 0.5 * (lo + hi)
        if np.minimum(1.0, c * base).sum() < nnz_per_row:
            lo = c
        else:
            hi = c
    prob = np.minimum(1.0, c * base)
    mask = rng.random((n, d)) < prob[None, :]
    vals = np.ones((n, d))
    w = rng.standard_normal(d)
    score = (mask * w[None, :]).sum(axis=1)
    b = -np.quantile(score, 1.0 - pos_frac)
    y = np.where(score + b > 0, 1.0, -1.0)
    nflip = int(round(0.05 * n)) if flips is None else flips
    flip = rng.choice(n, size=nflip, replace=False)
    y[flip] = -y[flip]
    indptr, indices, data = _csr_from_mask(mask, vals)
    return indptr, indices, data, y, d


def blob_mask_csr(n=60000, side=28, radius=10.5, active=149.0, centres=5, sigma=0.3, seed=1002):
    """Non-negative CSR rows on a side x side grid: each pixel inside the
    centred disc of the given radius is active with probability
    active/|disc|, values U(0,1].  Labels: sign of the log-likelihood ratio
    of two isotropic Gaussian mixtures (`centres` centres each, U(0,1) on the
    disc, 0 elsewhere, common sigma)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = side * side
    r, c = np.divmod(np.arange(d), side)
    disc = (r - (side - 1) / 2) ** 2 + (c - (side - 1) / 2) ** 2 <= radius ** 2
    pd = active / disc.sum()
    mus = np.zeros((2, centres, d))
    mus[:, :, disc] = rng.random((2, centres, int(disc.sum())))
    chunks_ip, chunks_ix, chunks_v, ys = [], [], [], []
    step = 10000
    offset = 0
    for s in range(0, n, step):
        m = min(step, n - s)
        mask = (rng.random((m, d)) < pd) & disc[None, :]
        vals = (1.0 - rng.random((m, d))).astype(np.float32).astype(np.float64)
        X = np.where(mask, vals, 0.0)
        llr = []
        for cls in range(2):
            d2 = ((X[:, None, :] - mus[cls][None, :, :]) ** 2).sum(axis=2)
            z = -d2 / (2 * sigma * sigma)
            zm = z.max(axis=1, keepdims=True)
            llr.append((zm[:, 0] + np.log(np.exp(z - zm).sum(axis=1))))
        ys.append(np.where(llr[0] - llr[1] > 0, 1.0, -1.0))
        ip, ix, v = _csr_from_mask(mask, vals)
        chunks_ip.append(ip[1:] + offset)
        offset += len(ix)
        chunks_ix.append(ix)
        chunks_v.append(v)
    indptr = np.concatenate([[0], *chunks_ip]).astype(np.int64)
    return indptr, np.concatenate(chunks_ix), np.concatenate(chunks_v), np.concatenate(ys), d


def gmm_dense(n=100000, d=256, components=8, seed=1003):
    """Dense rows from an equal-weight Gaussian mixture: means N(0, 4I),
    per-component diagonal sigma U(0.5, 2); label +1 for even component
    ids, -1 for odd.  Returned in CSR form with every entry stored."""
    X, y = gmm_dense_matrix(n, d, components, seed)
    indptr = np.arange(0, (n + 1) * d, d, dtype=np.int64)
    indices = np.tile(np.arange(d, dtype=np.int32), n)
    return indptr, indices, X.ravel().astype(np.float64), y, d


def gmm_dense_matrix(n=100000, d=256, components=8, seed=1003):
    rng = np.random.Generator(np.random.PCG64(seed))
    mu = 2.0 * rng.standard_normal((components, d))
    sig = rng.uniform(0.5, 2.0, size=(components, d))
    comp = rng.integers(0, components, size=n)
    X = (mu[comp] + sig[comp] * rng.standard_normal((n, d))).astype(np.float32)
    y = np.where(comp % 2 == 0, 1.0, -1.0)
    return X, y


def clusters_10class(classes=10, d=128, lo=500, hi=12000, total_range=(36000, 44000),
                     folds=5, seed=1004, scale=1.0):
    """One Gaussian cluster per class (means N(0, 1.5 I), unit noise), class
    sizes round(exp(U(ln lo, ln hi))) redrawn until the total falls in
    total_range; `scale` < 1 shrinks every class for quick tests.
    Returns (X float32 [n, d], y int labels 1..classes, fold ids)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    while True:
        sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), size=classes))).astype(int)
        if total_range[0] <= sizes.sum() <= total_range[1]:
            break
    sizes = np.maximum(4, np.round(sizes * scale).astype(int))
    mu = np.sqrt(1.5) * rng.standard_normal((classes, d))
    y = np.repeat(np.arange(1, classes + 1), sizes)
    X = (mu[y - 1] + rng.standard_normal((len(y), d))).astype(np.float32)
    perm = rng.permutation(len(y))
    X, y = X[perm], y[perm]
    fold = np.empty(len(y), dtype=np.int32)
    for c in range(1, classes + 1):  # stratified: deal each class round-robin
        idx = np.nonzero(y == c)[0]
        fold[idx[rng.permutation(len(idx))]] = np.arange(len(idx)) % folds
    return X, y, fold


def serving_request(k, base_seed=1005, n_range=(500, 20000)):
    """Request k of the serving stream: C-SVC (80%) or nu-SVC (nu = 0.3),
    n log-uniform in n_range, d in {32, 128, 512}, CSR density in
    {0.05, 0.3, 1.0}, U(0,1) features, labels from a sparse random
    hyperplane (10% support, at least one feature) split exactly at the
    median (ties broken at random, so both classes hold half the rows), then
    3% flips."""
    rng = np.random.Generator(np.random.PCG64(base_seed * 10 ** 6 + k))
    kind = "nu_svc" if rng.random() < 0.2 else "c_svc"
    n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
    d = int(rng.choice([32, 128, 512]))
    dens = float(rng.choice([0.05, 0.3, 1.0]))
    C = float(rng.choice([0.5, 1.0, 8.0]))
    mask = rng.random((n, d)) < dens
    vals = rng.random((n, d)).astype(np.float32).astype(np.float64)
    X = np.where(mask, vals, 0.0)
    w = np.where(rng.random(d) < 0.1, rng.standard_normal(d), 0.0)
    if not w.any():
        w[rng.integers(d)] = rng.standard_normal()
    s = X @ w
    order = np.lexsort((rng.random(n), s))
    y = np.full(n, -1.0)
    y[order[n // 2:]] = 1.0
    flip = rng.random(n) < 0.03
    y[flip] = -y[flip]
    indptr, indices, data = _csr_from_mask(mask, vals)
    return dict(kind=kind, n=n, d=d, density=dens, C=C, gamma=1.0 / d, nu=0.3,
                indptr=indptr, indices=indices, data=data, y=y)


def serving_rows(seed, n, d, density):
    """n test rows shaped like a serving request's: CSR, U(0,1) at positions
    active with probability `density`.  Returns (indptr, indices, data)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mask = rng.random((n, d)) < density
    vals = rng.random((n, d)).astype(np.float32).astype(np.float64)
    return _csr_from_mask(mask, vals)
