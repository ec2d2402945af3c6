"""Batched solves against the oracle: each subproblem of a batch, whether
one workgroup or a team of them solves it and however many launches it is
split over, must give the same iterates as the oracle on the same row subset
(identical iteration count, alpha to 1e-9, objective to 1e-9)."""
import ctypes as C

import numpy as np
import pytest

from libsvm_hip import datasets, ffi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("team_max,chunk", [(0, 0), (1, 0), (1, 53), (4, 37), (16, 0)])
def test_batch_subsets_match_oracle(hip, ora, team_max, chunk):
    ip, ix, v, y, d = datasets.zipf_binary(n=3000, seed=9)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, shrinking=0)
    rng = np.random.default_rng(0)
    print(f"  team_max={team_max} chunk={chunk}", flush=True)
    b = hip.BatchSolver(prob, prm, team_max=team_max, chunk=chunk)
    cases = []
    for k in range(12):
        n = int(rng.integers(200, 2500))
        rows = np.sort(rng.choice(prob.l, size=n, replace=False))
        rid = b.add_rowset(rows)
        sub = prob.subset(rows)
        svm_type = ffi.NU_SVC if k % 4 == 3 else ffi.C_SVC
        gamma = [1.0 / d, 0.02, 0.1][k % 3]
        sp = ffi.make_param(svm_type=svm_type, gamma=gamma, C_=[0.5, 1.0, 4.0][k % 3], nu=0.3,
                            eps=1e-3, shrinking=0)
        yy, p, a0, Cp, Cn, kind = hip.formulate(sub, sp)
        cases.append((sub, sp, yy))
        cases[-1] = cases[-1] + (dict(rowset=rid, kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn,
                                      gamma=gamma, eps=1e-3),)
    res, st = b.solve([c[3] for c in cases])
    for (sub, sp, yy, _), r in zip(cases, res):
        ref = ora.train_one(sub, sp)
        coef, rho, obj = hip.finish_solution(sub.y, sp, r, yy)
        print(f"  l={sub.l} type={sp.svm_type} iters {r['iters']} vs {ref['iters']} obj {obj:.10g} vs {ref['obj']:.10g}",
              flush=True)
        assert r["status"] == 0
        assert r["iters"] == ref["iters"]
        np.testing.assert_allclose(coef, ref["alpha"], rtol=0, atol=1e-9)
        assert abs(obj - ref["obj"]) <= 1e-9 * abs(ref["obj"])
        assert abs(rho - ref["rho"]) <= 1e-9
    assert st["iters"] == sum(r["iters"] for r in res)
    if chunk:
        assert st["launches"] > 1


@pytest.mark.parametrize("team_max", [1, 0])
def test_batch_small_arena_lru(hip, ora, team_max):
    """Arenas too small for a full cache: LRU eviction inside each workgroup
    (team_max 1) or across a team's shared arena (team_max 0)."""
    ip, ix, v, y, d = datasets.zipf_binary(n=2000, seed=2)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, shrinking=0)
    rows = np.arange(0, 2000, 2)
    b = hip.BatchSolver(prob, prm, cache_bytes=4 * 1024 * 5, team_max=team_max)  # 5 rows of 1024 floats
    rid = b.add_rowset(rows)
    sub = prob.subset(rows)
    yy, p, a0, Cp, Cn, kind = hip.formulate(sub, prm)
    res, st = b.solve([dict(rowset=rid, kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn, gamma=prm.gamma, eps=1e-3)] * 3)
    ref = ora.train_one(sub, prm)
    for r in res:
        assert r["iters"] == ref["iters"]
        np.testing.assert_allclose(r["alpha_raw"], ref["alpha_raw"], rtol=0, atol=1e-12)
    assert st["row_misses"] > 3 * 100


def test_svm_cross_validation_matches_oracle(hip, ora):
    """svm_cross_validation trains every fold's one-vs-one pairs in one batch;
    with the same rand() stream it must predict exactly what the oracle's
    fold-by-fold LIBSVM loop predicts."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    X, y, _ = datasets.clusters_10class(scale=0.05, seed=21)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y.astype(np.float64))
    prm = ffi.make_param(gamma=2.0 ** -7, C_=4.0, eps=1e-3, shrinking=0)
    libc.srand(7)
    dev = hip.svm_cross_validation(prob, prm, 5)
    libc.srand(7)
    ref = ora.svm_cross_validation(prob, prm, 5)
    agree = np.mean(dev == ref)
    print(f"  CV l={prob.l} accuracy dev={np.mean(dev == y):.4f} ref={np.mean(ref == y):.4f} agree={agree}", flush=True)
    assert agree == 1.0


def test_svm_train_binary_small_uses_batch_and_matches(hip, ora):
    ip, ix, v, y, d = datasets.zipf_binary(n=2000)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, shrinking=0)
    m_dev = hip.svm_train(prob, prm).arrays()
    m_ref = ora.svm_train(prob, prm).arrays()
    np.testing.assert_array_equal(m_dev["sv_indices"], m_ref["sv_indices"])
    np.testing.assert_allclose(m_dev["sv_coef"], m_ref["sv_coef"], atol=1e-9)
    np.testing.assert_allclose(m_dev["rho"], m_ref["rho"], atol=1e-9)


@pytest.mark.parametrize("kname,svm_type", [("rbf", ffi.C_SVC), ("poly", ffi.C_SVC), ("rbf", ffi.NU_SVC),
                                            ("linear", ffi.C_SVC), ("rbf", ffi.EPSILON_SVR),
                                            ("rbf", ffi.ONE_CLASS), ("sigmoid", ffi.C_SVC)])
def test_predict_batch_matches_oracle(hip, ora, kname, svm_type):
    kt = {"rbf": ffi.RBF, "poly": ffi.POLY, "linear": ffi.LINEAR, "sigmoid": ffi.SIGMOID}[kname]
    X, y, _ = datasets.clusters_10class(scale=0.04, seed=5)
    if svm_type == ffi.EPSILON_SVR:
        y = X[:, 0].astype(np.float64) + 0.1 * X[:, 1]
    prob = ffi.Problem.from_dense(X.astype(np.float64), y.astype(np.float64))
    prm = ffi.make_param(svm_type=svm_type, kernel_type=kt, gamma=2.0 ** -7, coef0=0.1, degree=2,
                         C_=4.0, nu=0.2, eps=1e-3, shrinking=0)
    m_ref = ora.svm_train(prob, prm)
    rows = np.arange(0, prob.l, 3)
    lab_r, dec_r = m_ref.predict_values(prob, rows=rows)
    m_dev = hip.svm_train(prob, prm)
    lab_d, dec_d = m_dev.predict_batch(prob, rows=rows)
    lab_h, dec_h = m_dev.predict_values(prob, rows=rows)  # host path of the same model
    np.testing.assert_allclose(dec_d, dec_h, rtol=0, atol=1e-9)
    np.testing.assert_allclose(dec_d, dec_r, rtol=0, atol=1e-4)
    if svm_type != ffi.EPSILON_SVR:
        assert np.mean(lab_d == lab_r) >= 0.999


@pytest.mark.parametrize("dense", [False, True])
def test_batch_speculative_rows_same_results(hip, dense):
    """Single-workgroup solves with a full cache compute, on a row miss, the
    rows of the pass's other strong candidates in the same pass; the iterates
    must be bit-identical to solves without it (speculative_rows=-1)."""
    if dense:
        X, y = datasets.gmm_dense_matrix(n=4000, d=64, seed=11)
        prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
        gammas = [1.0 / 64, 0.1]
    else:
        ip, ix, v, y, d = datasets.zipf_binary(n=4000, seed=5)
        prob = ffi.Problem(ip, ix, v, y, d)
        gammas = [1.0 / d, 0.05]
    prm = ffi.make_param(gamma=gammas[0], C_=1.0, eps=1e-3, shrinking=0)
    rng = np.random.default_rng(1)
    out = {}
    for spec in (0, -1):
        b = hip.BatchSolver(prob, prm, team_max=1, speculative_rows=spec)
        subs = []
        for k in range(6):
            rows = np.sort(rng.choice(prob.l, size=int(rng.integers(800, 3000)), replace=False)) if spec == 0 \
                else out[0]["rows"][k]
            sub = prob.subset(rows)
            yy, p, a0, Cp, Cn, kind = hip.formulate(sub, prm)
            subs.append((rows, dict(rowset=b.add_rowset(rows), kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn,
                                    gamma=gammas[k % 2], eps=1e-3)))
        res, st = b.solve([s[1] for s in subs])
        out[spec] = dict(rows=[s[0] for s in subs], res=res, st=st)
    print(f"  rows computed: speculative {out[0]['st']['row_misses']}, plain {out[-1]['st']['row_misses']}",
          flush=True)
    for a, b_ in zip(out[0]["res"], out[-1]["res"]):
        assert a["status"] == 0 and b_["status"] == 0
        assert a["iters"] == b_["iters"]
        np.testing.assert_array_equal(a["alpha_raw"], b_["alpha_raw"])
    # (the rows computed are not compared: where a stage pauses depends on
    # timing, and a paused solve restarts with an empty row cache)


def test_batch_more_subproblems_than_workgroups(hip, ora):
    """More subproblems than the grid has workgroups (300 > 256 on one
    MI355X): most run one workgroup each while the few large ones get small
    teams in the same launch, and workgroups take arenas as they go; every
    subproblem must still give the oracle's iterates."""
    ip, ix, v, y, d = datasets.zipf_binary(n=4000, seed=13)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, shrinking=0)
    rng = np.random.default_rng(3)
    b = hip.BatchSolver(prob, prm)
    cases = []
    for k in range(300):
        n = int(rng.integers(2500, 3500)) if k % 30 == 0 else int(rng.integers(60, 200))
        rows = np.sort(rng.choice(prob.l, size=n, replace=False))
        sub = prob.subset(rows)
        if len(np.unique(np.where(sub.y > 0, 1, -1))) < 2:
            continue
        yy, p, a0, Cp, Cn, kind = hip.formulate(sub, prm)
        cases.append((sub, dict(rowset=b.add_rowset(rows), kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn,
                                gamma=prm.gamma, eps=1e-3)))
    assert len(cases) > 256
    res, st = b.solve([c[1] for c in cases])
    print(f"  {len(cases)} subproblems, {st['iters']} iterations, {st['launches']} launches", flush=True)
    for (sub, _), r in zip(cases, res):
        ref = ora.train_one(sub, prm)
        assert r["status"] == 0
        assert r["iters"] == ref["iters"], (sub.l, r["iters"], ref["iters"])
        np.testing.assert_allclose(r["alpha_raw"], ref["alpha_raw"], rtol=0, atol=1e-12)


def test_ovo_cv_gamma_grid_full_size(hip, ora):
    """The 10-class configuration at full size (n=38,217, d=128): 5-fold CV x
    4 gammas x 45 one-vs-one pairs = 900 solves in one batched call.  Every
    solve must converge to a feasible point and match the oracle's
    full-size results (tests/golden/fullsize_ovo_cv_grid.npz, reference-exact
    variant): identical iteration count and support-vector set, objective
    within 1e-6 relative, rho within 1e-5; the 5 smallest are also compared
    with a live oracle solve (alpha to 1e-9)."""
    gammas = [2.0 ** -9, 2.0 ** -7, 2.0 ** -5, 2.0 ** -3]
    C = 4.0
    X, y, fold = datasets.clusters_10class()
    prob = ffi.Problem.from_dense(X.astype(np.float64), y.astype(np.float64), keep_zeros=True)
    b = hip.BatchSolver(prob, ffi.make_param(gamma=gammas[0], C_=C, eps=1e-3, shrinking=0))
    jobs, subs = [], []
    for f in range(int(fold.max()) + 1):
        train = np.nonzero(fold != f)[0]
        labels = list(dict.fromkeys(y[train].tolist()))
        by = {c: train[y[train] == c] for c in labels}
        for i in range(len(labels)):
            for j in range(i + 1, len(labels)):
                rows = np.concatenate([by[labels[i]], by[labels[j]]])
                yy = np.concatenate([np.ones(len(by[labels[i]])), -np.ones(len(by[labels[j]]))]).astype(np.int8)
                rid = b.add_rowset(rows)
                for g in gammas:
                    jobs.append((rows, yy, g))
                    subs.append(dict(rowset=rid, kind=hip.KIND_C, y=yy, p=-np.ones(len(rows)),
                                     alpha0=np.zeros(len(rows)), Cp=C, Cn=C, gamma=g, eps=1e-3))
    assert len(subs) == 900
    res, st = b.solve(subs)
    print(f"  900 solves: {st['iters']} iterations, {st['launches']} launches, {st['solve_ms']:.0f} ms", flush=True)
    for (rows, yy, g), r in zip(jobs, res):
        a = r["alpha_raw"]
        assert r["status"] == 0
        assert np.all(a >= 0) and np.all(a <= C)
        assert abs(np.dot(a, yy)) < 1e-8 * max(1.0, a.sum())
        assert r["obj"] < 0
    import hashlib
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_ovo_cv_grid.npz"))
    worst = {"obj": 0.0, "rho": 0.0, "obj_fp64": 0.0, "rho_fp64": 0.0}
    for k, r in enumerate(res):
        sv = np.nonzero(r["alpha_raw"] > 0)[0].astype(np.int32)
        dig = np.frombuffer(hashlib.sha256(sv.tobytes()).digest()[:8], dtype=np.uint64)[0]
        assert r["iters"] == fx["float_iters"][k], k
        assert len(sv) == fx["float_nSV"][k] and dig == fx["float_sv_digest"][k], k
        eo = abs(r["obj"] - fx["float_obj"][k]) / abs(fx["float_obj"][k])
        er = abs(r["rho"] - fx["float_rho"][k])
        assert eo <= 1e-6 and er <= 1e-5, k
        worst["obj"], worst["rho"] = max(worst["obj"], eo), max(worst["rho"], er)
        worst["obj_fp64"] = max(worst["obj_fp64"], abs(r["obj"] - fx["double_obj"][k]) / abs(fx["double_obj"][k]))
        worst["rho_fp64"] = max(worst["rho_fp64"], abs(r["rho"] - fx["double_rho"][k]))
    print(f"  900 solves vs the oracle: worst obj rel err {worst['obj']:.1e}, rho err {worst['rho']:.1e} "
          f"(reference-exact); {worst['obj_fp64']:.1e} / {worst['rho_fp64']:.1e} (pure FP64)", flush=True)
    order = np.argsort([len(j[0]) for j in jobs], kind="stable")[:5]
    for k in order:
        rows, yy, g = jobs[k]
        sub = ffi.Problem.from_dense(X[rows].astype(np.float64), yy.astype(np.float64), keep_zeros=True)
        ref = ora.train_one(sub, ffi.make_param(gamma=g, C_=C, eps=1e-3, shrinking=0))
        assert res[k]["iters"] == ref["iters"], (len(rows), g)
        np.testing.assert_allclose(res[k]["alpha_raw"], ref["alpha_raw"], rtol=0, atol=1e-9)
        assert abs(res[k]["obj"] - ref["obj"]) <= 1e-9 * abs(ref["obj"])


@pytest.mark.parametrize("dense", [True, False])
@pytest.mark.parametrize("kname", ["rbf", "linear", "poly", "sigmoid"])
def test_predict_batch_fp64_features_match_oracle(hip, ora, kname, dense):
    """Features that are not float32 values, in the model's support vectors
    and in the test rows: the device predicts from the FP64 features as
    svm_predict_values does (kernel values in k_function's order, each
    decision value summed in the reference's order), so its decision values
    equal the oracle's to 1e-10 relative (device exp / tanh against glibc's
    aside, the arithmetic is the same; sparse RBF rows add the terms in
    another order)."""
    kt = {"rbf": ffi.RBF, "poly": ffi.POLY, "linear": ffi.LINEAR, "sigmoid": ffi.SIGMOID}[kname]
    rng = np.random.default_rng(21)
    n, d = 900, 40

    def rows(m):
        X = rng.standard_normal((m, d)) * 1.37
        if not dense:
            X[rng.random((m, d)) < 0.7] = 0.0
        return X

    X = rows(n)
    assert np.any(X.astype(np.float32).astype(np.float64) != X)
    y = np.where(X[:, 0] + 0.5 * X[:, 1] > 0, 1.0, -1.0)
    prob = ffi.Problem.from_dense(X, y)
    prm = ffi.make_param(kernel_type=kt, gamma=1.0 / d, coef0=0.1, degree=2, C_=2.0, shrinking=0, cuda_flag=0)
    m = ora.svm_train(prob, prm)
    test = ffi.Problem.from_dense(rows(700), np.zeros(700))
    lab_r, dec_r = m.predict_values(test)
    lab_d, dec_d = hip.predict_batch(m.ptr, test)
    err = np.max(np.abs(dec_d - dec_r))
    print(f"  {kname} dense={dense}: nSV={m.arrays()['l']} max |dec - oracle| = {err:.2e}", flush=True)
    assert err <= 1e-10 * max(1.0, np.max(np.abs(dec_r)))
    np.testing.assert_array_equal(lab_d, lab_r)


@pytest.mark.parametrize("dense", [True, False])
def test_predict_batch_row_counts_across_launch_split(hip, ora, dense):
    """Row counts below, at and past whole rounds of 16-row workgroups over
    the device's compute units: the rows of the last partial round run at a
    smaller rows-per-workgroup, and a row's decision value does not depend on
    it (bit-identical to the same row predicted in a batch of 3 or alone),
    and matches the oracle to 1e-10 relative."""
    rng = np.random.default_rng(5)
    d = 48

    def rows(m):
        X = rng.standard_normal((m, d))
        if not dense:
            X[rng.random((m, d)) < 0.6] = 0.0
        return X

    X = rows(1500)
    y = np.where(X[:, 0] - 0.3 * X[:, 2] > 0, 1.0, -1.0)
    prm = ffi.make_param(gamma=0.5 / d, C_=4.0, shrinking=0, cuda_flag=0)
    m = ora.svm_train(ffi.Problem.from_dense(X, y), prm)
    T = rows(16 * 256 * 3 + 1001)
    for n in (1, 3, 255, 16 * 256 + 37, T.shape[0]):
        test = ffi.Problem.from_dense(T[:n], np.zeros(n))
        _, dec_d = hip.predict_batch(m.ptr, test)
        _, dec_r = m.predict_values(test)
        dec_d, dec_r = dec_d.ravel(), dec_r.ravel()
        assert np.max(np.abs(dec_d - dec_r)) <= 1e-10 * max(1.0, np.max(np.abs(dec_r))), n
        if n == 1:
            first = dec_d[:1].copy()
        elif n == 3:
            np.testing.assert_array_equal(dec_d[:1], first)
            first3 = dec_d[:3].copy()
        else:
            np.testing.assert_array_equal(dec_d[:3], first3)
        print(f"  dense={dense} rows={n}: max |dec - oracle| = {np.max(np.abs(dec_d - dec_r)):.2e}", flush=True)


@pytest.mark.parametrize("case", ["rbf-dense", "rbf-sparse", "linear-dense", "poly-sparse", "sigmoid-dense",
                                  "svr-rbf", "oneclass-rbf", "multiclass-rbf"])
def test_predict_batch_mfma_close_to_oracle(hip, ora, case):
    """The opt-in Gram-tile prediction (mfma_rows = 1: FP16 hi/lo split
    dots on the matrix cores, FP32 accumulation, FP64 kernel values and
    sums) against the oracle's FP64 predict_values: every decision value
    within 1e-4 of the decision values' scale (max(1, max |dec|)), the
    median within 1e-6 of it, and the labels equal wherever |dec| is above
    that bound.  A model with several coefficient rows (more than two
    classes) runs the FP64 kernel instead, bit for bit as without the
    option."""
    kname, layout = case.split("-")
    rng = np.random.default_rng(77)
    n, d = 1200, 96

    def rows(m):
        X = rng.standard_normal((m, d)) * 0.8
        if layout == "sparse":
            X[rng.random((m, d)) < 0.7] = 0.0
        return X

    X = rows(n)
    kt = {"rbf": ffi.RBF, "linear": ffi.LINEAR, "poly": ffi.POLY, "sigmoid": ffi.SIGMOID}.get(kname, ffi.RBF)
    if kname == "svr":
        y, st = X[:, 0] - 0.4 * X[:, 3] + 0.1 * rng.standard_normal(n), ffi.EPSILON_SVR
    elif kname == "oneclass":
        y, st = np.ones(n), ffi.ONE_CLASS
    elif kname == "multiclass":
        y, st = np.digitize(X[:, 0] + 0.5 * X[:, 1], [-0.6, 0.0, 0.7]).astype(float), ffi.C_SVC
    else:
        y, st = np.where(X[:, 0] + 0.5 * X[:, 1] > 0, 1.0, -1.0), ffi.C_SVC
    prm = ffi.make_param(svm_type=st, kernel_type=kt, gamma=0.5 / d, coef0=0.2, degree=3, C_=2.0, nu=0.3, p=0.1,
                         shrinking=0, cuda_flag=0)
    m = ora.svm_train(ffi.Problem.from_dense(X, y), prm)
    test = ffi.Problem.from_dense(rows(3000), np.zeros(3000))
    lab_r, dec_r = m.predict_values(test)
    lab_m, dec_m = hip.predict_batch(m.ptr, test, mfma=True)
    scale = max(1.0, float(np.max(np.abs(dec_r))))
    err = np.abs(dec_m - dec_r)
    print(f"  {case}: nSV={m.arrays()['l']} max |dec - oracle| = {err.max():.2e}, median {np.median(err):.2e}, "
          f"scale {scale:.2f}", flush=True)
    if kname == "multiclass":
        lab_e, dec_e = hip.predict_batch(m.ptr, test)
        np.testing.assert_array_equal(dec_m, dec_e)
        return
    assert err.max() <= 1e-4 * scale
    assert np.median(err) <= 1e-6 * scale
    sure = np.abs(dec_r).ravel() > 1e-4 * scale
    if st == ffi.C_SVC or st == ffi.ONE_CLASS:
        np.testing.assert_array_equal(lab_m[sure], lab_r[sure])


@pytest.mark.parametrize("kname", ["rbf", "poly", "linear"])
def test_predict_batch_matches_sklearn_decision_function(hip, kname):
    """The device decision values of a model trained by scikit-learn's
    bundled LIBSVM (tests/golden/sklearn_multiclass.npz, 4 classes) equal
    sklearn's decision_function to 1e-12."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from model_util import ArrayModel
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sklearn_multiclass.npz"))
    kt = {"rbf": ffi.RBF, "poly": ffi.POLY, "linear": ffi.LINEAR}[kname]
    prm = ffi.make_param(kernel_type=kt, gamma=0.2, coef0=0.5, degree=2, C_=2.0)
    sup = G[f"{kname}_support"]
    am = ArrayModel(prm, G["X"][sup], G[f"{kname}_dual_coef"], -G[f"{kname}_intercept"], label=[1, 2, 3, 4],
                    nSV=G[f"{kname}_n_support"])
    test = ffi.Problem.from_dense(G["Xt"], np.zeros(len(G["Xt"])))
    lab, dec = hip.predict_batch(am.ptr, test)
    np.testing.assert_allclose(dec, G[f"{kname}_dec"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(lab, G[f"{kname}_pred"])


def test_batch_hbm_fraction_caps_arenas_same_results(hip):
    """A small hbm_fraction (kernel-row arenas from 0.2% of free memory,
    LRU under pressure) changes no result."""
    X, y, fold = datasets.clusters_10class(scale=0.08, seed=2)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y.astype(np.float64))
    prm = ffi.make_param(gamma=2.0 ** -7, C_=4.0, eps=1e-3, shrinking=0)
    subs, rows_of = [], []
    b_def = hip.BatchSolver(prob, prm)
    L = hip.lib()
    opt = hip.svmhip_options(C.sizeof(hip.svmhip_options), 0, 0, 0, 0, 0.0, 0, 0, 0, 0, 0, 0.002)
    h = C.c_void_p()
    assert L.svmhip_batch_create(C.byref(prob.struct), C.byref(prm), C.byref(opt), C.byref(h)) == 0
    b_small = hip.BatchSolver.__new__(hip.BatchSolver)
    b_small._h, b_small.problem, b_small.param = h, prob, prm
    for c1, c2 in [(1, 2), (3, 7), (5, 9)]:
        rows = np.nonzero((y == c1) | (y == c2))[0]
        yy = np.where(y[rows] == c1, 1, -1).astype(np.int8)
        rid1, rid2 = b_def.add_rowset(rows), b_small.add_rowset(rows)
        assert rid1 == rid2
        subs.append(dict(rowset=rid1, kind=hip.KIND_C, y=yy, p=-np.ones(len(rows)), alpha0=np.zeros(len(rows)),
                         Cp=4.0, Cn=4.0, gamma=2.0 ** -7, eps=1e-3))
    r1, s1 = b_def.solve(subs)
    r2, s2 = b_small.solve(subs)
    assert s2["row_misses"] >= s1["row_misses"]
    for a, b in zip(r1, r2):
        assert a["iters"] == b["iters"]
        np.testing.assert_array_equal(a["alpha_raw"], b["alpha_raw"])
    b_def.close()
    b_small.close()


def test_batch_shared_dots_dense_match_oracle(hip, ora, monkeypatch):
    """Dense problems: kernel rows gathered from the handle's exact dot
    products of the base rows (shared_dots = 1) give the oracle's iterates
    (C-SVC, nu-SVC and epsilon-SVR subsets, several gammas), bit-identical to
    rows computed from the subsets' own rows (SVMHIP_SHARED_DOTS=0).  A
    second solve on the same handle, with fewer and smaller subproblems,
    allocates nothing and gives the results of a fresh handle."""
    X, y = datasets.gmm_dense_matrix(n=3000, d=64, seed=21)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
    prm = ffi.make_param(gamma=1.0 / 64, C_=1.0, eps=1e-3, shrinking=0)
    rng = np.random.default_rng(4)
    specs = []
    for k in range(9):
        rows = np.sort(rng.choice(prob.l, size=int(rng.integers(300, 2400)), replace=False))
        svm_type = [ffi.C_SVC, ffi.NU_SVC, ffi.EPSILON_SVR][k % 3]
        gamma = [1.0 / 64, 0.05, 0.003][k % 3]
        specs.append((rows, ffi.make_param(svm_type=svm_type, gamma=gamma, C_=[0.5, 2.0, 1.0][k % 3], nu=0.4,
                                           p=0.1, eps=1e-3, shrinking=0)))

    def subs_for(b, which):
        out = []
        for rows, sp in (specs[i] for i in which):
            sub = prob.subset(rows)
            yy, p, a0, Cp, Cn, kind = hip.formulate(sub, sp)
            out.append((sub, sp, yy, dict(rowset=b.add_rowset(rows), kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp,
                                          Cn=Cn, gamma=sp.gamma, eps=1e-3)))
        return out

    b = hip.BatchSolver(prob, prm)
    cases = subs_for(b, range(9))
    res, st = b.solve([c[3] for c in cases])
    assert st["shared_dots"] == 1
    for (sub, sp, yy, _), r in zip(cases, res):
        ref = ora.train_one(sub, sp)
        print(f"  l={sub.l} type={sp.svm_type} iters {r['iters']} vs {ref['iters']}", flush=True)
        assert r["status"] == 0
        assert r["iters"] == ref["iters"]
        coef, rho, obj = hip.finish_solution(sub.y, sp, r, yy)
        np.testing.assert_allclose(coef, ref["alpha"], rtol=0, atol=1e-9)
        assert abs(obj - ref["obj"]) <= 1e-9 * abs(ref["obj"])
    # the same handle again: fewer, smaller subproblems over its row sets
    # (adding a row set allocates; a solve that fits the kept buffers not)
    small = sorted(range(9), key=lambda i: len(specs[i][0]))[:4]
    res2, st2 = b.solve([cases[i][3] for i in small])
    assert st2["device_allocs"] == st["device_allocs"]
    monkeypatch.setenv("SVMHIP_SHARED_DOTS", "0")
    b0 = hip.BatchSolver(prob, prm)
    plain = subs_for(b0, range(9))
    res0, st0 = b0.solve([c[3] for c in plain])
    assert st0["shared_dots"] == 0
    for a, c in zip(res, res0):
        assert a["iters"] == c["iters"]
        np.testing.assert_array_equal(a["alpha_raw"], c["alpha_raw"])
    for i, r in zip(small, res2):
        assert r["iters"] == res0[i]["iters"]
        np.testing.assert_array_equal(r["alpha_raw"], res0[i]["alpha_raw"])
    b.close()
    b0.close()


def test_batch_kept_buffers_grow_for_a_larger_solve(hip):
    """A handle whose first solve was small and whose second needs more of
    every buffer (more subproblems, more variables, SVR doubling them)
    reallocates and gives the results of a fresh handle."""
    X, y = datasets.gmm_dense_matrix(n=2500, d=32, seed=31)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
    prm = ffi.make_param(gamma=1.0 / 32, C_=1.0, eps=1e-3, shrinking=0)
    rng = np.random.default_rng(8)

    def subs(b, sizes, svr_every):
        out = []
        for k, n in enumerate(sizes):
            rows = np.sort(rng.choice(prob.l, size=n, replace=False))
            sub = prob.subset(rows)
            sp = ffi.make_param(svm_type=ffi.EPSILON_SVR if k % svr_every == 0 else ffi.C_SVC, gamma=1.0 / 32,
                                C_=1.0, p=0.1, eps=1e-3, shrinking=0)
            yy, p, a0, Cp, Cn, kind = hip.formulate(sub, sp)
            out.append((rows, dict(rowset=b.add_rowset(rows), kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn,
                                   gamma=sp.gamma, eps=1e-3), sp))
        return out

    b = hip.BatchSolver(prob, prm)
    small = subs(b, [200, 300], 99)
    _, st1 = b.solve([s[1] for s in small])
    big = subs(b, [900, 1500, 2200, 1100, 700], 2)
    res, st2 = b.solve([s[1] for s in big])
    assert st2["device_allocs"] > st1["device_allocs"]
    fresh = hip.BatchSolver(prob, prm)
    ref = []
    for rows, d, sp in big:
        d2 = dict(d)
        d2["rowset"] = fresh.add_rowset(rows)
        ref.append(d2)
    res0, _ = fresh.solve(ref)
    for a, c in zip(res, res0):
        assert a["status"] == 0 and a["iters"] == c["iters"]
        np.testing.assert_array_equal(a["alpha_raw"], c["alpha_raw"])
    b.close()
    fresh.close()


def test_batch_default_arena_budget(hip):
    """With neither cache_bytes nor hbm_fraction given, the kernel-row arenas
    take at most 32 GiB in all, so a handle per call (svm_train,
    svm_cross_validation) does not re-allocate most of the device memory each
    time; the solve matches one with smaller arenas given by cache_bytes."""
    X, y = datasets.gmm_dense_matrix(n=20000, d=16, seed=41)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
    prm = ffi.make_param(gamma=1.0 / 16, C_=1.0, eps=1e-3, shrinking=0)
    rows = np.arange(0, 20000, 2)
    n = len(rows)
    stride = (n + 63) // 64 * 64
    sub = prob.subset(rows)
    yy, p, a0, Cp, Cn, kind = hip.formulate(sub, prm)
    out = []
    for cache_bytes in (0, 64 << 20):
        b = hip.BatchSolver(prob, prm, cache_bytes=cache_bytes)
        rid = b.add_rowset(rows)
        res, st = b.solve([dict(rowset=rid, kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn, gamma=prm.gamma,
                                eps=1e-3)])
        total = st["cache_slots"] * stride * 4 * st["grid_blocks"]
        print(f"  cache_bytes={cache_bytes}: {st['cache_slots']} slots x {st['grid_blocks']} arenas = "
              f"{total / 2 ** 30:.1f} GiB, iters {res[0]['iters']}", flush=True)
        assert res[0]["status"] == 0
        out.append((res[0], st))
        b.close()
    (r0, st0), (r1, st1) = out
    assert st0["cache_slots"] * stride * 4 * st0["grid_blocks"] <= 32 * 2 ** 30
    assert st0["cache_slots"] < n  # the whole matrix in each arena would take 100 GB
    assert st1["cache_slots"] == (64 << 20) // (stride * 4)
    assert r0["iters"] == r1["iters"]
    np.testing.assert_array_equal(r0["alpha_raw"], r1["alpha_raw"])
