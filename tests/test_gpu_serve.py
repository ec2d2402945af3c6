"""Serving (svmhip_server_*): requests solved together must give exactly what
each gives alone, and what the oracle's svm_train_one gives on the same rows
(identical iteration count, coefficients to 1e-9, objective to 1e-9
relative, rho to 1e-9); predictions against a resident model must equal
sum_i coef_i K(x_i, t) - rho computed in FP64 from that model (1e-9); the
HBM budget must bound what is admitted."""
import numpy as np
import pytest

from libsvm_hip import datasets, ffi

pytestmark = pytest.mark.gpu


def request_problem(r):
    return ffi.Problem(r["indptr"], r["indices"], r["data"], r["y"], r["d"])


def request_param(r):
    t = ffi.NU_SVC if r["kind"] == "nu_svc" else ffi.C_SVC
    return ffi.make_param(svm_type=t, gamma=r["gamma"], C_=r["C"], nu=r["nu"], eps=1e-3, shrinking=0)


def oracle_view(prob):
    """The same rows with y = +1 for the first label seen, -1 otherwise."""
    y = np.where(prob.y == prob.y[0], 1.0, -1.0)
    return ffi.Problem(prob.indptr, prob.indices, prob.data, y, prob.n_features)


def solve_alone(hip, prob, prm, stage_ms):
    s = hip.Server()
    rid = s.submit_train(prob, prm)
    comps = s.run_until_idle(stage_ms=stage_ms)
    coef, lab = s.train_result(rid, prob.l)
    s.close()
    return comps[0], coef


def test_server_mix_matches_alone_and_oracle(hip, ora):
    reqs = [datasets.serving_request(k, n_range=(300, 3000)) for k in range(10)]
    probs = [request_problem(r) for r in reqs]
    prms = [request_param(r) for r in reqs]
    s = hip.Server()
    ids = [s.submit_train(p, q) for p, q in zip(probs, prms)]
    comps = {c["id"]: c for c in s.run_until_idle(stage_ms=1.0)}  # short stages: many resumes
    assert sorted(comps) == ids
    for rid, r, prob, prm in zip(ids, reqs, probs, prms):
        c = comps[rid]
        coef, lab = s.train_result(rid, prob.l)
        assert lab[0] == prob.y[0] and lab[1] != lab[0]
        alone, coef_alone = solve_alone(hip, prob, prm, stage_ms=50.0)
        ref = ora.train_one(oracle_view(prob), prm)
        print(f"  {r['kind']} n={r['n']} d={r['d']} dens={r['density']}: iters mix {c['iters']} alone "
              f"{alone['iters']} oracle {ref['iters']}, obj {c['obj']:.10g} vs {ref['obj']:.10g}", flush=True)
        assert c["status"] == 0 and alone["status"] == 0
        assert c["iters"] == alone["iters"] == ref["iters"]
        np.testing.assert_array_equal(coef, coef_alone)
        assert c["rho"] == alone["rho"] and c["obj"] == alone["obj"]
        np.testing.assert_allclose(coef, ref["alpha"], rtol=0, atol=1e-9)
        assert abs(c["obj"] - ref["obj"]) <= 1e-9 * abs(ref["obj"])
        assert abs(c["rho"] - ref["rho"]) <= 1e-9
        assert c["nSV"] == int(np.count_nonzero(ref["alpha"]))
    s.close()


def test_server_predict_against_resident_model(hip):
    r = datasets.serving_request(3, n_range=(800, 2000))
    prob, prm = request_problem(r), request_param(r)
    s = hip.Server()
    mid = s.submit_train(prob, prm)
    (comp,) = s.run_until_idle()
    coef, lab = s.train_result(mid, prob.l)
    ip, ix, v = datasets.serving_rows(77, 3000, r["d"], r["density"])
    test = ffi.Problem(ip, ix, v, np.zeros(3000), r["d"])
    pid = s.submit_predict(mid, test)
    # a second train request in flight at the same time as the prediction
    r2 = datasets.serving_request(4, n_range=(800, 2000))
    s.submit_train(request_problem(r2), request_param(r2))
    assert len(s.run_until_idle()) == 2
    dec, plab = s.predict_result(pid, 3000)
    # FP64 reference from the model's own coefficients and rho
    X = prob.to_dense()
    T = test.to_dense()
    sv = np.nonzero(coef)[0]
    assert len(sv) == comp["nSV"] > 0
    Xs = X[sv]
    d2 = (Xs * Xs).sum(1)[None, :] + (T * T).sum(1)[:, None] - 2.0 * T @ Xs.T
    ref = np.exp(-r["gamma"] * d2) @ coef[sv] - comp["rho"]
    np.testing.assert_allclose(dec, ref, rtol=0, atol=1e-9)
    assert np.all(plab == np.where(dec > 0, lab[0], lab[1]))
    s.close()


def test_server_budget_admission(hip):
    reqs = [datasets.serving_request(k, n_range=(1000, 1500)) for k in range(20, 28)]
    probs = [request_problem(r) for r in reqs]
    budget = 16 * 2 ** 20  # two or three requests at a time
    s = hip.Server(budget_bytes=budget, cache_bytes=4 * 2 ** 20)
    ids = [s.submit_train(p, request_param(r)) for p, r in zip(probs, reqs)]
    peak, seen_queue, done = 0.0, False, []
    for _ in range(100000):
        a, q = s.step(2.0)
        peak = max(peak, s.bytes_in_use())
        seen_queue |= q > 0
        done += s.poll()
        if a == 0 and q == 0:
            break
    print(f"  peak bytes {peak / 2 ** 20:.1f} MiB of {budget / 2 ** 20:.0f}, queued at some step: {seen_queue}")
    assert sorted(c["id"] for c in done) == ids
    assert all(c["status"] == 0 for c in done)
    assert peak <= budget
    assert seen_queue  # the budget held requests back
    s.close()
    # a request larger than the whole budget completes with SVMHIP_ERR_NOMEM
    tiny = hip.Server(budget_bytes=2 * 2 ** 20, cache_bytes=4 * 2 ** 20)
    rid = tiny.submit_train(probs[0], request_param(reqs[0]))
    res = tiny.run_until_idle()
    assert [(c["id"], c["status"]) for c in res] == [(rid, 3)]
    tiny.close()


@pytest.mark.parametrize("k", [32, 40])  # nu-SVC, d=512 dense / d=128 at 30% density
def test_server_nu_initialisation_split_over_stages(hip, ora, k):
    """A nu-SVC request starts from alpha0 with nu*l/2 nonzeros per class, so
    G = p + Q alpha0 is most of its first work; with short stages the server
    spreads it over many launches.  The result must not change."""
    r = datasets.serving_request(k, n_range=(1500, 2500))
    assert r["kind"] == "nu_svc"
    prob, prm = request_problem(r), request_param(r)
    s = hip.Server()
    rid = s.submit_train(prob, prm)
    (c,) = s.run_until_idle(stage_ms=0.05)
    coef, _ = s.train_result(rid, prob.l)
    launches = s.stats()["launches"]
    s.close()
    alone, coef_alone = solve_alone(hip, prob, prm, stage_ms=1000.0)
    ref = ora.train_one(oracle_view(prob), prm)
    print(f"  n={r['n']} d={r['d']}: {launches} launches, iters {c['iters']} vs {ref['iters']}", flush=True)
    assert launches > 5
    assert c["status"] == 0 and c["iters"] == alone["iters"] == ref["iters"]
    np.testing.assert_array_equal(coef, coef_alone)
    np.testing.assert_allclose(coef, ref["alpha"], rtol=0, atol=1e-9)
    assert abs(c["obj"] - ref["obj"]) <= 1e-9 * abs(ref["obj"])
    assert abs(c["rho"] - ref["rho"]) <= 1e-9


def test_server_release_model_before_fetching_prediction(hip):
    """A finished prediction stays readable after its model is released: the
    request keeps the model's two labels."""
    r = datasets.serving_request(5, n_range=(600, 1200))
    prob, prm = request_problem(r), request_param(r)
    s = hip.Server()
    mid = s.submit_train(prob, prm)
    s.run_until_idle()
    ip, ix, v = datasets.serving_rows(77, 300, r["d"], r["density"])
    test = ffi.Problem(ip, ix, v, np.zeros(300), r["d"])
    pid = s.submit_predict(mid, test)
    s.run_until_idle()
    dec0, lab0 = s.predict_result(pid, 300)
    s.release(mid)
    dec1, lab1 = s.predict_result(pid, 300)
    np.testing.assert_array_equal(dec0, dec1)
    np.testing.assert_array_equal(lab0, lab1)
    assert set(np.unique(lab1)) <= set(np.unique(prob.y))
    s.release(pid)
    s.close()


def test_server_serving_slice_at_real_sizes(hip):
    """A fixed slice of the serving stream at its real sizes (n up to 19,974,
    dense d=512 rows, nu-SVC requests, C up to 8; ids in
    tests/golden/fullsize_serving_slice.npz) solved together in one server,
    with a 10,000-row prediction submitted against the dense d=512 model as
    soon as it is resident (while the others still train).  Every result
    must be bit-identical to the same request solved alone, and equal the
    oracle's (reference-exact variant): same iteration count and support
    vectors, objective within 1e-6 relative, rho within 1e-5, coefficients
    within 1e-6 C; decision values within 1e-9 of the oracle's."""
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fullsize_serving_slice.npz"))
    ids = [int(k) for k in fx["ids"]]
    pm = int(fx["pred_model"])
    reqs = {k: datasets.serving_request(k) for k in ids}
    probs = {k: request_problem(reqs[k]) for k in ids}
    prms = {k: request_param(reqs[k]) for k in ids}
    r = reqs[pm]
    ip, ix, v = datasets.serving_rows(1005 + r["d"], 10000, r["d"], 0.3)
    test = ffi.Problem(ip, ix, v, np.zeros(10000), r["d"])
    s = hip.Server()
    rid = {k: s.submit_train(probs[k], prms[k]) for k in ids}
    comps, pid = {}, None
    for _ in range(200000):
        a, q = s.step(10.0)
        for c in s.poll():
            comps[c["id"]] = c
            if c["id"] == rid[pm]:
                pid = s.submit_predict(rid[pm], test)
        if a == 0 and q == 0 and pid is not None and pid in comps:
            break
    assert len(comps) == len(ids) + 1
    dec, lab = s.predict_result(pid, 10000)
    err = np.max(np.abs(dec - fx["pred_dec"]))
    print(f"  prediction of 10,000 rows against request {pm}: max |dec - oracle| = {err:.2e}", flush=True)
    assert err <= 1e-9 * max(1.0, np.max(np.abs(fx["pred_dec"])))
    np.testing.assert_array_equal(lab, np.where(fx["pred_dec"] > 0, fx["pred_labels"][0], fx["pred_labels"][1]))
    for k in ids:
        c = comps[rid[k]]
        coef, _ = s.train_result(rid[k], probs[k].l)
        alone, coef_alone = solve_alone(hip, probs[k], prms[k], stage_ms=10.0)
        sv = np.nonzero(coef != 0)[0]
        it, obj, rho = int(fx[f"float_{k}_iters"]), float(fx[f"float_{k}_obj"]), float(fx[f"float_{k}_rho"])
        rr = reqs[k]
        print(f"  request {k}: {rr['kind']} n={rr['n']} d={rr['d']} dens={rr['density']} C={rr['C']}: iters "
              f"{c['iters']} (oracle {it}), obj rel err {abs(c['obj'] - obj) / abs(obj):.1e}, rho err "
              f"{abs(c['rho'] - rho):.1e}, nSV {len(sv)}; pure-FP64 oracle obj rel err "
              f"{abs(c['obj'] - float(fx[f'double_{k}_obj'])) / abs(obj):.1e}", flush=True)
        assert c["status"] == 0 and alone["status"] == 0
        assert c["iters"] == alone["iters"] == it
        np.testing.assert_array_equal(coef, coef_alone)
        assert c["rho"] == alone["rho"] and c["obj"] == alone["obj"]
        assert abs(c["obj"] - obj) <= 1e-6 * abs(obj)
        assert abs(c["rho"] - rho) <= 1e-5
        np.testing.assert_array_equal(sv, fx[f"float_{k}_sv"])
        assert np.max(np.abs(coef[sv] - fx[f"float_{k}_coef_sv"])) <= 1e-6 * rr["C"]
    s.close()
