"""Batched solves (one workgroup per subproblem) against the oracle: each
subproblem of a batch must give the same iterates as the oracle on the same
row subset (identical iteration count, alpha to 1e-9, objective to 1e-9)."""
import numpy as np
import pytest

from libsvm_hip import datasets, ffi

pytestmark = pytest.mark.gpu


def test_batch_subsets_match_oracle(hip, ora):
    ip, ix, v, y, d = datasets.zipf_binary(n=3000, seed=9)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, shrinking=0)
    rng = np.random.default_rng(0)
    b = hip.BatchSolver(prob, prm)
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


def test_batch_small_arena_lru(hip, ora):
    """Arenas too small for a full cache: LRU eviction inside each workgroup."""
    ip, ix, v, y, d = datasets.zipf_binary(n=2000, seed=2)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=1.0 / d, C_=1.0, eps=1e-3, shrinking=0)
    rows = np.arange(0, 2000, 2)
    b = hip.BatchSolver(prob, prm, cache_bytes=4 * 1024 * 5)  # 5 rows of 1024 floats
    rid = b.add_rowset(rows)
    sub = prob.subset(rows)
    yy, p, a0, Cp, Cn, kind = hip.formulate(sub, prm)
    res, st = b.solve([dict(rowset=rid, kind=kind, y=yy, p=p, alpha0=a0, Cp=Cp, Cn=Cn, gamma=prm.gamma, eps=1e-3)] * 3)
    ref = ora.train_one(sub, prm)
    for r in res:
        assert r["iters"] == ref["iters"]
        np.testing.assert_allclose(r["alpha_raw"], ref["alpha_raw"], rtol=0, atol=1e-12)
    assert st["row_misses"] > 3 * 100
