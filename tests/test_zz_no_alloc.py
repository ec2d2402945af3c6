"""svmhip_predict_batch_noalloc: the prediction entry point that allocates
no device memory (its rows, coefficients and results are staged in the
library's fixed device scratch area).  Same kernels as svmhip_predict_batch,
so the decision values and labels must be bit-identical to it, call after
call, for smaller inputs after larger ones, and for test sets that pass
through the scratch area in several chunks; a model too large for the
scratch area is refused with SVMHIP_ERR_NOMEM."""
import ctypes as C

import numpy as np
import pytest

from libsvm_hip import datasets, ffi

pytestmark = pytest.mark.gpu


def _model_and_test(ora, n_test, seed=3):
    X, y, _ = datasets.clusters_10class(scale=0.05, seed=seed)
    prob = ffi.Problem.from_dense(X.astype(np.float64) * 1.01, y.astype(np.float64))  # not float32 values
    prm = ffi.make_param(gamma=2.0 ** -7, C_=4.0, eps=1e-3, shrinking=0, cuda_flag=0)
    m = ora.svm_train(prob, prm)
    rng = np.random.default_rng(seed)
    Xt = X[rng.integers(0, len(X), n_test)].astype(np.float64) + 0.3 * rng.standard_normal((n_test, X.shape[1]))
    return prob, m, ffi.Problem.from_dense(Xt, np.zeros(n_test))


def test_noalloc_matches_allocating_entry_point(hip, ora):
    prob, m, test = _model_and_test(ora, 3000)
    for rows in (np.arange(3000), np.arange(0, 3000, 7)):  # the second call smaller than the first
        lab_a, dec_a = hip.predict_batch(m.ptr, test, rows)
        lab_n, dec_n = hip.predict_batch(m.ptr, test, rows, noalloc=True)
        np.testing.assert_array_equal(dec_n, dec_a)
        np.testing.assert_array_equal(lab_n, lab_a)
        lab_r, dec_r = m.predict_values(test, rows=rows)
        assert np.max(np.abs(dec_n - dec_r)) <= 1e-9 * max(1.0, np.max(np.abs(dec_r)))
        print(f"  {len(rows)} rows x {dec_n.shape[1]} pairs: bit-identical to svmhip_predict_batch", flush=True)


def test_noalloc_test_rows_in_chunks(hip, ora):
    """90,000 dense test rows of 256 features (184 MB as FP64) pass through
    the 128 MiB scratch area in several chunks."""
    X, y = datasets.gmm_dense_matrix(n=3000, d=256, seed=9)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y, keep_zeros=True)
    prm = ffi.make_param(gamma=1.0 / 256, C_=1.0, eps=1e-3, shrinking=0, cuda_flag=0)
    m = ora.svm_train(prob, prm)
    Xt, _ = datasets.gmm_dense_matrix(n=90000, d=256, seed=10)
    test = ffi.Problem.from_dense(Xt.astype(np.float64), np.zeros(90000), keep_zeros=True)
    lab_a, dec_a = hip.predict_batch(m.ptr, test)
    lab_n, dec_n = hip.predict_batch(m.ptr, test, noalloc=True)
    np.testing.assert_array_equal(dec_n, dec_a)
    np.testing.assert_array_equal(lab_n, lab_a)


def test_noalloc_refuses_model_beyond_scratch(hip):
    """Support vectors of 70,000 x 256 dense FP64 features (143 MB) do not
    fit in the 128 MiB scratch area: SVMHIP_ERR_NOMEM, nothing computed."""
    nsv, d = 70000, 256
    rng = np.random.default_rng(1)
    sv = ffi.Problem.from_dense(rng.standard_normal((nsv, d)), np.ones(nsv), keep_zeros=True)
    coef = np.full(nsv, 0.5)
    coef_rows = (C.POINTER(C.c_double) * 1)(coef.ctypes.data_as(C.POINTER(C.c_double)))
    rho = (C.c_double * 1)(0.25)
    label = (C.c_int * 2)(1, -1)
    nSV = (C.c_int * 2)(nsv // 2, nsv - nsv // 2)
    model = ffi.svm_model()
    model.param = ffi.make_param(gamma=1.0 / d)
    model.nr_class, model.l = 2, nsv
    model.SV = sv.struct.x
    model.sv_coef = C.cast(coef_rows, C.POINTER(C.POINTER(C.c_double)))
    model.rho = C.cast(rho, C.POINTER(C.c_double))
    model.label = C.cast(label, C.POINTER(C.c_int))
    model.nSV = C.cast(nSV, C.POINTER(C.c_int))
    test = ffi.Problem.from_dense(rng.standard_normal((10, d)), np.zeros(10), keep_zeros=True)
    ptrs = (C.POINTER(ffi.svm_node) * 10)(*[test.row_nodes(r) for r in range(10)])
    dec = np.zeros(10)
    rc = hip.lib().svmhip_predict_batch_noalloc(C.byref(model), ptrs, 10, None,
                                                dec.ctypes.data_as(C.POINTER(C.c_double)), None)
    assert rc == 3  # SVMHIP_ERR_NOMEM
    lab_a, dec_a = hip.predict_batch(C.pointer(model), test)  # the allocating entry point takes it
    assert np.all(np.isfinite(dec_a))
