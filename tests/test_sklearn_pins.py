"""Pins of the oracle functions outside the SMO solve, and of the library's
host-side prediction code, against the LIBSVM scikit-learn bundles
(committed vectors in tests/golden/sklearn_multiclass.npz, written by
tests/golden/make_sklearn_multiclass_golden.py):

  * multi-class svm_train (libsvm/svm.cpp:2246-2386): the support-vector
    set and class packing, the sv_coef layout and the rho order;
  * svm_predict_values (:2569-2643): one-vs-one decision values;
  * svm_predict_probability (:2644-2700): Platt sigmoid and pairwise
    coupling, on a model carrying sklearn's probA / probB;
  * the model-file text (:2709-2801, :2803-3044): a hand-written model file
    in LIBSVM's format loads and saves back byte for byte.
"""
import ctypes as C
import os

import numpy as np
import pytest

from libsvm_hip import ffi
from model_util import ArrayModel

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(GOLDEN, "sklearn_multiclass.npz"))
KERNELS = {"rbf": ffi.RBF, "poly": ffi.POLY, "linear": ffi.LINEAR}


def _param(kname):
    return ffi.make_param(kernel_type=KERNELS[kname], gamma=0.2, coef0=0.5, degree=2, C_=2.0, eps=1e-3,
                          shrinking=0, cuda_flag=0)


@pytest.mark.parametrize("kname", list(KERNELS))
def test_oracle_multiclass_train_matches_sklearn(ora, kname):
    prob = ffi.Problem.from_dense(G["X"], G["y"])
    m = ora.svm_train(prob, _param(kname)).arrays()
    np.testing.assert_array_equal(m["label"], [1, 2, 3, 4])
    np.testing.assert_array_equal(m["nSV"], G[f"{kname}_n_support"])
    np.testing.assert_array_equal(m["sv_indices"] - 1, G[f"{kname}_support"])
    np.testing.assert_allclose(m["sv_coef"], G[f"{kname}_dual_coef"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(m["rho"], -G[f"{kname}_intercept"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("kname", list(KERNELS))
def test_oracle_predict_values_match_sklearn(ora, kname):
    prob = ffi.Problem.from_dense(G["X"], G["y"])
    m = ora.svm_train(prob, _param(kname))
    test = ffi.Problem.from_dense(G["Xt"], np.zeros(len(G["Xt"])))
    lab, dec = m.predict_values(test)
    np.testing.assert_allclose(dec, G[f"{kname}_dec"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(lab, G[f"{kname}_pred"])


def _sklearn_model(kname_or_prob):
    pre = "prob" if kname_or_prob == "prob" else kname_or_prob
    kname = "rbf" if pre == "prob" else pre
    sup = G[f"{pre}_support"]
    extra = {}
    if pre == "prob":
        extra = dict(probA=G["prob_probA"], probB=G["prob_probB"])
    return ArrayModel(_param(kname), G["X"][sup], G[f"{pre}_dual_coef"], -G[f"{pre}_intercept"],
                      label=[1, 2, 3, 4], nSV=G[f"{pre}_n_support"], **extra)


@pytest.mark.parametrize("kname", list(KERNELS))
def test_library_host_predict_values_match_sklearn(kname):
    import libsvm_hip
    L = libsvm_hip.lib()
    am = _sklearn_model(kname)
    test = ffi.Problem.from_dense(G["Xt"], np.zeros(len(G["Xt"])))
    dec = np.zeros((test.l, 6))
    lab = np.zeros(test.l)
    for r in range(test.l):
        lab[r] = L.svm_predict_values(am.ptr, test.row_nodes(r), dec[r].ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_allclose(dec, G[f"{kname}_dec"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(lab, G[f"{kname}_pred"])


def test_library_predict_probability_matches_sklearn():
    import libsvm_hip
    L = libsvm_hip.lib()
    L.svm_predict_probability.argtypes = [C.POINTER(ffi.svm_model), C.POINTER(ffi.svm_node), C.POINTER(C.c_double)]
    L.svm_predict_probability.restype = C.c_double
    L.svm_check_probability_model.argtypes = [C.POINTER(ffi.svm_model)]
    am = _sklearn_model("prob")
    assert L.svm_check_probability_model(am.ptr) == 1
    test = ffi.Problem.from_dense(G["Xt"], np.zeros(len(G["Xt"])))
    P = np.zeros((test.l, 4))
    lab = np.zeros(test.l)
    for r in range(test.l):
        lab[r] = L.svm_predict_probability(am.ptr, test.row_nodes(r), P[r].ctypes.data_as(C.POINTER(C.c_double)))
    np.testing.assert_allclose(P, G["prob_proba"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(lab, 1 + np.argmax(P, axis=1))
    assert np.allclose(P.sum(axis=1), 1.0)


def test_model_file_text_round_trip(tmp_path):
    """tests/golden/poly_3class.model is written by hand in LIBSVM 3.18's
    format (keys in svm_save_model's order; rho, gamma, coef0, probA/probB
    printed with %g; coefficients with %.16g; feature values with %.8g):
    svm_load_model then svm_save_model reproduces it byte for byte, and the
    loaded numbers are the ones the text states."""
    import libsvm_hip
    L = libsvm_hip.lib()
    src = os.path.join(GOLDEN, "poly_3class.model")
    text = open(src).read()
    m = libsvm_hip.svm_load_model(src)
    mm = m.ptr.contents
    assert (mm.param.svm_type, mm.param.kernel_type, mm.param.degree) == (ffi.C_SVC, ffi.POLY, 3)
    assert mm.param.gamma == 0.125 and mm.param.coef0 == -1.5
    assert (mm.nr_class, mm.l) == (3, 5)
    np.testing.assert_array_equal([mm.rho[i] for i in range(3)], [0.123457, -2.5e-07, 3.0])
    np.testing.assert_array_equal([mm.probA[i] for i in range(3)], [-1.25, -0.5, -3.0])
    np.testing.assert_array_equal([mm.label[i] for i in range(3)], [7, -2, 5])
    np.testing.assert_array_equal([mm.nSV[i] for i in range(3)], [2, 2, 1])
    assert mm.sv_coef[0][1] == float("0.3333333333333333") and mm.sv_coef[1][4] == -1e-300
    assert mm.SV[2][0].index == 3 and mm.SV[2][0].value == float("0.12345679")
    out = str(tmp_path / "saved.model")
    m.save(out)
    assert open(out).read() == text
    # the printf formats themselves, on values whose renderings differ by format
    assert "%g %.16g %.8g" % (1 / 3, 1 / 3, 1 / 3) == "0.333333 0.3333333333333333 0.33333333"
    assert L.svm_check_probability_model(m.ptr) == 1
