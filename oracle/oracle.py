"""ORACLE — test infrastructure only.

ctypes binding of oracle/_build/liboracle_svm.so, the FP64 CPU restatement
of LIBSVM 3.18 (oracle_svm.cpp).  Imported only by tests/, the smoke check
in __graft_entry__.py and bench.py's CPU-baseline leg, always as the
checker or the baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "libsvm-cuda_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from libsvm_hip import ffi  # noqa: E402

LIB_PATH = os.path.join(_HERE, "_build", "liboracle_svm.so")


class ora_options(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("q_double", C.c_int), ("trace_cap", C.c_int),
                ("trace_ij", C.POINTER(C.c_int))]


class ora_result(C.Structure):
    _fields_ = [("l", C.c_int), ("n_var", C.c_int), ("alpha", C.POINTER(C.c_double)),
                ("alpha_raw", C.POINTER(C.c_double)), ("G", C.POINTER(C.c_double)),
                ("rho", C.c_double), ("obj", C.c_double), ("r", C.c_double),
                ("upper_bound_p", C.c_double), ("upper_bound_n", C.c_double),
                ("iters", C.c_int), ("trace_len", C.c_int), ("nSV", C.c_int), ("nBSV", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.ora_set_options.argtypes = [C.POINTER(ora_options)]
        L.ora_set_verbose.argtypes = [C.c_int]
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_train_one.argtypes = [C.POINTER(ffi.svm_problem), C.POINTER(ffi.svm_parameter),
                                    C.c_double, C.c_double, C.POINTER(ora_result)]
        L.ora_kernel_row.argtypes = [C.POINTER(ffi.svm_problem), C.POINTER(ffi.svm_parameter),
                                     C.c_int, C.POINTER(C.c_float)]
        L.ora_kernel_row_f64.argtypes = [C.POINTER(ffi.svm_problem), C.POINTER(ffi.svm_parameter),
                                         C.c_int, C.POINTER(C.c_double)]
        L.ora_k_function.argtypes = [C.POINTER(ffi.svm_node), C.POINTER(ffi.svm_node),
                                     C.POINTER(ffi.svm_parameter)]
        L.ora_k_function.restype = C.c_double
        L.ora_svm_train.argtypes = [C.POINTER(ffi.svm_problem), C.POINTER(ffi.svm_parameter)]
        L.ora_svm_train.restype = C.POINTER(ffi.svm_model)
        L.ora_svm_predict_values.argtypes = [C.POINTER(ffi.svm_model), C.POINTER(ffi.svm_node),
                                             C.POINTER(C.c_double)]
        L.ora_svm_predict_values.restype = C.c_double
        L.ora_svm_predict.argtypes = [C.POINTER(ffi.svm_model), C.POINTER(ffi.svm_node)]
        L.ora_svm_predict.restype = C.c_double
        L.ora_cross_validation_folds.argtypes = [C.POINTER(ffi.svm_problem), C.POINTER(ffi.svm_parameter),
                                                 C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ora_free_model.argtypes = [C.POINTER(ffi.svm_model)]
        L.ora_svm_cross_validation.argtypes = [C.POINTER(ffi.svm_problem), C.POINTER(ffi.svm_parameter),
                                               C.c_int, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def set_threads(n):
    """Threads that fill a missing kernel row (default 1, LIBSVM's one
    thread); results do not depend on it."""
    lib().ora_set_threads(int(n))


def train_one(problem: ffi.Problem, param, Cp=None, Cn=None, max_iter=0, q_double=False,
              trace=0):
    """One svm_train_one solve (libsvm/svm.cpp:1707-1769).  Returns a dict
    with alpha (signed, length l), alpha_raw / G (solver variables), rho,
    obj, r, iters, nSV, nBSV and, if trace > 0, the first `trace` (i, j)
    working-set pairs."""
    L = lib()
    Cp = param.C if Cp is None else Cp
    Cn = param.C if Cn is None else Cn
    l = problem.l
    nvar = 2 * l if param.svm_type in (ffi.EPSILON_SVR, ffi.NU_SVR) else l
    alpha = np.zeros(l)
    alpha_raw = np.zeros(nvar)
    G = np.zeros(nvar)
    tr = np.zeros(2 * max(trace, 1), dtype=np.int32)
    opt = ora_options(max_iter, int(q_double), trace,
                      tr.ctypes.data_as(C.POINTER(C.c_int)) if trace else None)
    L.ora_set_options(C.byref(opt))
    res = ora_result()
    res.alpha = alpha.ctypes.data_as(C.POINTER(C.c_double))
    res.alpha_raw = alpha_raw.ctypes.data_as(C.POINTER(C.c_double))
    res.G = G.ctypes.data_as(C.POINTER(C.c_double))
    L.ora_train_one(C.byref(problem.struct), C.byref(param), Cp, Cn, C.byref(res))
    L.ora_set_options(None)
    out = dict(alpha=alpha, alpha_raw=alpha_raw, G=G, rho=res.rho, obj=res.obj, r=res.r,
               iters=res.iters, nSV=res.nSV, nBSV=res.nBSV,
               upper_bound_p=res.upper_bound_p, upper_bound_n=res.upper_bound_n)
    if trace:
        out["trace"] = tr[: 2 * res.trace_len].reshape(-1, 2)
    return out


def kernel_row(problem: ffi.Problem, param, i, f64=False):
    L = lib()
    if f64:
        out = np.zeros(problem.l)
        L.ora_kernel_row_f64(C.byref(problem.struct), C.byref(param), i,
                             out.ctypes.data_as(C.POINTER(C.c_double)))
    else:
        out = np.zeros(problem.l, dtype=np.float32)
        L.ora_kernel_row(C.byref(problem.struct), C.byref(param), i,
                         out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


class Model:
    """Owns a struct svm_model* trained by the oracle."""

    def __init__(self, ptr, problem):
        self.ptr = ptr
        self.problem = problem  # SVs point into the problem's node buffer

    def arrays(self):
        return ffi.model_arrays(self.ptr)

    def predict_values(self, problem: ffi.Problem, rows=None):
        L = lib()
        m = self.ptr.contents
        k = m.nr_class
        st = m.param.svm_type
        ndec = 1 if st in (ffi.ONE_CLASS, ffi.EPSILON_SVR, ffi.NU_SVR) else k * (k - 1) // 2
        rows = range(problem.l) if rows is None else rows
        dec = np.zeros((len(rows), ndec))
        lab = np.zeros(len(rows))
        buf = np.zeros(max(ndec, 1))
        for n, r in enumerate(rows):
            lab[n] = L.ora_svm_predict_values(self.ptr, problem.row_nodes(r),
                                              buf.ctypes.data_as(C.POINTER(C.c_double)))
            dec[n] = buf[:ndec]
        return lab, dec

    def __del__(self):
        try:
            lib().ora_free_model(self.ptr)
        except Exception:
            pass


def svm_train(problem: ffi.Problem, param, max_iter=0, q_double=False):
    L = lib()
    L.ora_set_options(C.byref(ora_options(max_iter, int(q_double), 0, None)))
    ptr = L.ora_svm_train(C.byref(problem.struct), C.byref(param))
    L.ora_set_options(None)
    return Model(ptr, problem)


def cross_validation(problem: ffi.Problem, param, folds):
    L = lib()
    folds = np.ascontiguousarray(folds, dtype=np.int32)
    target = np.zeros(problem.l)
    L.ora_cross_validation_folds(C.byref(problem.struct), C.byref(param), int(folds.max()) + 1,
                                 folds.ctypes.data_as(C.POINTER(C.c_int)),
                                 target.ctypes.data_as(C.POINTER(C.c_double)))
    return target


def svm_cross_validation(problem: ffi.Problem, param, nr_fold):
    """rand()-drawn folds exactly as the reference (call srand first)."""
    L = lib()
    target = np.zeros(problem.l)
    L.ora_svm_cross_validation(C.byref(problem.struct), C.byref(param), nr_fold,
                               target.ctypes.data_as(C.POINTER(C.c_double)))
    return target
