"""Python binding of libsvm_hip.so, the MI355X SMO solver.

Two layers, both straight ctypes over the C-ABI declared in include/:
  * the kept LIBSVM API (include/svm.h): svm_train, svm_predict_values,
    svm_save_model / svm_load_model, svm_cross_validation, ...
  * the device-solver handle (include/svm_hip.h): DeviceSolver mirrors the
    reference's CudaSolver methods (libsvm/cuda_solver.h:224-244) and adds
    the fused on-device solve.

The library must be built (`make -C libsvm-cuda_amd`); importing this module
without it raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import ffi
from .ffi import (C_SVC, EPSILON_SVR, LINEAR, NU_SVC, NU_SVR, ONE_CLASS, POLY,  # noqa: F401
                  PRECOMPUTED, RBF, SIGMOID, Problem, make_param, model_arrays)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsvm_hip.so")

KIND_C, KIND_NU = 0, 1
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEVICE", 3: "ERR_NOMEM", 4: "ERR_STATE",
                5: "ERR_TIMEOUT", 6: "ERR_MAXITER"}


class svmhip_options(C.Structure):
    _fields_ = [("struct_size", C.c_size_t), ("device", C.c_int), ("grid_blocks", C.c_int),
                ("block_threads", C.c_int), ("max_iter", C.c_int), ("cache_bytes", C.c_double),
                ("iters_per_launch", C.c_int), ("verbose", C.c_int), ("speculative_rows", C.c_int),
                ("mfma_rows", C.c_int), ("exchange_copies", C.c_int), ("hbm_fraction", C.c_double)]


class svmhip_stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("optimal", C.c_int), ("obj", C.c_double), ("rho", C.c_double),
                ("r", C.c_double), ("row_hits", C.c_longlong), ("row_misses", C.c_longlong),
                ("solve_ms", C.c_double), ("kernel_ms", C.c_double), ("launches", C.c_int),
                ("setup_ms", C.c_double), ("cache_slots", C.c_int),
                ("grid_blocks", C.c_int), ("gram_ms", C.c_double), ("device_allocs", C.c_longlong),
                ("resident_vectors", C.c_int), ("shared_dots", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class svmhip_subproblem(C.Structure):
    _fields_ = [("rowset", C.c_int), ("kind", C.c_int), ("n_var", C.c_int),
                ("y", C.POINTER(C.c_byte)), ("p", C.POINTER(C.c_double)),
                ("alpha", C.POINTER(C.c_double)), ("Cp", C.c_double), ("Cn", C.c_double),
                ("gamma", C.c_double), ("eps", C.c_double), ("rho", C.c_double),
                ("obj", C.c_double), ("r", C.c_double), ("iters", C.c_int), ("status", C.c_int)]


class svmhip_completion(C.Structure):
    _fields_ = [("id", C.c_int), ("type", C.c_int), ("status", C.c_int), ("n", C.c_int),
                ("iters", C.c_int), ("nSV", C.c_int), ("rho", C.c_double), ("obj", C.c_double),
                ("r", C.c_double), ("device_ms", C.c_double)]


# every extern "C" symbol declared in include/svm.h and include/svm_hip.h
SVM_H_SYMBOLS = [
    "svm_train", "svm_cross_validation", "svm_save_model", "svm_load_model", "svm_get_svm_type",
    "svm_get_nr_class", "svm_get_labels", "svm_get_sv_indices", "svm_get_nr_sv",
    "svm_get_svr_probability", "svm_predict_values", "svm_predict", "svm_predict_probability",
    "svm_free_model_content", "svm_free_and_destroy_model", "svm_destroy_param",
    "svm_check_parameter", "svm_check_probability_model", "svm_set_print_string_function",
    "libsvm_version"]
SVM_HIP_SYMBOLS = [
    "svmhip_create", "svmhip_setup_rbf", "svmhip_setup", "svmhip_select_working_set",
    "svmhip_compute_alpha", "svmhip_update_gradient", "svmhip_update_alpha_status", "svmhip_fetch",
    "svmhip_solve", "svmhip_finalize", "svmhip_get_stats", "svmhip_kernel_row", "svmhip_destroy",
    "svmhip_last_error", "svmhip_device_count", "svmhip_batch_create", "svmhip_batch_add_rowset",
    "svmhip_batch_solve", "svmhip_batch_destroy", "svmhip_batch_last_error", "svmhip_predict_batch",
    "svmhip_predict_batch_noalloc",
    "svmhip_server_create", "svmhip_server_submit_train", "svmhip_server_submit_predict",
    "svmhip_server_step", "svmhip_server_poll", "svmhip_server_train_result",
    "svmhip_server_predict_result", "svmhip_server_release", "svmhip_server_bytes_in_use",
    "svmhip_server_stats", "svmhip_server_destroy", "svmhip_server_last_error"]

_lib = None
PRINT_FUNC = C.CFUNCTYPE(None, C.c_char_p)
_quiet_cb = PRINT_FUNC(lambda s: None)


def lib():
    """Load libsvm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libsvm_hip.so not built: {LIB_PATH} (run `make -C libsvm-cuda_amd`)")
    L = C.CDLL(LIB_PATH)
    P, D, I = C.POINTER, C.c_double, C.c_int
    h = C.c_void_p
    L.svm_train.argtypes = [P(ffi.svm_problem), P(ffi.svm_parameter)]
    L.svm_train.restype = P(ffi.svm_model)
    L.svm_cross_validation.argtypes = [P(ffi.svm_problem), P(ffi.svm_parameter), I, P(D)]
    L.svm_save_model.argtypes = [C.c_char_p, P(ffi.svm_model)]
    L.svm_load_model.argtypes = [C.c_char_p]
    L.svm_load_model.restype = P(ffi.svm_model)
    L.svm_predict_values.argtypes = [P(ffi.svm_model), P(ffi.svm_node), P(D)]
    L.svm_predict_values.restype = D
    L.svm_predict.argtypes = [P(ffi.svm_model), P(ffi.svm_node)]
    L.svm_predict.restype = D
    L.svm_free_and_destroy_model.argtypes = [P(P(ffi.svm_model))]
    L.svm_check_parameter.argtypes = [P(ffi.svm_problem), P(ffi.svm_parameter)]
    L.svm_check_parameter.restype = C.c_char_p
    L.svm_set_print_string_function.argtypes = [PRINT_FUNC]
    L.svm_get_nr_class.argtypes = [P(ffi.svm_model)]
    L.svm_get_nr_sv.argtypes = [P(ffi.svm_model)]
    L.svmhip_create.argtypes = [P(ffi.svm_problem), P(ffi.svm_parameter), I, P(svmhip_options),
                                P(h)]
    L.svmhip_setup_rbf.argtypes = [h, I]
    L.svmhip_setup.argtypes = [h, P(C.c_byte), P(D), P(D), P(C.c_char), D, D, I]
    L.svmhip_select_working_set.argtypes = [h, I, P(I), P(I), P(I)]
    L.svmhip_compute_alpha.argtypes = [h]
    L.svmhip_update_gradient.argtypes = [h, I]
    L.svmhip_update_alpha_status.argtypes = [h]
    L.svmhip_fetch.argtypes = [h, P(D), P(D), P(C.c_char), I]
    L.svmhip_solve.argtypes = [h, I, P(svmhip_stats)]
    L.svmhip_finalize.argtypes = [h, P(svmhip_stats)]
    L.svmhip_get_stats.argtypes = [h, P(svmhip_stats)]
    L.svmhip_kernel_row.argtypes = [h, I, P(C.c_float)]
    L.svmhip_destroy.argtypes = [h]
    L.svmhip_destroy.restype = None
    L.svmhip_last_error.argtypes = [h]
    L.svmhip_last_error.restype = C.c_char_p
    L.svmhip_device_count.restype = I
    L.svmhip_batch_create.argtypes = [P(ffi.svm_problem), P(ffi.svm_parameter), P(svmhip_options), P(h)]
    L.svmhip_batch_create.restype = I
    L.svmhip_batch_add_rowset.argtypes = [h, P(I), I]
    L.svmhip_batch_add_rowset.restype = I
    L.svmhip_batch_solve.argtypes = [h, P(svmhip_subproblem), I, P(svmhip_stats)]
    L.svmhip_batch_solve.restype = I
    L.svmhip_batch_destroy.argtypes = [h]
    L.svmhip_batch_destroy.restype = None
    L.svmhip_batch_last_error.argtypes = [h]
    L.svmhip_batch_last_error.restype = C.c_char_p
    L.svmhip_predict_batch.argtypes = [P(ffi.svm_model), P(P(ffi.svm_node)), I, P(svmhip_options), P(D), P(D)]
    L.svmhip_predict_batch.restype = I
    L.svmhip_predict_batch_noalloc.argtypes = [P(ffi.svm_model), P(P(ffi.svm_node)), I, P(svmhip_options), P(D),
                                               P(D)]
    L.svmhip_predict_batch_noalloc.restype = I
    L.svmhip_server_create.argtypes = [P(svmhip_options), D, P(h)]
    L.svmhip_server_create.restype = I
    L.svmhip_server_submit_train.argtypes = [h, P(ffi.svm_problem), P(ffi.svm_parameter)]
    L.svmhip_server_submit_train.restype = I
    L.svmhip_server_submit_predict.argtypes = [h, I, P(P(ffi.svm_node)), I]
    L.svmhip_server_submit_predict.restype = I
    L.svmhip_server_step.argtypes = [h, D, P(I), P(I)]
    L.svmhip_server_step.restype = I
    L.svmhip_server_poll.argtypes = [h, P(svmhip_completion), I]
    L.svmhip_server_poll.restype = I
    L.svmhip_server_train_result.argtypes = [h, I, P(D), I, P(D)]
    L.svmhip_server_train_result.restype = I
    L.svmhip_server_predict_result.argtypes = [h, I, P(D), P(D), I]
    L.svmhip_server_predict_result.restype = I
    L.svmhip_server_release.argtypes = [h, I]
    L.svmhip_server_release.restype = I
    L.svmhip_server_bytes_in_use.argtypes = [h]
    L.svmhip_server_bytes_in_use.restype = D
    L.svmhip_server_stats.argtypes = [h, P(svmhip_stats)]
    L.svmhip_server_stats.restype = I
    L.svmhip_server_destroy.argtypes = [h]
    L.svmhip_server_destroy.restype = None
    L.svmhip_server_last_error.argtypes = [h]
    L.svmhip_server_last_error.restype = C.c_char_p
    for name in ("svmhip_create", "svmhip_setup_rbf", "svmhip_setup", "svmhip_select_working_set",
                 "svmhip_compute_alpha", "svmhip_update_gradient", "svmhip_update_alpha_status",
                 "svmhip_fetch", "svmhip_solve", "svmhip_finalize", "svmhip_get_stats",
                 "svmhip_kernel_row"):
        getattr(L, name).restype = I
    _lib = L
    return L


def set_quiet(quiet=True):
    """Route LIBSVM's info() output to nowhere (quiet) or stdout."""
    lib().svm_set_print_string_function(_quiet_cb if quiet else PRINT_FUNC())


def device_count():
    return lib().svmhip_device_count()


class SolverError(RuntimeError):
    pass


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class DeviceSolver:
    """One device solve: a Python view of the svmhip_* handle.

    Method names follow the reference's CudaSolver (libsvm/cuda_solver.h:224-244)."""

    def __init__(self, problem: Problem, param, kind=KIND_C, grid_blocks=0, max_iter=0,
                 cache_bytes=0.0, iters_per_launch=0, device=0, verbose=0, speculative_rows=0, mfma_rows=0,
                 exchange_copies=0):
        L = lib()
        self.problem = problem  # keeps the node buffers alive
        self.param = param
        opt = svmhip_options(C.sizeof(svmhip_options), device, grid_blocks, 0, max_iter,
                             float(cache_bytes), iters_per_launch, verbose, speculative_rows, mfma_rows,
                             exchange_copies)
        self._h = C.c_void_p()
        rc = L.svmhip_create(C.byref(problem.struct), C.byref(param), kind, C.byref(opt),
                             C.byref(self._h))
        if rc != 0:
            raise SolverError(f"svmhip_create failed: {STATUS_NAMES.get(rc, rc)}")
        self.n_var = 0

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().svmhip_last_error(self._h).decode()
            raise SolverError(f"{what}: {STATUS_NAMES.get(rc, rc)}: {msg}")

    def setup_rbf_variables(self):
        self._check(lib().svmhip_setup_rbf(self._h, self.problem.l), "setup_rbf")

    def setup_solver(self, y, G, alpha, status, Cp, Cn):
        y = np.ascontiguousarray(y, dtype=np.int8)
        G = np.ascontiguousarray(G, dtype=np.float64)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        status = np.ascontiguousarray(status, dtype=np.int8)
        self.n_var = len(y)
        self._check(lib().svmhip_setup(self._h, _ptr(y, C.c_byte), _ptr(G, C.c_double),
                                       _ptr(alpha, C.c_double), _ptr(status, C.c_char),
                                       float(Cp), float(Cn), self.n_var), "setup")

    def select_working_set(self, want_pair=False):
        opt, i, j = C.c_int(), C.c_int(-1), C.c_int(-1)
        if want_pair:
            rc = lib().svmhip_select_working_set(self._h, self.n_var, C.byref(opt), C.byref(i),
                                                 C.byref(j))
        else:
            rc = lib().svmhip_select_working_set(self._h, self.n_var, C.byref(opt), None, None)
        self._check(rc, "select_working_set")
        return opt.value, i.value, j.value

    def compute_alpha(self):
        self._check(lib().svmhip_compute_alpha(self._h), "compute_alpha")

    def update_gradient(self):
        self._check(lib().svmhip_update_gradient(self._h, self.n_var), "update_gradient")

    def update_alpha_status(self):
        self._check(lib().svmhip_update_alpha_status(self._h), "update_alpha_status")

    def fetch_vectors(self):
        G = np.zeros(self.n_var)
        a = np.zeros(self.n_var)
        st = np.zeros(self.n_var, dtype=np.int8)
        self._check(lib().svmhip_fetch(self._h, _ptr(G, C.c_double), _ptr(a, C.c_double),
                                       _ptr(st, C.c_char), self.n_var), "fetch")
        return G, a, st

    def solve(self, max_iter=0):
        s = svmhip_stats()
        self._check(lib().svmhip_solve(self._h, max_iter, C.byref(s)), "solve")
        return s.as_dict()

    def finalize(self):
        s = svmhip_stats()
        self._check(lib().svmhip_finalize(self._h, C.byref(s)), "finalize")
        return s.as_dict()

    def kernel_row(self, i):
        out = np.zeros(self.problem.l, dtype=np.float32)
        self._check(lib().svmhip_kernel_row(self._h, int(i), _ptr(out, C.c_float)), "kernel_row")
        return out

    def close(self):
        if self._h:
            lib().svmhip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def formulate(problem: Problem, param, Cp=None, Cn=None):
    """Initial (y, p, alpha, Cp, Cn, kind) of one subproblem, as the
    formulations of libsvm/svm.cpp:1500-1696 build them."""
    l = problem.l
    yl = problem.y
    st = param.svm_type
    if st == C_SVC:
        y = np.where(yl > 0, 1, -1).astype(np.int8)
        return y, -np.ones(l), np.zeros(l), (param.C if Cp is None else Cp), \
            (param.C if Cn is None else Cn), KIND_C
    if st == NU_SVC:
        y = np.where(yl > 0, 1, -1).astype(np.int8)
        a = np.zeros(l)
        pos = neg = param.nu * l / 2
        for i in range(l):
            if y[i] == 1:
                a[i] = min(1.0, pos)
                pos -= a[i]
            else:
                a[i] = min(1.0, neg)
                neg -= a[i]
        return y, np.zeros(l), a, 1.0, 1.0, KIND_NU
    if st == ONE_CLASS:
        n = int(param.nu * l)
        a = np.zeros(l)
        a[:n] = 1
        if n < l:
            a[n] = param.nu * l - n
        return np.ones(l, dtype=np.int8), np.zeros(l), a, 1.0, 1.0, KIND_C
    if st == EPSILON_SVR:
        y = np.concatenate([np.ones(l), -np.ones(l)]).astype(np.int8)
        p = np.concatenate([param.p - yl, param.p + yl])
        return y, p, np.zeros(2 * l), param.C, param.C, KIND_C
    if st == NU_SVR:
        C_ = param.C
        a = np.zeros(2 * l)
        s = C_ * param.nu * l / 2
        for i in range(l):
            a[i] = a[i + l] = min(s, C_)
            s -= a[i]
        y = np.concatenate([np.ones(l), -np.ones(l)]).astype(np.int8)
        p = np.concatenate([-yl, yl])
        return y, p, a, C_, C_, KIND_NU
    raise ValueError(st)


def initial_status(y, alpha, Cp, Cn):
    C_ = np.where(y > 0, Cp, Cn)
    return np.where(alpha >= C_, 1, np.where(alpha <= 0, 0, 2)).astype(np.int8)


def solve_one(problem: Problem, param, Cp=None, Cn=None, step_wise=False, max_iter=0, **opts):
    """Train one subproblem on the device; returns the raw solver result
    (alpha_raw, G, status, stats) plus the signed coefficients `alpha` in the
    convention of svm_train_one (libsvm/svm.cpp:1707-1769)."""
    y, p, a0, Cp_, Cn_, kind = formulate(problem, param, Cp, Cn)
    s = DeviceSolver(problem, param, kind=kind, max_iter=max_iter, **opts)
    s.setup_rbf_variables()
    s.setup_solver(y, p, a0, initial_status(y, a0, Cp_, Cn_), Cp_, Cn_)
    if step_wise:
        limit = max_iter if max_iter > 0 else max(10_000_000, 100 * problem.l)
        it = 0
        while it < limit:
            opt, _, _ = s.select_working_set()
            if opt:
                break
            it += 1
            s.compute_alpha()
            s.update_gradient()
            s.update_alpha_status()
        stats = s.finalize()
    else:
        stats = s.solve(max_iter)
    G, a, st = s.fetch_vectors()
    s.close()
    l = problem.l
    rho, obj, r = stats["rho"], stats["obj"], stats["r"]
    if param.svm_type == C_SVC:
        coef = a * y
    elif param.svm_type == NU_SVC:
        coef = a * y / r
        rho, obj = rho / r, obj / (r * r)
    elif param.svm_type == ONE_CLASS:
        coef = a.copy()
    else:
        coef = a[:l] - a[l:]
    return dict(alpha=coef, alpha_raw=a, G=G, status=st, rho=rho, obj=obj, r=r,
                iters=stats["iters"], stats=stats)


class Model:
    """Owns a struct svm_model* returned by svm_train / svm_load_model."""

    def __init__(self, ptr, problem=None):
        if not ptr:
            raise SolverError("svm_train returned no model (device failure)")
        self.ptr = ptr
        self.problem = problem

    def arrays(self):
        return model_arrays(self.ptr)

    def predict_values(self, problem: Problem, rows=None):
        L = lib()
        m = self.ptr.contents
        k = m.nr_class
        ndec = 1 if m.param.svm_type in (ONE_CLASS, EPSILON_SVR, NU_SVR) else k * (k - 1) // 2
        rows = range(problem.l) if rows is None else rows
        dec = np.zeros((len(rows), max(ndec, 1)))
        lab = np.zeros(len(rows))
        buf = np.zeros(max(ndec, 1))
        for n, r in enumerate(rows):
            lab[n] = L.svm_predict_values(self.ptr, problem.row_nodes(r), _ptr(buf, C.c_double))
            dec[n] = buf
        return lab, dec

    def predict_batch(self, problem: Problem, rows=None, noalloc=False, mfma=False):
        """Decision values and labels of many rows on the device
        (svmhip_predict_batch, or svmhip_predict_batch_noalloc)."""
        return predict_batch(self.ptr, problem, rows, noalloc, mfma)

    def save(self, path):
        if lib().svm_save_model(path.encode(), self.ptr) != 0:
            raise SolverError(f"svm_save_model({path}) failed")

    def __del__(self):
        try:
            if self.ptr:
                lib().svm_free_and_destroy_model(C.byref(self.ptr))
        except Exception:
            pass


def predict_batch(model_ptr, problem: Problem, rows=None, noalloc=False, mfma=False):
    """Decision values and labels of problem's rows (all, or `rows`) against
    the struct svm_model* model_ptr, on the device.  mfma: the opt-in Gram-
    tile path (mfma_rows = 1) for models with one coefficient per support
    vector."""
    m = model_ptr.contents
    k = m.nr_class
    ndec = 1 if m.param.svm_type in (ONE_CLASS, EPSILON_SVR, NU_SVR) else k * (k - 1) // 2
    rows = np.arange(problem.l) if rows is None else np.asarray(rows)
    ptrs = (C.POINTER(ffi.svm_node) * max(len(rows), 1))()
    for n, r in enumerate(rows):
        ptrs[n] = problem.row_nodes(int(r))
    dec = np.zeros((len(rows), ndec))
    lab = np.zeros(len(rows))
    opt = svmhip_options(C.sizeof(svmhip_options), 0, 0, 0, 0, 0.0, 0, 0)
    opt.mfma_rows = 1 if mfma else 0
    fn = lib().svmhip_predict_batch_noalloc if noalloc else lib().svmhip_predict_batch
    rc = fn(model_ptr, ptrs, len(rows), C.byref(opt), _ptr(dec, C.c_double), _ptr(lab, C.c_double))
    if rc != 0:
        raise SolverError(f"{'svmhip_predict_batch_noalloc' if noalloc else 'svmhip_predict_batch'}: "
                          f"{STATUS_NAMES.get(rc, rc)}")
    return lab, dec


def svm_train(problem: Problem, param):
    return Model(lib().svm_train(C.byref(problem.struct), C.byref(param)), problem)


def svm_load_model(path):
    ptr = lib().svm_load_model(path.encode())
    if not ptr:
        raise SolverError(f"svm_load_model({path}) failed")
    return Model(ptr)


def check_parameter(problem: Problem, param):
    r = lib().svm_check_parameter(C.byref(problem.struct), C.byref(param))
    return None if r is None else r.decode()


class BatchSolver:
    """Many subproblems over row subsets of one uploaded problem, solved in
    one launch (one workgroup each): the svmhip_batch_* C-ABI."""

    def __init__(self, problem: Problem, param, cache_bytes=0.0, max_iter=0, device=0, team_max=0, chunk=0,
                 verbose=0, speculative_rows=0):
        """team_max: largest team of workgroups per subproblem (0 = default,
        1 = one workgroup each); chunk: iterations per subproblem per launch
        (0 = sized by the scheduler); speculative_rows: -1 turns off the
        speculative kernel rows of single-workgroup solves."""
        L = lib()
        self.problem = problem
        self.param = param
        opt = svmhip_options(C.sizeof(svmhip_options), device, team_max, 0, max_iter, float(cache_bytes), chunk,
                             verbose, speculative_rows)
        self._h = C.c_void_p()
        rc = L.svmhip_batch_create(C.byref(problem.struct), C.byref(param), C.byref(opt), C.byref(self._h))
        if rc != 0:
            raise SolverError(f"svmhip_batch_create failed: {STATUS_NAMES.get(rc, rc)}")

    def add_rowset(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        rid = lib().svmhip_batch_add_rowset(self._h, _ptr(rows, C.c_int), len(rows))
        if rid < 0:
            raise SolverError(f"add_rowset: {STATUS_NAMES.get(-rid, -rid)}: "
                              f"{lib().svmhip_batch_last_error(self._h).decode()}")
        return rid

    def solve(self, subs):
        """subs: list of dicts with rowset, kind, y, p, alpha0, Cp, Cn, gamma, eps.
        Returns (list of result dicts with alpha/rho/obj/r/iters/status, stats)."""
        n = len(subs)
        arr = (svmhip_subproblem * max(n, 1))()
        keep = []
        for k, sp in enumerate(subs):
            y = np.ascontiguousarray(sp["y"], dtype=np.int8)
            p = np.ascontiguousarray(sp["p"], dtype=np.float64)
            a = np.array(sp["alpha0"], dtype=np.float64)
            keep += [y, p, a]
            arr[k] = svmhip_subproblem(sp["rowset"], sp["kind"], len(y), _ptr(y, C.c_byte),
                                       _ptr(p, C.c_double), _ptr(a, C.c_double), sp["Cp"], sp["Cn"],
                                       sp["gamma"], sp["eps"], 0.0, 0.0, 0.0, 0, 0)
        st = svmhip_stats()
        rc = lib().svmhip_batch_solve(self._h, arr, n, C.byref(st))
        if rc != 0:
            raise SolverError(f"batch_solve: {STATUS_NAMES.get(rc, rc)}: "
                              f"{lib().svmhip_batch_last_error(self._h).decode()}")
        out = []
        for k in range(n):
            out.append(dict(alpha_raw=keep[3 * k + 2], rho=arr[k].rho, obj=arr[k].obj, r=arr[k].r,
                            iters=arr[k].iters, status=arr[k].status))
        return out, st.as_dict()

    def close(self):
        if self._h:
            lib().svmhip_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def finish_solution(problem_y, param, res, y, r=None):
    """Signed coefficients, rho and obj from a raw solver result, as the
    formulations of libsvm/svm.cpp:1500-1696 post-process them."""
    a = res["alpha_raw"]
    rho, obj, rr = res["rho"], res["obj"], res["r"]
    st = param.svm_type
    l = len(problem_y)
    if st == C_SVC:
        coef = a * y
    elif st == NU_SVC:
        coef = a * y / rr
        rho, obj = rho / rr, obj / (rr * rr)
    elif st == ONE_CLASS:
        coef = a.copy()
    else:
        coef = a[:l] - a[l:]
    return coef, rho, obj


def svm_cross_validation(problem: Problem, param, nr_fold):
    """The kept svm_cross_validation (rand()-drawn stratified folds)."""
    target = np.zeros(problem.l)
    lib().svm_cross_validation(C.byref(problem.struct), C.byref(param), nr_fold, _ptr(target, C.c_double))
    return target


class Server:
    """Device-resident request server (the svmhip_server_* C-ABI): binary
    C-SVC / nu-SVC training solves and predictions against their resident
    models, admitted by HBM budget and advanced together one stage per step."""

    def __init__(self, budget_bytes=0.0, cache_bytes=0.0, team_max=0, max_iter=0, device=0, speculative_rows=0):
        L = lib()
        opt = svmhip_options(C.sizeof(svmhip_options), device, team_max, 0, max_iter, float(cache_bytes), 0, 0,
                             speculative_rows)
        self._h = C.c_void_p()
        rc = L.svmhip_server_create(C.byref(opt), float(budget_bytes), C.byref(self._h))
        if rc != 0:
            raise SolverError(f"svmhip_server_create failed: {STATUS_NAMES.get(rc, rc)}")
        self._keep = {}

    def _err(self):
        return lib().svmhip_server_last_error(self._h).decode()

    def submit_train(self, problem: Problem, param):
        rid = lib().svmhip_server_submit_train(self._h, C.byref(problem.struct), C.byref(param))
        if rid < 0:
            raise SolverError(f"submit_train: {STATUS_NAMES.get(-rid, -rid)}: {self._err()}")
        return rid

    def submit_predict(self, model_id, problem: Problem):
        rid = lib().svmhip_server_submit_predict(self._h, model_id, problem.struct.x, problem.l)
        if rid < 0:
            raise SolverError(f"submit_predict: {STATUS_NAMES.get(-rid, -rid)}: {self._err()}")
        return rid

    def step(self, stage_ms=10.0):
        a, q = C.c_int(), C.c_int()
        rc = lib().svmhip_server_step(self._h, float(stage_ms), C.byref(a), C.byref(q))
        if rc != 0:
            raise SolverError(f"server step: {STATUS_NAMES.get(rc, rc)}: {self._err()}")
        return a.value, q.value

    def poll(self, max_n=4096):
        buf = (svmhip_completion * max_n)()
        n = lib().svmhip_server_poll(self._h, buf, max_n)
        return [dict(id=c.id, type=c.type, status=c.status, n=c.n, iters=c.iters, nSV=c.nSV, rho=c.rho,
                     obj=c.obj, r=c.r, device_ms=c.device_ms) for c in buf[:n]]

    def train_result(self, rid, n):
        coef = np.zeros(n)
        lab = np.zeros(2)
        rc = lib().svmhip_server_train_result(self._h, rid, _ptr(coef, C.c_double), n, _ptr(lab, C.c_double))
        if rc != 0:
            raise SolverError(f"train_result: {STATUS_NAMES.get(rc, rc)}: {self._err()}")
        return coef, lab

    def predict_result(self, rid, n):
        dec = np.zeros(n)
        lab = np.zeros(n)
        rc = lib().svmhip_server_predict_result(self._h, rid, _ptr(dec, C.c_double), _ptr(lab, C.c_double), n)
        if rc != 0:
            raise SolverError(f"predict_result: {STATUS_NAMES.get(rc, rc)}: {self._err()}")
        return dec, lab

    def release(self, rid):
        rc = lib().svmhip_server_release(self._h, rid)
        if rc != 0:
            raise SolverError(f"release: {STATUS_NAMES.get(rc, rc)}: {self._err()}")

    def bytes_in_use(self):
        return lib().svmhip_server_bytes_in_use(self._h)

    def stats(self):
        st = svmhip_stats()
        lib().svmhip_server_stats(self._h, C.byref(st))
        return st.as_dict()

    def run_until_idle(self, stage_ms=10.0, max_steps=100000):
        """Steps until nothing is active or queued; returns all completions."""
        out = []
        for _ in range(max_steps):
            a, q = self.step(stage_ms)
            out += self.poll()
            if a == 0 and q == 0:
                return out
        raise SolverError("server did not drain")

    def close(self):
        if self._h:
            lib().svmhip_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
