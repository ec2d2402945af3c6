"""ctypes mirrors of the C structs in include/svm.h and numpy helpers that
build LIBSVM problems (struct svm_problem) from CSR or dense arrays.

The struct layouts restate libsvm/svm.h:31-92 of the reference; the
svm_node array is built in bulk with numpy (16-byte records: int32 index,
4 pad bytes, float64 value) so a 60k x 784 problem is assembled in well
under a second.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

C_SVC, NU_SVC, ONE_CLASS, EPSILON_SVR, NU_SVR = range(5)
LINEAR, POLY, RBF, SIGMOID, PRECOMPUTED = range(5)


class svm_node(C.Structure):
    _fields_ = [("index", C.c_int), ("value", C.c_double)]


class svm_problem(C.Structure):
    _fields_ = [("l", C.c_int), ("y", C.POINTER(C.c_double)),
                ("x", C.POINTER(C.POINTER(svm_node)))]


class svm_parameter(C.Structure):
    _fields_ = [("cuda_flag", C.c_int), ("svm_type", C.c_int), ("kernel_type", C.c_int),
                ("degree", C.c_int), ("gamma", C.c_double), ("coef0", C.c_double),
                ("cache_size", C.c_double), ("eps", C.c_double), ("C", C.c_double),
                ("nr_weight", C.c_int), ("weight_label", C.POINTER(C.c_int)),
                ("weight", C.POINTER(C.c_double)), ("nu", C.c_double), ("p", C.c_double),
                ("shrinking", C.c_int), ("probability", C.c_int)]


class svm_model(C.Structure):
    _fields_ = [("param", svm_parameter), ("nr_class", C.c_int), ("l", C.c_int),
                ("SV", C.POINTER(C.POINTER(svm_node))),
                ("sv_coef", C.POINTER(C.POINTER(C.c_double))),
                ("rho", C.POINTER(C.c_double)), ("probA", C.POINTER(C.c_double)),
                ("probB", C.POINTER(C.c_double)), ("sv_indices", C.POINTER(C.c_int)),
                ("label", C.POINTER(C.c_int)), ("nSV", C.POINTER(C.c_int)),
                ("free_sv", C.c_int)]


NODE_DTYPE = np.dtype({"names": ["index", "value"], "formats": ["<i4", "<f8"],
                       "offsets": [0, 8], "itemsize": 16})


def make_param(svm_type=C_SVC, kernel_type=RBF, degree=3, gamma=0.0, coef0=0.0,
               cache_size=100.0, eps=1e-3, C_=1.0, nu=0.5, p=0.1, shrinking=1,
               probability=0, cuda_flag=1, weights=None):
    """svm_parameter with svm-train's defaults (svm-train/svm-train.c:164-180)."""
    prm = svm_parameter()
    prm.cuda_flag = cuda_flag
    prm.svm_type = svm_type
    prm.kernel_type = kernel_type
    prm.degree = degree
    prm.gamma = gamma
    prm.coef0 = coef0
    prm.cache_size = cache_size
    prm.eps = eps
    prm.C = C_
    prm.nu = nu
    prm.p = p
    prm.shrinking = shrinking
    prm.probability = probability
    prm._keep = []
    if weights:
        labels = (C.c_int * len(weights))(*[int(k) for k in weights])
        vals = (C.c_double * len(weights))(*[float(v) for v in weights.values()])
        prm.nr_weight = len(weights)
        prm.weight_label = C.cast(labels, C.POINTER(C.c_int))
        prm.weight = C.cast(vals, C.POINTER(C.c_double))
        prm._keep += [labels, vals]
    else:
        prm.nr_weight = 0
    return prm


class Problem:
    """A struct svm_problem plus the numpy buffers that back it.

    Rows are taken from CSR arrays (indptr, indices 0-based, data); LIBSVM
    feature indices are 1-based, so index = column + 1.  Explicit zeros in
    `data` are kept as stored nodes (LIBSVM keeps whatever the file holds).
    """

    def __init__(self, indptr, indices, data, y, n_features=None):
        indptr = np.asarray(indptr, dtype=np.int64)
        indices = np.asarray(indices, dtype=np.int32)
        data = np.asarray(data, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        l = len(indptr) - 1
        if len(y) != l:
            raise ValueError("len(y) != number of rows")
        nnz = len(indices)
        nodes = np.empty(nnz + l, dtype=NODE_DTYPE)
        row_len = np.diff(indptr)
        # position of each row's first node in the node array (rows are
        # laid out back to back, each followed by its -1 terminator)
        starts = indptr[:-1] + np.arange(l, dtype=np.int64)
        term = starts + row_len
        mask = np.ones(nnz + l, dtype=bool)
        mask[term] = False
        nodes["index"][mask] = indices + 1
        nodes["value"][mask] = data
        nodes["index"][term] = -1
        nodes["value"][term] = 0.0
        self.nodes = nodes
        self.y = y
        self.indptr, self.indices, self.data = indptr, indices, data
        self.l = l
        self.n_features = int(n_features if n_features is not None
                              else (indices.max() + 1 if nnz else 0))
        base = nodes.ctypes.data
        self._rowptr = (C.POINTER(svm_node) * max(l, 1))()
        addrs = (base + starts * 16).astype(np.uint64)
        C.memmove(self._rowptr, addrs.ctypes.data, 8 * l)
        self.struct = svm_problem()
        self.struct.l = l
        self.struct.y = y.ctypes.data_as(C.POINTER(C.c_double))
        self.struct.x = C.cast(self._rowptr, C.POINTER(C.POINTER(svm_node)))

    @classmethod
    def from_dense(cls, X, y, keep_zeros=False):
        X = np.asarray(X, dtype=np.float64)
        l, d = X.shape
        if keep_zeros:
            indptr = np.arange(0, (l + 1) * d, d, dtype=np.int64)
            indices = np.tile(np.arange(d, dtype=np.int32), l)
            data = X.ravel()
        else:
            nzr, nzc = np.nonzero(X)
            indptr = np.zeros(l + 1, dtype=np.int64)
            np.cumsum(np.bincount(nzr, minlength=l), out=indptr[1:])
            indices = nzc.astype(np.int32)
            data = X[nzr, nzc]
        return cls(indptr, indices, data, y, n_features=d)

    def subset(self, rows, y=None):
        """Problem over a subset of rows (copies the node data)."""
        rows = np.asarray(rows, dtype=np.int64)
        lens = np.diff(self.indptr)[rows]
        indptr = np.zeros(len(rows) + 1, dtype=np.int64)
        np.cumsum(lens, out=indptr[1:])
        sel = np.concatenate([np.arange(self.indptr[r], self.indptr[r + 1]) for r in rows]) \
            if len(rows) else np.zeros(0, dtype=np.int64)
        return Problem(indptr, self.indices[sel], self.data[sel],
                       self.y[rows] if y is None else y, n_features=self.n_features)

    def row_nodes(self, r):
        """POINTER(svm_node) to row r (for svm_predict calls)."""
        return self._rowptr[r]

    def to_dense(self):
        X = np.zeros((self.l, self.n_features))
        rows = np.repeat(np.arange(self.l), np.diff(self.indptr))
        X[rows, self.indices] = self.data
        return X


def model_arrays(model_ptr):
    """Copy the interesting fields of a struct svm_model* into numpy."""
    m = model_ptr.contents
    k = m.nr_class
    npair = k * (k - 1) // 2 if m.param.svm_type in (C_SVC, NU_SVC) else 1
    out = {"nr_class": k, "l": m.l,
           "rho": np.ctypeslib.as_array(m.rho, shape=(npair,)).copy()}
    out["sv_coef"] = np.array([np.ctypeslib.as_array(m.sv_coef[c], shape=(m.l,)).copy()
                               for c in range(max(k - 1, 1))]) if m.l else np.zeros((max(k - 1, 1), 0))
    out["sv_indices"] = (np.ctypeslib.as_array(m.sv_indices, shape=(m.l,)).copy()
                         if m.l and m.sv_indices else np.zeros(0, dtype=np.int32))
    if m.label:
        out["label"] = np.ctypeslib.as_array(m.label, shape=(k,)).copy()
    if m.nSV:
        out["nSV"] = np.ctypeslib.as_array(m.nSV, shape=(k,)).copy()
    return out
