"""Builds a struct svm_model (include/svm.h) in Python from arrays, for
tests that feed a model trained elsewhere (scikit-learn's bundled LIBSVM)
to the library and the oracle."""
import ctypes as C

import numpy as np

from libsvm_hip import ffi


class ArrayModel:
    """svm_model over numpy arrays; .ptr is a POINTER(svm_model)."""

    def __init__(self, param, SV, sv_coef, rho, label=None, nSV=None, probA=None, probB=None):
        SV = np.asarray(SV, dtype=np.float64)
        self.sv = ffi.Problem.from_dense(SV, np.zeros(len(SV)))
        self.coef = np.ascontiguousarray(np.atleast_2d(sv_coef), dtype=np.float64)
        self.coef_rows = (C.POINTER(C.c_double) * len(self.coef))(
            *[self.coef[i].ctypes.data_as(C.POINTER(C.c_double)) for i in range(len(self.coef))])
        self.keep = []
        m = ffi.svm_model()
        m.param = param
        m.nr_class = 2 if label is None else len(label)
        m.l = len(SV)
        m.SV = self.sv.struct.x
        m.sv_coef = C.cast(self.coef_rows, C.POINTER(C.POINTER(C.c_double)))
        m.rho = self._d(rho)
        m.probA = self._d(probA) if probA is not None else None
        m.probB = self._d(probB) if probB is not None else None
        m.label = self._i(label) if label is not None else None
        m.nSV = self._i(nSV) if nSV is not None else None
        m.sv_indices = None
        m.free_sv = 0
        self.model = m
        self.ptr = C.pointer(m)

    def _d(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def _i(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        self.keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int))
