"""The C-ABI boundary, checked without a GPU: libsvm_hip.so loads, exports
every function include/*.h declares, the structs have the byte layout a C
compiler gives the headers, parameter checking matches the reference's
messages, and without a device the product fails loudly (no CPU solver)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import libsvm_hip
from libsvm_hip import datasets, ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "svm.h"), os.path.join(ROOT, "include", "svm_hip.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = set()
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(svm\w*)\s*\(([^;{]*)\)\s*;", text):
        names.add(m.group(2))
    return names


def test_library_exports_every_declared_symbol():
    L = libsvm_hip.lib()
    declared = set()
    for h in HEADERS:
        declared |= declared_functions(h)
    assert len(declared) >= 30
    assert declared | {"libsvm_version"} == set(libsvm_hip.SVM_H_SYMBOLS) | set(libsvm_hip.SVM_HIP_SYMBOLS)
    missing = [n for n in declared | {"libsvm_version"} if not hasattr(L, n)]
    assert not missing, missing
    assert C.c_int.in_dll(L, "libsvm_version").value == 318


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "svm_hip.h"
#define P(T, F) printf(#T "." #F " %zu\n", offsetof(struct T, F));
#define S(T) printf(#T " %zu\n", sizeof(struct T));
int main(void) {
  S(svm_node) S(svm_problem) S(svm_parameter) S(svm_model)
  P(svm_parameter, cuda_flag) P(svm_parameter, svm_type) P(svm_parameter, gamma)
  P(svm_parameter, cache_size) P(svm_parameter, weight) P(svm_parameter, nu)
  P(svm_parameter, shrinking) P(svm_parameter, probability)
  P(svm_model, nr_class) P(svm_model, SV) P(svm_model, sv_indices) P(svm_model, free_sv)
  printf("svmhip_options %zu\nsvmhip_stats %zu\n", sizeof(svmhip_options), sizeof(svmhip_stats));
  printf("svmhip_subproblem %zu\nsvmhip_completion %zu\n", sizeof(svmhip_subproblem), sizeof(svmhip_completion));
  printf("svmhip_completion.rho %zu\nsvmhip_completion.device_ms %zu\n", offsetof(svmhip_completion, rho),
         offsetof(svmhip_completion, device_ms));
  return 0;
}
"""


def test_struct_layout_matches_c_compiler(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.check_output([str(exe)]).decode().splitlines())
    got = {k: int(v) for k, v in out.items()}
    assert got["svm_node"] == C.sizeof(ffi.svm_node) == 16
    assert got["svm_problem"] == C.sizeof(ffi.svm_problem)
    assert got["svm_parameter"] == C.sizeof(ffi.svm_parameter)
    assert got["svm_model"] == C.sizeof(ffi.svm_model)
    assert got["svm_parameter.cuda_flag"] == 0  # first member, as in the reference
    for f in ("svm_type", "gamma", "cache_size", "weight", "nu", "shrinking", "probability"):
        assert got[f"svm_parameter.{f}"] == getattr(ffi.svm_parameter, f).offset
    for f in ("nr_class", "SV", "sv_indices", "free_sv"):
        assert got[f"svm_model.{f}"] == getattr(ffi.svm_model, f).offset
    assert got["svmhip_options"] == C.sizeof(libsvm_hip.svmhip_options)
    assert got["svmhip_stats"] == C.sizeof(libsvm_hip.svmhip_stats)
    assert got["svmhip_subproblem"] == C.sizeof(libsvm_hip.svmhip_subproblem)
    assert got["svmhip_completion"] == C.sizeof(libsvm_hip.svmhip_completion)
    assert got["svmhip_completion.rho"] == libsvm_hip.svmhip_completion.rho.offset
    assert got["svmhip_completion.device_ms"] == libsvm_hip.svmhip_completion.device_ms.offset


def test_check_parameter_messages():
    ip, ix, v, y, d = datasets.zipf_binary(n=200)
    prob = ffi.Problem(ip, ix, v, y, d)
    ok = ffi.make_param(gamma=0.1)
    assert libsvm_hip.check_parameter(prob, ok) is None
    cases = [
        (dict(cuda_flag=0), "this library has only the GPU solver: set cuda_flag = 1"),
        (dict(eps=0.0), "eps <= 0"),
        (dict(C_=0.0), "C <= 0"),
        (dict(gamma=-1.0), "gamma < 0"),
        (dict(cache_size=0.0), "cache_size <= 0"),
        (dict(svm_type=ffi.NU_SVC, nu=1.5), "nu <= 0 or nu > 1"),
        (dict(svm_type=ffi.NU_SVC, nu=0.9), "specified nu is infeasible"),
        (dict(shrinking=2), "shrinking != 0 and shrinking != 1"),
        (dict(probability=1), "probability estimates are not supported by the GPU solver"),
        (dict(svm_type=7), "unknown svm type"),
        (dict(kernel_type=9), "unknown kernel type"),
    ]
    for kw, msg in cases:
        prm = ffi.make_param(**{"gamma": 0.1, **kw})
        assert libsvm_hip.check_parameter(prob, prm) == msg, kw


@pytest.mark.skipif(libsvm_hip.device_count() > 0, reason="a device is present")
def test_no_device_fails_loudly():
    ip, ix, v, y, d = datasets.zipf_binary(n=100)
    prob = ffi.Problem(ip, ix, v, y, d)
    prm = ffi.make_param(gamma=0.1)
    with pytest.raises(libsvm_hip.SolverError, match="ERR_DEVICE"):
        libsvm_hip.DeviceSolver(prob, prm)
    libsvm_hip.set_quiet(True)
    with pytest.raises(libsvm_hip.SolverError, match="no model"):
        libsvm_hip.svm_train(prob, prm)


def test_datasets_deterministic_and_float32_valued():
    a = datasets.zipf_binary(n=300)
    b = datasets.zipf_binary(n=300)
    for x, z in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, z)
    ip, ix, v, y, d = a
    assert d == 123 and set(np.unique(v)) == {1.0}
    assert abs(np.diff(ip).mean() - 14) < 1.5
    assert 0.15 < np.mean(y > 0) < 0.35
    ip, ix, v, y, d = datasets.blob_mask_csr(n=2000)
    assert d == 784 and 0.15 < len(v) / (2000 * 784) < 0.23
    np.testing.assert_array_equal(v, v.astype(np.float32).astype(np.float64))
    assert np.all((v > 0) & (v <= 1))
    X, yy = datasets.gmm_dense_matrix(n=1000, d=256)
    assert X.dtype == np.float32 and set(np.unique(yy)) == {-1.0, 1.0}
    Xc, yc, fold = datasets.clusters_10class()
    assert 36000 <= len(yc) <= 44000 and set(np.unique(fold)) == set(range(5))


def test_launch_table_parses_and_covers_the_shapes():
    """libsvm-cuda_amd/tune/launch_table.txt (written by tune/autotune.py):
    every row has the nine fields svmhip_create reads, grids within 1..512
    (the tuner also tries grids between the powers of two), copies within
    1..4, and for each mode and row kind the size ranges leave no gap from 1
    row upwards."""
    path = os.path.join(ROOT, "libsvm-cuda_amd", "tune", "launch_table.txt")
    rows = []
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split("#")[0].split()
        assert len(f) == 9, line
        mode, full, dense, lmin, lmax, dmin, dmax, grid, xrep = f[0], *map(int, f[1:])
        assert mode in ("exact", "gram") and full in (0, 1) and dense in (0, 1)
        assert 1 <= grid <= 512 and 1 <= xrep <= 4
        rows.append((mode, full, dense, dmin, dmax, lmin, lmax))
    groups = {}
    for mode, full, dense, dmin, dmax, lmin, lmax in rows:
        groups.setdefault((mode, full, dense, dmin, dmax), []).append((lmin, lmax))
    assert ("gram", 1, 1, 65, 2 ** 31 - 1) in groups and ("exact", 1, 1, 65, 2 ** 31 - 1) in groups
    for key, rs in groups.items():
        rs.sort()
        assert rs[0][0] == 1 and rs[-1][1] == 2 ** 31 - 1, key
        for (a0, a1), (b0, b1) in zip(rs, rs[1:]):
            assert b0 == a1 + 1, key
