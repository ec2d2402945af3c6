"""svm-train on the device: the model file it writes is byte-identical to
svm_save_model of the same library's svm_train on the same rows, and its -v
report is the oracle's cross-validation result with the same rand()
stream (a fresh process, seed 1)."""
import ctypes
import subprocess

import numpy as np
import pytest

from cli_util import SVM_TRAIN, write_svmlight
from libsvm_hip import datasets, ffi

pytestmark = pytest.mark.gpu


def test_model_file_matches_library(hip, tmp_path):
    ip, ix, v, y, d = datasets.zipf_binary(n=1500, seed=31)
    prob = ffi.Problem(ip, ix, v, y, d)
    data = tmp_path / "train.txt"
    write_svmlight(data, prob)
    model = tmp_path / "cli.model"
    r = subprocess.run([SVM_TRAIN, "-C", "-q", "-c", "2", str(data), str(model)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    gamma = 1.0 / (int(prob.indices.max()) + 1)  # svm-train's default: 1 / largest feature index
    prm = ffi.make_param(gamma=gamma, C_=2.0, eps=1e-3, shrinking=0)
    m = hip.svm_train(prob, prm)
    ref = tmp_path / "lib.model"
    m.save(str(ref))
    assert model.read_text() == ref.read_text()


def test_cross_validation_report_matches_oracle(hip, ora, tmp_path):
    X, y, _ = datasets.clusters_10class(scale=0.04, seed=8)
    prob = ffi.Problem.from_dense(X.astype(np.float64), y.astype(np.float64))
    data = tmp_path / "cv.txt"
    write_svmlight(data, prob)
    r = subprocess.run([SVM_TRAIN, "-q", "-v", "5", "-g", "0.0078125", "-c", "4", str(data)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)  # the CLI process starts from glibc's default seed
    prm = ffi.make_param(gamma=0.0078125, C_=4.0, eps=1e-3, shrinking=0)
    target = ora.svm_cross_validation(prob, prm, 5)
    acc = 100.0 * np.mean(target == prob.y)
    print(f"  {r.stdout.strip()} (oracle {acc:g}%)")
    assert r.stdout.strip() == "Cross Validation Accuracy = %s%%" % ("%g" % acc)
