"""Helpers for the svm-train command-line tests: the binary's path and an
svmlight writer that round-trips float32-valued features exactly."""
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SVM_TRAIN = os.path.join(ROOT, "libsvm-cuda_amd", "bin", "svm-train")


def write_svmlight(path, prob):
    with open(path, "w") as f:
        for r in range(prob.l):
            a, b = prob.indptr[r], prob.indptr[r + 1]
            # svm_node indices are column + 1 (ffi.Problem's convention)
            feats = " ".join(f"{int(i) + 1}:{float(v)!r}" for i, v in zip(prob.indices[a:b], prob.data[a:b]))
            f.write(f"{float(prob.y[r])!r} {feats}\n")
