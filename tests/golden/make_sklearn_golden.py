"""Writes tests/golden/sklearn_libsvm.npz: input/output vectors from the
LIBSVM that scikit-learn bundles (sklearn/svm/src/libsvm, LIBSVM 3.10 with
sklearn's modifications), an independent third-party implementation of the
same SMO solver.  The oracle is checked against these vectors.

Run: python tests/golden/make_sklearn_golden.py   (needs scikit-learn)
"""
import os

import numpy as np
from sklearn.svm import SVC, SVR, NuSVC, NuSVR, OneClassSVM

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sklearn_libsvm.npz")

KERNELS = ["linear", "poly", "rbf", "sigmoid"]
TYPES = ["c_svc", "nu_svc", "one_class", "epsilon_svr", "nu_svr"]


def data(seed=0, n=150, d=6):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32).astype(np.float64)
    y = np.where(X[:, 0] + 0.5 * X[:, 1] + 0.3 * rng.standard_normal(n) > 0, 1.0, -1.0)
    t = (2 * X[:, 0] + np.sin(X[:, 1])).astype(np.float32).astype(np.float64)
    return X, y, t


def main():
    X, y, t = data()
    out = {"X": X, "y": y, "t": t}
    for kn in KERNELS:
        for st in TYPES:
            kw = dict(kernel=kn, gamma=0.15, coef0=0.5, degree=3, tol=1e-3, shrinking=False)
            # sklearn sorts class labels, so its internal +1 class is our -1:
            # fit on -y to get the same internal labelling as LIBSVM's
            if st == "c_svc":
                m = SVC(C=2.0, **kw).fit(X, -y)
            elif st == "nu_svc":
                m = NuSVC(nu=0.4, **kw).fit(X, -y)
            elif st == "one_class":
                m = OneClassSVM(nu=0.4, **kw).fit(X)
            elif st == "epsilon_svr":
                m = SVR(C=2.0, epsilon=0.1, **kw).fit(X, t)
            else:
                m = NuSVR(C=2.0, nu=0.4, **kw).fit(X, t)
            a = np.zeros(len(X))
            a[m.support_] = m.dual_coef_[0]
            if st in ("c_svc", "nu_svc"):
                a, rho = -a, m.intercept_[0]
            else:
                rho = -m.intercept_[0]
            key = f"{kn}_{st}"
            out[key + "_alpha"] = a
            out[key + "_rho"] = np.array(rho)
            out[key + "_iters"] = np.array(int(np.ravel(m.n_iter_)[0]))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
