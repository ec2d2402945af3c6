"""Writes tests/golden/sklearn_multiclass.npz: one-vs-one multi-class
training, decision values and probability outputs from the LIBSVM that
scikit-learn bundles (LIBSVM 3.10 with sklearn's changes), to pin the
oracle's multi-class svm_train (support-vector packing, sv_coef layout and
rho order, libsvm/svm.cpp:2246-2386), its svm_predict_values
(:2569-2643) and the library's svm_predict_probability (:2644-2700).

The rows are ordered so that the classes first appear in increasing label
order, which makes LIBSVM's class order (first appearance) and sklearn's
(sorted labels) the same.

Run: python tests/golden/make_sklearn_multiclass_golden.py   (needs scikit-learn)
"""
import os

import numpy as np
from sklearn.svm import SVC

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sklearn_multiclass.npz")


def data(seed=4, k=4, per=60, d=5):
    rng = np.random.default_rng(seed)
    mu = 1.2 * rng.standard_normal((k, d))
    y = np.repeat(np.arange(1, k + 1), per)
    X = mu[y - 1] + rng.standard_normal((len(y), d))
    perm = rng.permutation(len(y))
    X, y = X[perm], y[perm]
    first = [int(np.nonzero(y == c)[0][0]) for c in range(1, k + 1)]
    order = first + [i for i in range(len(y)) if i not in first]  # classes appear as 1, 2, ..., k
    X, y = X[order], y[order]
    X = X.astype(np.float32).astype(np.float64)
    Xt = (mu[rng.integers(0, k, 200)] + rng.standard_normal((200, d))).astype(np.float32).astype(np.float64)
    return X, y.astype(np.float64), Xt


def main():
    X, y, Xt = data()
    out = {"X": X, "y": y, "Xt": Xt}
    for kn in ("rbf", "poly", "linear"):
        m = SVC(C=2.0, kernel=kn, gamma=0.2, coef0=0.5, degree=2, tol=1e-3, shrinking=False,
                decision_function_shape="ovo").fit(X, y)
        out[f"{kn}_support"] = m.support_.astype(np.int32)
        out[f"{kn}_n_support"] = m.n_support_.astype(np.int32)
        out[f"{kn}_dual_coef"] = m.dual_coef_
        out[f"{kn}_intercept"] = m.intercept_
        out[f"{kn}_dec"] = m.decision_function(Xt)
        out[f"{kn}_pred"] = m.predict(Xt)
    # probability outputs of a model with Platt parameters (the calibration
    # itself runs sklearn's own cross-validation; only prediction is pinned)
    mp = SVC(C=2.0, kernel="rbf", gamma=0.2, tol=1e-3, shrinking=False, probability=True, random_state=0,
             decision_function_shape="ovo").fit(X, y)
    out["prob_support"] = mp.support_.astype(np.int32)
    out["prob_n_support"] = mp.n_support_.astype(np.int32)
    out["prob_dual_coef"] = mp.dual_coef_
    out["prob_intercept"] = mp.intercept_
    out["prob_probA"] = mp.probA_
    out["prob_probB"] = mp.probB_
    out["prob_proba"] = mp.predict_proba(Xt)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
