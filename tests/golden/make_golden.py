"""Write the golden fixtures of tests/golden/ from the compiled reference CPU
k-means (oracle/_ref/libkmc_ref.so, built by oracle/Makefile from the
reference's comparisons/CPU_Kmeans sources).

    python tests/golden/make_golden.py

Fixtures (inputs and the reference's outputs, nothing else):
  ref_c1_blobs.npz      digits-like blobs, N=1797 D=64 K=10, seed 42, tol 1e-4,
                        maxiter 300: initial centroids, final centroids,
                        labels, convergence iteration
  ref_digits.npz        the reference's own data file
                        datasets/N1797_D64_digits-sklearn.csv (features only;
                        its last column is the class label), seed 7
  ref_steps.npz         single assign+update stages on small random inputs:
                        distances, labels, counts, updated centroids
  ref_runs.npz          whole runs from explicit centroids on small inputs
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "matrix-centric-k-means_amd"))

from oracle import Reference, build  # noqa: E402
from kmc import datagen  # noqa: E402

REF_DATA = "/root/reference/datasets/N1797_D64_digits-sklearn.csv"


def main():
    build()
    R = Reference()

    X, truth = datagen.digits_like()
    conv, C0, C, lab = R.run(X, 10, 1e-4, 42, 300)
    np.savez_compressed(os.path.join(HERE, "ref_c1_blobs.npz"), X=X, k=10, tol=np.float32(1e-4),
                        seed=42, maxiter=300, C0=C0, C=C, labels=lab, converged=conv)
    print("c1 blobs: converged at", conv)

    if os.path.exists(REF_DATA):
        raw = np.loadtxt(REF_DATA, delimiter=",", skiprows=1, dtype=np.float32)
        Xd = np.ascontiguousarray(raw[:, :64])
        conv, C0, C, lab = R.run(Xd, 10, 1e-4, 7, 300)
        np.savez_compressed(os.path.join(HERE, "ref_digits.npz"), X=Xd, k=10,
                            tol=np.float32(1e-4), seed=7, maxiter=300, C0=C0, C=C, labels=lab,
                            converged=conv)
        print("digits: converged at", conv)

    rng = np.random.default_rng(20261203)
    steps = {}
    cases = [(1, 1, 1), (2, 3, 2), (17, 5, 3), (64, 16, 8), (257, 33, 16), (400, 64, 37),
             (200, 130, 5), (300, 7, 64)]
    for idx, (n, d, k) in enumerate(cases):
        Xs = rng.normal(0, 3, size=(n, d)).astype(np.float32)
        if idx % 2:
            Xs = np.rint(Xs)  # integer-valued data: exact ties are common
        Cs = Xs[rng.choice(n, k, replace=False)].copy() if k <= n else Xs[:k].copy()
        if idx == 5:
            Cs[3] = Cs[2]  # duplicated centroid: tie in every distance
        D, lab, cnt, Cn = R.step(Xs, Cs)
        for name, arr in (("X", Xs), ("C", Cs), ("D", D), ("labels", lab), ("counts", cnt),
                          ("Cnew", Cn)):
            steps["%d_%s" % (idx, name)] = arr
    steps["ncases"] = len(cases)
    np.savez_compressed(os.path.join(HERE, "ref_steps.npz"), **steps)

    runs = {}
    rcases = [(200, 2, 3, 1e-4), (500, 10, 8, 0.0), (800, 24, 20, 1e-3), (600, 64, 10, 1e-5)]
    for idx, (n, d, k, tol) in enumerate(rcases):
        Xs, _, _ = datagen.blobs(n, d, k, 100 + idx, -5.0, 5.0, 0.5, 1.5)
        Cs = datagen.init_rows(Xs, k, 200 + idx)
        conv, Cf, lab = R.run_from(Xs, Cs, tol, 100)
        for name, arr in (("X", Xs), ("C0", Cs), ("C", Cf), ("labels", lab)):
            runs["%d_%s" % (idx, name)] = arr
        runs["%d_tol" % idx] = np.float32(tol)
        runs["%d_converged" % idx] = conv
        runs["%d_maxiter" % idx] = 100
    runs["ncases"] = len(rcases)
    np.savez_compressed(os.path.join(HERE, "ref_runs.npz"), **runs)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
