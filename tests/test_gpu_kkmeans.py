"""GPU parity of kernel k-means (the reference's Popcorn path) through the C ABI.

The reference's GPU arithmetic (cuBLAS GEMM/SYRK, cuSPARSE SpMM/SpMV, RAFT
argmin) cannot run here, so the oracle is the FP64 restatement
(oracle/kmc_oracle.cpp, orc_kk_*), pinned in tests/test_oracle.py against the
reference unit test's own host formula. Tolerances:
  kernel matrix   fp32 store: |K - K64| <= 2e-6 * (|K64| + scale)
                  bf16 store: |K - K64| <= 2**-8 * |K64| + 2e-6 * scale
  distances       |D - D64| <= 4 (n 2^-24 + u_store) max|K|, u_store = 0 / 2^-9
  labels          identical to the FP64 argmin wherever the FP64 top-2 margin
                  exceeds twice the distance tolerance
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNELS = ["linear", "polynomial", "sigmoid", "rbf"]
PRODUCTS = ["dense", "sorted"]
KIND = {"linear": 0, "polynomial": 1, "sigmoid": 2, "rbf": 3}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _data(n, d, k, seed, with_truth=False):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 1, size=(k, d))
    lab = rng.integers(0, k, size=n)
    X = np.clip(centres[lab] + 0.1 * rng.standard_normal((n, d)), 0, 1).astype(np.float32)
    X[rng.random((n, d)) < 0.5] = 0.0
    if with_truth:
        noisy = lab.copy()
        flip = rng.random(n) < 0.3
        noisy[flip] = rng.integers(0, k, size=int(flip.sum()))
        return X, noisy.astype(np.int32)
    return X


def _dist_tol(n, store, scale):
    """Worst-case distance error: every K entry carries the storage rounding
    (bf16: 2^-9 relative) and every sum over a cluster up to gamma_n of fp32
    accumulation; three such terms in d(i,c), one more for headroom."""
    u_store = 0.0 if store == "fp32" else 2.0 ** -9
    return 4.0 * (n * 2.0 ** -24 + u_store) * scale


@pytest.mark.parametrize("product", PRODUCTS)
@pytest.mark.parametrize("store", ["fp32", "bf16"])
@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_matrix_and_distances(kmc_mod, oracle, kernel, store, product):
    n, d, k = 333, 37, 5
    X = _data(n, d, k, 1 + KIND[kernel])
    gamma = 1.0 / d
    _say("kk", kernel, store, n, d, k)
    K64 = oracle.kernel_matrix(X, KIND[kernel], gamma)
    scale = np.abs(K64).max()
    with kmc_mod.KernelEngine(n, d, k, kernel=kernel, store=store, product=product) as e:
        e.set_points(X)
        Kg = e.kernel_rows(0, n)
        if store == "fp32":
            tol = 2e-6 * (np.abs(K64) + scale)
        else:
            tol = 2.0 ** -8 * np.abs(K64) + 2e-6 * scale
        assert np.all(np.abs(Kg - K64) <= tol), np.max(np.abs(Kg - K64) - tol)
        labels = (np.arange(n) % k).astype(np.int32)
        labels[labels == k - 1] = 0  # cluster k-1 empty: +inf everywhere
        D = e.distances(labels)
        D64 = oracle.kk_distances(K64.astype(np.float64), labels.astype(np.uint32), k,
                                  add_diag=False)
        assert np.all(np.isinf(D[:, k - 1])) and np.all(np.isinf(D64[:, k - 1]))
        fin = np.isfinite(D64)
        dtol = _dist_tol(n, store, scale)
        assert np.all(np.abs(D[fin] - D64[fin]) <= dtol), np.max(np.abs(D[fin] - D64[fin]))


def _bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _gram_as_built(X, build):
    """The Gram values each build form computes exactly (before fp32
    accumulation): fp32 x.y; bf16x3 hi.hi + hi.lo + lo.hi with x = hi + lo;
    bf16 hi.hi."""
    X64 = X.astype(np.float64)
    if build == "fp32":
        return X64 @ X64.T
    hi = _bf16(X)
    if build == "bf16":
        return hi @ hi.T
    lo = _bf16((X64 - hi).astype(np.float32))
    return hi @ hi.T + hi @ lo.T + lo @ hi.T


def _kernel_of(V, kind, xn2, gamma):
    if kind == "linear":
        return V.copy()
    if kind == "polynomial":
        return (V + 1.0) ** 2
    if kind == "sigmoid":
        return np.tanh(V + 1.0)
    K = np.exp(-gamma * np.maximum(xn2[:, None] + xn2[None, :] - 2.0 * V, 0.0))
    np.fill_diagonal(K, 1.0)
    return K


@pytest.mark.parametrize("build", ["fp32", "bf16x3", "bf16"])
@pytest.mark.parametrize("store", ["fp32", "bf16"])
@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_matrix_build_forms(kmc_mod, kernel, store, build):
    """Every build form: K is exactly symmetric (the lower-triangle tiles are
    mirrored), and K equals the kernel of the Gram values the form computes
    (its operand rounding restated in FP64 here) within fp32 accumulation
    and the storage rounding. 700 points: off-diagonal and diagonal tiles,
    a padded last tile."""
    n, d, k = 700, 45, 4
    X = _data(n, d, k, 7 + KIND[kernel])
    gamma = 1.0 / d
    V = _gram_as_built(X, build)
    xn2 = (X.astype(np.float64) ** 2).sum(1)
    Kref = _kernel_of(V, kernel, xn2, gamma)
    scale = np.abs(Kref).max()
    _say("kk build", kernel, store, build)
    with kmc_mod.KernelEngine(n, d, k, kernel=kernel, store=store, build=build) as e:
        e.set_points(X)
        Kg = e.kernel_rows(0, n)
    assert np.array_equal(Kg, Kg.T)
    # accumulation: gamma_{3d} of sum |x||y| <= 3d 2^-24 (||x||^2 + ||y||^2)/2,
    # propagated through the kernel function (|f'| <= 2(|v|+1) for the
    # polynomial, 2 gamma K for the RBF), plus float evaluation of f
    xa = np.abs(X.astype(np.float64))
    sabs = xa @ xa.T
    acc = 3 * d * 2.0 ** -24 * sabs * 2.0
    fprime = {"linear": 1.0, "sigmoid": 1.0, "polynomial": 2.0 * (np.abs(V) + 1.0),
              "rbf": 2.0 * gamma * np.abs(Kref)}[kernel]
    tol = fprime * acc + 4e-7 * (np.abs(Kref) + scale)
    if store == "bf16":
        tol = tol + 2.0 ** -8 * np.abs(Kref)
    err = np.abs(Kg - Kref)
    assert np.all(err <= tol), float(np.max(err - tol))


@pytest.mark.parametrize("product", PRODUCTS)
@pytest.mark.parametrize("store", ["fp32", "bf16"])
@pytest.mark.parametrize("kernel", ["polynomial", "rbf"])
def test_one_iteration_labels(kmc_mod, oracle, kernel, store, product):
    n, d, k = 1000, 24, 7
    X, lab0 = _data(n, d, k, 11, with_truth=True)
    K64 = oracle.kernel_matrix(X, KIND[kernel], 1.0 / d)
    scale = np.abs(K64).max()
    _say("kk one iteration", kernel, store)
    with kmc_mod.KernelEngine(n, d, k, kernel=kernel, store=store, product=product) as e:
        e.set_points(X)
        e.set_labels(lab0)
        e.run(1)
        lab = e.labels()
    D64 = oracle.kk_distances(K64, lab0.astype(np.uint32), k, add_diag=True)
    want = np.argmin(D64, axis=1)
    srt = np.sort(D64, axis=1)
    margin = srt[:, 1] - srt[:, 0]
    dtol = _dist_tol(n, store, scale)
    safe = margin > 2 * dtol
    assert np.array_equal(lab[safe], want[safe])
    assert safe.mean() > 0.5


@pytest.mark.parametrize("product", PRODUCTS)
def test_linear_kernel_is_one_lloyd_step(kmc_mod, oracle, product):
    """With the linear kernel, one kernel k-means iteration from labels L is a
    Lloyd assignment against the means of L."""
    n, d, k = 2000, 16, 9
    from kmc import datagen
    X, _, _ = datagen.blobs(n, d, k, 5, -4, 4, 0.5, 1.5)
    lab0 = (np.arange(n) % k).astype(np.int32)
    C, _ = oracle.f64_centroids(X, lab0.astype(np.uint32), k)
    lab_ll, margin, _ = oracle.f64_assign(X, C)
    with kmc_mod.KernelEngine(n, d, k, kernel="linear", store="fp32", product=product) as e:
        e.set_points(X)
        e.set_labels(lab0)
        e.run(1)
        lab = e.labels()
    safe = margin > 1e-2
    assert np.array_equal(lab[safe], lab_ll[safe].astype(np.int32))


@pytest.mark.parametrize("product", PRODUCTS)
@pytest.mark.parametrize("store", ["fp32", "bf16"])
def test_run_to_stability_matches_fp64(kmc_mod, oracle, store, product):
    n, d, k = 800, 20, 6
    X = _data(n, d, k, 21)
    K64 = oracle.kernel_matrix(X, 1, 0.0)
    lab0 = (np.arange(n) % k).astype(np.int32)
    it64, lab64 = oracle.kk_lloyd(K64, lab0.astype(np.uint32), k, 100)
    _say("kk run", store)
    with kmc_mod.KernelEngine(n, d, k, kernel="polynomial", store=store,
                              product=product) as e:
        e.set_points(X)
        e.set_labels(lab0)
        res = e.run(100, check_converged=True, stop_stable=True)
        lab = e.labels()
    agree = np.mean(lab == lab64.astype(np.int32))
    assert agree >= (0.999 if store == "fp32" else 0.98), (agree, res, it64)
    D64 = oracle.kk_distances(K64, lab.astype(np.uint32), k, add_diag=True)
    obj64 = D64[np.arange(n), lab].sum()
    assert abs(res["objective"] - obj64) <= (1e-5 if store == "fp32" else 1e-2) * abs(obj64)


@pytest.mark.parametrize("store", ["fp32", "bf16"])
def test_many_clusters_sorted_product(kmc_mod, oracle, store):
    """k = 300 (beyond the dense product's 64): distances for labels with
    uneven and empty clusters, then one iteration, against the FP64 oracle."""
    n, d, k = 3000, 16, 300
    X = _data(n, d, 40, 31)
    K64 = oracle.kernel_matrix(X, 1, 0.0)
    scale = np.abs(K64).max()
    rng = np.random.default_rng(5)
    lab0 = (rng.integers(0, k, size=n) * rng.integers(0, 2, size=n)).astype(np.int32)
    lab0[lab0 == 7] = 8  # an empty cluster
    _say("kk many clusters", store)
    with kmc_mod.KernelEngine(n, d, k, kernel="polynomial", store=store, product="sorted") as e:
        e.set_points(X)
        D = e.distances(lab0)
        D64 = oracle.kk_distances(K64, lab0.astype(np.uint32), k, add_diag=False)
        fin = np.isfinite(D64)
        assert np.array_equal(np.isfinite(D), fin)
        dtol = _dist_tol(n, store, scale)
        assert np.all(np.abs(D[fin] - D64[fin]) <= dtol), np.max(np.abs(D[fin] - D64[fin]))
        e.set_labels(lab0)
        e.run(1)
        lab = e.labels()
    D64a = oracle.kk_distances(K64, lab0.astype(np.uint32), k, add_diag=True)
    want = np.argmin(D64a, axis=1)
    srt = np.sort(D64a, axis=1)
    safe = srt[:, 1] - srt[:, 0] > 2 * _dist_tol(n, store, scale)
    assert np.array_equal(lab[safe], want[safe])
    # bf16 storage widens the tolerance: fewer points are decided by it
    assert safe.mean() > (0.5 if store == "fp32" else 0.1), safe.mean()
