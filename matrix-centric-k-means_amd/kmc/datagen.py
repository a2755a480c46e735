"""Seeded synthetic inputs for the benchmark and test workloads.

All generators use numpy's PCG64 (``np.random.default_rng(seed)``) and return
float32 row-major arrays, so the same seed gives the same bytes everywhere.
"""
import numpy as np


def blobs(n, d, k, seed, centre_lo, centre_hi, sigma_lo, sigma_hi, chunk=1 << 18):
    """Gaussian blobs: centres ~ U[lo,hi]^d, per-blob sigma ~ U[sigma_lo, sigma_hi],
    blob of each point uniform over k, x = centre + sigma * N(0, I).
    Returns (X float32 n x d, truth int32 n, centres float32 k x d)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(centre_lo, centre_hi, size=(k, d)).astype(np.float32)
    sig = rng.uniform(sigma_lo, sigma_hi, size=k).astype(np.float32)
    truth = rng.integers(0, k, size=n, dtype=np.int32)
    X = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        z = rng.standard_normal((e - s, d), dtype=np.float32)
        t = truth[s:e]
        X[s:e] = centres[t] + sig[t, None] * z
    return X, truth, centres


def digits_like(seed=1797):
    """N=1797, D=64, 10 isotropic blobs; centres U[0,16]^64, sigma 2, values
    clipped to [0,16] and rounded to integers, stored as float32."""
    X, truth, _ = blobs(1797, 64, 10, seed, 0.0, 16.0, 2.0, 2.0)
    X = np.rint(np.clip(X, 0.0, 16.0)).astype(np.float32)
    return X, truth


def large_blobs(n=1_000_000, d=128, k=256, seed=20261203):
    """Centres U[-10,10]^d, sigma_k ~ U[0.5, 2.0]."""
    return blobs(n, d, k, seed, -10.0, 10.0, 0.5, 2.0)


def huge_blobs(n=4_000_000, d=256, k=4096, seed=20261204):
    """Centres U[-5,5]^d, sigma 1."""
    return blobs(n, d, k, seed, -5.0, 5.0, 1.0, 1.0)


def sparse_unit_blobs(n=65536, d=784, k=10, seed=65536, zero_frac=0.7):
    """Values in [0,1] around 10 blob centres, 70% of entries zeroed."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, size=(k, d)).astype(np.float32)
    truth = rng.integers(0, k, size=n, dtype=np.int32)
    X = centres[truth] + 0.15 * rng.standard_normal((n, d), dtype=np.float32)
    np.clip(X, 0.0, 1.0, out=X)
    X[rng.random((n, d)) < zero_frac] = 0.0
    return X.astype(np.float32), truth


def init_rows(X, k, seed):
    """k distinct (by value) rows of X chosen with the seed."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    chosen, seen = [], set()
    order = rng.permutation(n)
    for r in order:
        key = X[r].tobytes()
        if key in seen:
            continue
        seen.add(key)
        chosen.append(r)
        if len(chosen) == k:
            break
    if len(chosen) < k:
        raise ValueError("fewer than k distinct rows")
    return X[np.array(chosen)].copy()


def batch_problems(count=512, seed=20261205):
    """Independent Lloyd problems of uneven shape: N log-uniform in [1e3, 2e5],
    D in {16, 32, 64, 128}, K uniform in [4, 128], Gaussian blobs (centres
    U[-10,10]^D, sigma U[0.5, 2]) with per-instance seeds, K seed-chosen
    distinct rows as initial centroids. Returns [(X, C0), ...]."""
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        d = int(rng.choice([16, 32, 64, 128]))
        k = int(rng.integers(4, 129))
        X, _, _ = blobs(n, d, k, seed + 1 + i, -10.0, 10.0, 0.5, 2.0)
        probs.append((X, init_rows(X, k, seed + 10_000 + i)))
    return probs
