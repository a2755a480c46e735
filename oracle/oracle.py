"""ctypes loader for the CPU oracle (oracle/libkmc_oracle.so) and the compiled
reference CPU k-means (oracle/_ref/libkmc_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libkmc_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkmc_ref.so")

_P = ctypes.c_void_p
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


def build():
    """Compile the oracle (and the reference, where its sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _ptr(a):
    return a.ctypes.data


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_cpu_init_centroids.restype = ctypes.c_int
        L.orc_cpu_init_centroids.argtypes = [_P, _sz, _u32, _u32, ctypes.c_int, _P]
        L.orc_cpu_distances.argtypes = [_P, _sz, _u32, _P, _u32, _P]
        L.orc_cpu_argmin.argtypes = [_P, _sz, _u32, _P, _P]
        L.orc_cpu_centroids.argtypes = [_P, _sz, _u32, _P, _P, _u32, _P]
        L.orc_cpu_converged.restype = ctypes.c_int
        L.orc_cpu_converged.argtypes = [_P, _P, _u32, _u32, ctypes.c_float]
        L.orc_cpu_lloyd.restype = _u64
        L.orc_cpu_lloyd.argtypes = [_P, _sz, _u32, _u32, ctypes.c_float, _u64, _P, _P, _P, _P]
        L.orc_f64_row_norms.argtypes = [_P, _sz, _u32, _P]
        L.orc_plusplus.argtypes = [_P, _sz, _u32, _u32, _P, _P, _P]
        L.orc_plusplus_trials.argtypes = [_P, _sz, _u32, _u32, _u32, _P, _P, _P]
        L.orc_f64_distances.argtypes = [_P, _sz, _u32, _P, _u32, _P]
        L.orc_f64_inertia.restype = ctypes.c_double
        L.orc_f64_inertia.argtypes = [_P, _sz, _u32, _P, _P]
        L.orc_f64_centroids.argtypes = [_P, _sz, _u32, _P, _u32, _P, _P]
        L.orc_f64_assign.restype = ctypes.c_double
        L.orc_f64_assign.argtypes = [_P, _sz, _u32, _P, _u32, _P, _P]
        L.orc_kk_kernel_matrix.argtypes = [_P, _sz, _u32, ctypes.c_int, ctypes.c_double, _P]
        L.orc_kk_distances.argtypes = [_P, _sz, _P, _u32, ctypes.c_int, _P]
        L.orc_kk_lloyd.restype = _u64
        L.orc_kk_lloyd.argtypes = [_P, _sz, _u32, _u64, _P]
        L.orc_gpu_lloyd.restype = _u64
        L.orc_gpu_lloyd.argtypes = [_P, _sz, _u32, _u32, ctypes.c_float, _u64, ctypes.c_int,
                                    _P, _P, _P, _P]
        self.L = L

    # -- reference CPU k-means restatement (float) --
    def init_centroids(self, X, k, seed):
        X = np.ascontiguousarray(X, np.float32)
        C = np.empty((k, X.shape[1]), np.float32)
        if self.L.orc_cpu_init_centroids(_ptr(X), X.shape[0], X.shape[1], k, seed, _ptr(C)):
            raise ValueError("fewer than k distinct rows")
        return C

    def distances(self, X, C):
        X = np.ascontiguousarray(X, np.float32)
        C = np.ascontiguousarray(C, np.float32)
        D = np.empty((X.shape[0], C.shape[0]), np.float32)
        self.L.orc_cpu_distances(_ptr(X), X.shape[0], X.shape[1], _ptr(C), C.shape[0], _ptr(D))
        return D

    def argmin(self, D):
        D = np.ascontiguousarray(D, np.float32)
        lab = np.empty(D.shape[0], np.uint32)
        cnt = np.empty(D.shape[1], np.uint32)
        self.L.orc_cpu_argmin(_ptr(D), D.shape[0], D.shape[1], _ptr(lab), _ptr(cnt))
        return lab, cnt

    def centroids(self, X, labels, k):
        X = np.ascontiguousarray(X, np.float32)
        lab = np.ascontiguousarray(labels, np.uint32)
        cnt = np.bincount(lab, minlength=k).astype(np.uint32)
        C = np.empty((k, X.shape[1]), np.float32)
        self.L.orc_cpu_centroids(_ptr(X), X.shape[0], X.shape[1], _ptr(lab), _ptr(cnt), k, _ptr(C))
        return C, cnt

    def converged(self, C, Clast, tol):
        C = np.ascontiguousarray(C, np.float32)
        Clast = np.ascontiguousarray(Clast, np.float32)
        return bool(self.L.orc_cpu_converged(_ptr(C), _ptr(Clast), C.shape[0], C.shape[1],
                                             ctypes.c_float(tol)))

    def lloyd(self, X, C0, tol, maxiter):
        X = np.ascontiguousarray(X, np.float32)
        C = np.array(C0, dtype=np.float32, copy=True)
        n, d = X.shape
        k = C.shape[0]
        lab = np.empty(n, np.uint32)
        cnt = np.empty(k, np.uint32)
        ran = np.zeros(1, np.uint64)
        conv = self.L.orc_cpu_lloyd(_ptr(X), n, d, k, ctypes.c_float(tol), maxiter, _ptr(C),
                                    _ptr(lab), _ptr(cnt), _ptr(ran))
        return int(conv), lab, C, cnt, int(ran[0])

    def gpu_lloyd(self, X, labels0, k, tol, maxiter, check=True):
        """The reference GPU engine's loop rules (src/kmeans.cu:772-1122) on the
        linear path from initial labels: returns (converged, labels, C, scores,
        iterations run); scores are the float per-iteration scores."""
        X = np.ascontiguousarray(X, np.float32)
        n, d = X.shape
        lab = np.array(labels0, dtype=np.uint32, copy=True)
        C = np.empty((k, d), np.float32)
        sc = np.zeros(max(1, int(maxiter)), np.float32)
        ran = np.zeros(1, np.uint64)
        conv = self.L.orc_gpu_lloyd(_ptr(X), n, d, k, ctypes.c_float(tol), maxiter,
                                    int(bool(check)), _ptr(lab), _ptr(C), _ptr(sc), _ptr(ran))
        r = int(ran[0])
        return int(conv), lab, C, sc[:r].copy(), r

    # -- FP64 --
    def f64_distances(self, X, C):
        X = np.ascontiguousarray(X, np.float32)
        C = np.ascontiguousarray(C, np.float32)
        D = np.empty((X.shape[0], C.shape[0]), np.float64)
        self.L.orc_f64_distances(_ptr(X), X.shape[0], X.shape[1], _ptr(C), C.shape[0], _ptr(D))
        return D

    def plusplus(self, X, k, uniforms, trials=1):
        """k-means++ seeding with the given uniforms (1 + (k-1)*trials of
        them): (chosen indices, C). trials > 1: greedy k-means++."""
        X = np.ascontiguousarray(X, np.float32)
        u = np.ascontiguousarray(uniforms, np.float64)
        assert u.size >= 1 + (k - 1) * max(1, trials)
        chosen = np.empty(k, np.int64)
        C = np.empty((k, X.shape[1]), np.float32)
        self.L.orc_plusplus_trials(_ptr(X), X.shape[0], X.shape[1], k, trials, _ptr(u),
                                   _ptr(chosen), _ptr(C))
        return chosen, C

    def lloyd_explicit(self, X, C0, tol, maxiter, check):
        """The explicit-Lloyd rules of the reference's RAFT benchmark
        (experiments/raft-bench/src/kmeans.cu:95-212), over this oracle's
        exact labels (the reference CPU float direct form, first minimum) and
        FP64 means rounded once: per iteration t, cost_t = FP64 inertia of the
        assignment, an empty cluster keeps its centroid, then stop when
        (check and sum of squared centroid changes < tol) or (t >= 2 and
        cost_t / cost_{t-1} > 1 - tol). Returns (stop iteration or maxiter,
        labels, C, list of costs)."""
        X = np.ascontiguousarray(X, np.float32)
        C = np.array(C0, dtype=np.float32, copy=True)
        k = C.shape[0]
        prior = 0.0
        costs = []
        lab = None
        for t in range(1, int(maxiter) + 1):
            lab, _ = self.argmin(self.distances(X, C))
            cost = self.f64_inertia(X, C, lab)
            costs.append(cost)
            Cn, cnt = self.f64_centroids(X, lab, k)
            Cn[cnt == 0] = C[cnt == 0]
            shift = float(((Cn.astype(np.float64) - C.astype(np.float64)) ** 2).sum())
            C = Cn
            done = bool(check) and shift < tol
            if t >= 2 and prior > 0.0 and cost / prior > 1.0 - tol:
                done = True
            prior = cost
            if done:
                return t, lab, C, costs
        return int(maxiter), lab, C, costs

    def f64_inertia(self, X, C, labels):
        X = np.ascontiguousarray(X, np.float32)
        C = np.ascontiguousarray(C, np.float32)
        lab = np.ascontiguousarray(labels, np.uint32)
        return self.L.orc_f64_inertia(_ptr(X), X.shape[0], X.shape[1], _ptr(C), _ptr(lab))

    def f64_centroids(self, X, labels, k):
        X = np.ascontiguousarray(X, np.float32)
        lab = np.ascontiguousarray(labels, np.uint32)
        C = np.empty((k, X.shape[1]), np.float32)
        cnt = np.empty(k, np.uint32)
        self.L.orc_f64_centroids(_ptr(X), X.shape[0], X.shape[1], _ptr(lab), k, _ptr(C), _ptr(cnt))
        return C, cnt

    def f64_assign(self, X, C):
        X = np.ascontiguousarray(X, np.float32)
        C = np.ascontiguousarray(C, np.float32)
        lab = np.empty(X.shape[0], np.uint32)
        margin = np.empty(X.shape[0], np.float64)
        inert = self.L.orc_f64_assign(_ptr(X), X.shape[0], X.shape[1], _ptr(C), C.shape[0],
                                      _ptr(lab), _ptr(margin))
        return lab, margin, inert

    # -- kernel k-means (FP64) --
    def kernel_matrix(self, X, kind, gamma=0.0):
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        K = np.empty((n, n), np.float64)
        self.L.orc_kk_kernel_matrix(_ptr(X), n, X.shape[1], kind, gamma, _ptr(K))
        return K

    def kk_distances(self, K, labels, k, add_diag=True):
        K = np.ascontiguousarray(K, np.float64)
        lab = np.ascontiguousarray(labels, np.uint32)
        D = np.empty((K.shape[0], k), np.float64)
        self.L.orc_kk_distances(_ptr(K), K.shape[0], _ptr(lab), k, int(add_diag), _ptr(D))
        return D

    def kk_lloyd(self, K, labels0, k, maxiter):
        K = np.ascontiguousarray(K, np.float64)
        lab = np.array(labels0, dtype=np.uint32, copy=True)
        it = self.L.orc_kk_lloyd(_ptr(K), K.shape[0], k, maxiter, _ptr(lab))
        return int(it), lab


class Reference:
    """The reference CPU k-means, compiled from its own sources (oracle/_ref)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.kref_run.restype = _u64
        L.kref_run.argtypes = [_P, _sz, _u32, _u32, ctypes.c_float, ctypes.c_int, _u64, _P, _P, _P]
        L.kref_run_from.restype = _u64
        L.kref_run_from.argtypes = [_P, _sz, _u32, _u32, ctypes.c_float, _u64, _P, _P]
        L.kref_step.argtypes = [_P, _sz, _u32, _u32, _P, _P, _P, _P, _P]
        L.kref_converged.restype = ctypes.c_int
        L.kref_converged.argtypes = [_P, _sz, _u32, _u32, ctypes.c_float, _P, _P]
        L.kref_time_run.restype = ctypes.c_double
        L.kref_time_run.argtypes = [_P, _sz, _u32, _u32, _u64, _P, _P]
        self.L = L

    @staticmethod
    def available(path=REF_SO):
        return os.path.exists(path)

    def run(self, X, k, tol, seed, maxiter):
        X = np.ascontiguousarray(X, np.float32)
        n, d = X.shape
        C0 = np.empty((k, d), np.float32)
        C = np.empty((k, d), np.float32)
        lab = np.empty(n, np.uint32)
        conv = self.L.kref_run(_ptr(X), n, d, k, ctypes.c_float(tol), seed, maxiter, _ptr(C0),
                               _ptr(C), _ptr(lab))
        return int(conv), C0, C, lab

    def run_from(self, X, C0, tol, maxiter):
        X = np.ascontiguousarray(X, np.float32)
        C = np.array(C0, dtype=np.float32, copy=True)
        n, d = X.shape
        lab = np.empty(n, np.uint32)
        conv = self.L.kref_run_from(_ptr(X), n, d, C.shape[0], ctypes.c_float(tol), maxiter,
                                    _ptr(C), _ptr(lab))
        return int(conv), C, lab

    def step(self, X, C):
        X = np.ascontiguousarray(X, np.float32)
        C = np.ascontiguousarray(C, np.float32)
        n, d = X.shape
        k = C.shape[0]
        D = np.empty((n, k), np.float32)
        lab = np.empty(n, np.uint32)
        cnt = np.empty(k, np.uint32)
        Cn = np.empty((k, d), np.float32)
        self.L.kref_step(_ptr(X), n, d, k, _ptr(C), _ptr(D), _ptr(lab), _ptr(cnt), _ptr(Cn))
        return D, lab, cnt, Cn

    def time_run(self, X, C0, maxiter):
        """Seconds spent in the reference's Kmeans::run(maxiter) alone."""
        X = np.ascontiguousarray(X, np.float32)
        C0 = np.ascontiguousarray(C0, np.float32)
        conv = np.zeros(1, np.uint64)
        return self.L.kref_time_run(_ptr(X), X.shape[0], X.shape[1], C0.shape[0], maxiter,
                                    _ptr(C0), _ptr(conv))

    def converged(self, X, C, Clast, tol):
        X = np.ascontiguousarray(X, np.float32)
        C = np.ascontiguousarray(C, np.float32)
        Clast = np.ascontiguousarray(Clast, np.float32)
        return bool(self.L.kref_converged(_ptr(X), X.shape[0], X.shape[1], C.shape[0],
                                          ctypes.c_float(tol), _ptr(C), _ptr(Clast)))


def splitmix_uniforms(seed, k):
    """The k uniforms in [0,1) of the engine's k-means++ draws: splitmix64
    (Steele, Lea, Flood 2014), top 53 bits / 2^53."""
    mask = (1 << 64) - 1
    st = seed & mask
    out = []
    for _ in range(k):
        st = (st + 0x9E3779B97F4A7C15) & mask
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out.append((z >> 11) * 2.0 ** -53)
    return np.array(out, dtype=np.float64)
