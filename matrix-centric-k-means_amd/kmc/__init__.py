"""Python host binding of the gfx950 k-means engine (C ABI in include/kmc.h).

The engine is the HIP library ``lib/libkmc.so`` next to this package; there is
no CPU fallback. Importing works without a GPU (so that the library's exports
can be checked); creating an engine without a HIP device raises KmcError.

The class mirrors the reference engine ``Kmeans`` (src/kmeans.cuh:117-139):
construct with (n, d, k, tol), give points and initial centroids, ``run``
returns the convergence iteration, ``labels()`` is the label write-back
(src/kmeans.cu:1084-1103) and ``score`` the reference's objective
(sum of min distances without ||x||^2, src/kmeans.cu:1053-1061).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libkmc.so")

OK, ERR_ARG, ERR_NO_DEVICE, ERR_OUT_OF_BOUNDS, ERR_INFILE, ERR_HIP, ERR_OOM, \
    ERR_STATE, ERR_UNSUPPORTED = range(9)

PRECISIONS = {"fp32": 0, "bf16x3": 1, "bf16": 2, "fp8": 3}
UPDATES = {"ordered": 0, "tree": 1, "fixed": 2}
RUN_CHECK_CONVERGED = 0x1
RUN_TIME_KERNELS = 0x2


class KmcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("kmc status %d (%s): %s" % (status, _status_text(status), msg))
        self.status = status


class _Config(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("d", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("tol", ctypes.c_float), ("precision", ctypes.c_int32),
                ("update", ctypes.c_int32), ("device", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class _Result(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("converged_at", ctypes.c_uint64),
                ("inertia", ctypes.c_double), ("score", ctypes.c_double),
                ("max_shift", ctypes.c_double), ("refined_two", ctypes.c_uint64),
                ("refined_full", ctypes.c_uint64), ("ms_total", ctypes.c_double),
                ("ms_assign", ctypes.c_double), ("assign_launches", ctypes.c_uint64),
                ("ms_update", ctypes.c_double), ("update_launches", ctypes.c_uint64),
                ("full_passes", ctypes.c_uint64), ("moved_points", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class _SetupTimes(ctypes.Structure):
    _fields_ = [("memcpy_ms", ctypes.c_double), ("operands_ms", ctypes.c_double),
                ("init_ms", ctypes.c_double)]


class _StageTimes(ctypes.Structure):
    _fields_ = [("distances_ms", ctypes.c_float), ("argmin_ms", ctypes.c_float),
                ("centroids_ms", ctypes.c_float)]


def _stage_rows(fn, h, check):
    cnt = ctypes.c_uint64(0)
    check(fn(h, None, 0, ctypes.byref(cnt)), "stage_times")
    arr = (_StageTimes * max(1, cnt.value))()
    check(fn(h, arr, cnt.value, ctypes.byref(cnt)), "stage_times")
    return [(a.distances_ms, a.argmin_ms, a.centroids_ms) for a in arr[:cnt.value]]


MAX_VARIANTS = 16


class _Tuning(ctypes.Structure):
    _fields_ = [("nvariants", ctypes.c_int32), ("chosen", ctypes.c_int32),
                ("ms", ctypes.c_float * MAX_VARIANTS), ("threads", ctypes.c_int32 * MAX_VARIANTS),
                ("tile", ctypes.c_int32 * MAX_VARIANTS)]


class _KKConfig(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("d", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("tol", ctypes.c_float), ("kernel", ctypes.c_int32), ("store", ctypes.c_int32),
                ("gamma", ctypes.c_float), ("device", ctypes.c_int32),
                ("product", ctypes.c_int32), ("build", ctypes.c_int32)]


class _KKResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("converged_at", ctypes.c_uint64),
                ("objective", ctypes.c_double), ("score", ctypes.c_double),
                ("changed", ctypes.c_uint64), ("ms_total", ctypes.c_double),
                ("ms_build", ctypes.c_double), ("ms_spmm", ctypes.c_double),
                ("spmm_launches", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class _Problem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("d", ctypes.c_uint32), ("k", ctypes.c_uint32),
                ("X", ctypes.c_void_p), ("ld", ctypes.c_size_t), ("C0", ctypes.c_void_p)]


class _BatchConfig(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_float), ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class _BatchResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("converged_at", ctypes.c_uint64),
                ("inertia", ctypes.c_double)]


KERNELS = {"linear": 0, "polynomial": 1, "sigmoid": 2, "rbf": 3}
STORES = {"fp32": 0, "bf16": 1}
RUN_CONVERGE_SCORE = 0x4
RUN_EXCLUDE_EMPTY = 0x8
RUN_STOP_STABLE = 0x10
RUN_GRAPH = 0x20
KK_PRODUCTS = {"auto": 0, "dense": 1, "sorted": 2}
KBUILDS = {"fp32": 0, "bf16x3": 1, "bf16": 2}
RUN_CONVERGE_SHIFT_SUM = 0x40
RUN_COST_RATIO = 0x80
RUN_KEEP_EMPTY = 0x100
CONVERGE = {"shift": 0, "score": RUN_CONVERGE_SCORE, "shift_sum": RUN_CONVERGE_SHIFT_SUM}
EMPTY = {"zero": 0, "exclude": RUN_EXCLUDE_EMPTY, "keep": RUN_KEEP_EMPTY}

_lib = None


def load_library(path=LIB_PATH):
    """Load libkmc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise KmcError(ERR_STATE, "%s not built: run matrix-centric-k-means_amd/build.py" % path)
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    sigs = {
        "kmc_create": (ctypes.c_int, [P(_Config), P(vp)]),
        "kmc_destroy": (ctypes.c_int, [vp]),
        "kmc_set_points": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
        "kmc_set_points_device": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
        "kmc_set_centroids": (ctypes.c_int, [vp, vp]),
        "kmc_set_labels": (ctypes.c_int, [vp, vp]),
        "kmc_init_plusplus": (ctypes.c_int, [vp, ctypes.c_uint64, vp]),
        "kmc_init_plusplus_trials": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32, vp]),
        "kmc_run": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32, P(_Result)]),
        "kmc_get_labels": (ctypes.c_int, [vp, vp]),
        "kmc_get_centroids": (ctypes.c_int, [vp, vp]),
        "kmc_assign": (ctypes.c_int, [vp, vp, vp, P(_Result)]),
        "kmc_update": (ctypes.c_int, [vp, vp, vp, vp, vp, P(_Result)]),
        "kmc_last_error": (ctypes.c_char_p, [vp]),
        "kmc_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "kmc_device_count": (ctypes.c_int, []),
        "kmc_memory_footprint": (ctypes.c_int, [ctypes.c_int, P(ctypes.c_size_t)]),
        "kmc_autotune": (ctypes.c_int, [vp, P(_Tuning)]),
        "kmc_set_assign_variant": (ctypes.c_int, [vp, ctypes.c_int]),
        "kmc_set_prune_min": (ctypes.c_int, [vp, ctypes.c_int32]),
        "kmc_set_short_prune": (ctypes.c_int, [vp, ctypes.c_int32]),
        "kmc_kk_create": (ctypes.c_int, [P(_KKConfig), P(vp)]),
        "kmc_kk_destroy": (ctypes.c_int, [vp]),
        "kmc_kk_set_points": (ctypes.c_int, [vp, vp, ctypes.c_size_t]),
        "kmc_kk_set_labels": (ctypes.c_int, [vp, vp]),
        "kmc_kk_run": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32, P(_KKResult)]),
        "kmc_kk_get_labels": (ctypes.c_int, [vp, vp]),
        "kmc_kk_distances": (ctypes.c_int, [vp, vp, vp]),
        "kmc_kk_get_kernel_rows": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.c_size_t, vp]),
        "kmc_kk_last_error": (ctypes.c_char_p, [vp]),
        "kmc_batch_create": (ctypes.c_int, [P(_Problem), ctypes.c_uint32, P(_BatchConfig),
                                            P(vp)]),
        "kmc_batch_run": (ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_uint32, P(_BatchResult),
                                         P(ctypes.c_double)]),
        "kmc_batch_get_labels": (ctypes.c_int, [vp, ctypes.c_uint32, vp]),
        "kmc_batch_get_centroids": (ctypes.c_int, [vp, ctypes.c_uint32, vp]),
        "kmc_batch_last_error": (ctypes.c_char_p, [vp]),
        "kmc_batch_destroy": (ctypes.c_int, [vp]),
        "kmc_assign_static": (ctypes.c_int, [vp, ctypes.c_size_t, ctypes.c_uint32,
                                             ctypes.c_size_t, vp, ctypes.c_uint32, vp, vp]),
        "kmc_probe_mfma": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint32, vp, vp, vp, vp,
                                          vp]),
        "kmc_get_setup_times": (ctypes.c_int, [vp, P(_SetupTimes)]),
        "kmc_get_stage_times": (ctypes.c_int, [vp, vp, ctypes.c_uint64, P(ctypes.c_uint64)]),
        "kmc_kk_get_setup_times": (ctypes.c_int, [vp, P(_SetupTimes)]),
        "kmc_kk_get_stage_times": (ctypes.c_int, [vp, vp, ctypes.c_uint64,
                                                  P(ctypes.c_uint64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _status_text(s):
    try:
        return load_library().kmc_status_string(s).decode()
    except Exception:  # noqa: BLE001 - only used to format an error
        return "?"


def device_count():
    return load_library().kmc_device_count()


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


class Engine:
    """One k-means problem resident on one GPU."""

    def __init__(self, n, d, k, tol=1e-4, precision="fp32", device=0, update="ordered"):
        self._lib = load_library()
        self.n, self.d, self.k = int(n), int(d), int(k)
        self.precision = precision
        cfg = _Config(self.n, self.d, self.k, float(tol), PRECISIONS[precision], UPDATES[update],
                      int(device), 0)
        h = ctypes.c_void_p()
        st = self._lib.kmc_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != OK:
            raise KmcError(st, "kmc_create(n=%d, d=%d, k=%d, %s)" % (n, d, k, precision))
        self._h = h

    def _check(self, st, what):
        if st != OK:
            msg = self._lib.kmc_last_error(self._h)
            raise KmcError(st, "%s: %s" % (what, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_points(self, X):
        X = _f32(X, (self.n, self.d))
        self._check(self._lib.kmc_set_points(self._h, X.ctypes.data, self.d), "set_points")

    def set_points_device(self, ptr, ld=None):
        self._check(self._lib.kmc_set_points_device(self._h, ctypes.c_void_p(ptr),
                                                    self.d if ld is None else ld),
                    "set_points_device")

    def set_centroids(self, C0):
        C0 = _f32(C0, (self.k, self.d))
        self._check(self._lib.kmc_set_centroids(self._h, C0.ctypes.data), "set_centroids")

    def set_labels(self, labels):
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        self._check(self._lib.kmc_set_labels(self._h, lab.ctypes.data), "set_labels")

    def init_plusplus(self, seed, trials=1):
        """k-means++ seeding of the initial centroids; returns the chosen
        point indices. trials > 1: greedy k-means++ with that many D^2
        trials per round; 0: 2 + ceil(ln k) trials."""
        out = np.empty(self.k, dtype=np.int64)
        self._check(self._lib.kmc_init_plusplus_trials(self._h, int(seed) & ((1 << 64) - 1),
                                                       int(trials), out.ctypes.data),
                    "init_plusplus")
        return out

    def run(self, maxiter, check_converged=True, time_kernels=False, graph=False,
            converge="shift", cost_ratio=False, empty="zero"):
        """Lloyd iterations. converge: "shift" (max per-centroid shift <= tol,
        from iteration 2: the reference CPU k-means), "score" (|score change|
        < tol: the reference GPU engine) or "shift_sum" (sum of squared
        centroid changes < tol: the explicit-Lloyd benchmark); cost_ratio:
        also stop when cost_t / cost_{t-1} > 1 - tol; empty: what an empty
        cluster becomes ("zero" vector, "exclude"d from later assignments,
        or "keep" its centroid)."""
        flags = (RUN_CHECK_CONVERGED if check_converged else 0) | \
                (RUN_TIME_KERNELS if time_kernels else 0) | (RUN_GRAPH if graph else 0) | \
                CONVERGE[converge] | EMPTY[empty] | (RUN_COST_RATIO if cost_ratio else 0)
        r = _Result()
        self._check(self._lib.kmc_run(self._h, int(maxiter), flags, ctypes.byref(r)), "run")
        return r.as_dict()

    def setup_times(self):
        """Device ms of the H2D copy, the derived operands and the initial
        centroids (the reference's memcpy / b-mult / init_centroids timers)."""
        t = _SetupTimes()
        self._check(self._lib.kmc_get_setup_times(self._h, ctypes.byref(t)), "setup_times")
        return {"memcpy_ms": t.memcpy_ms, "operands_ms": t.operands_ms, "init_ms": t.init_ms}

    def stage_times(self):
        """Per-iteration (distances, argmin, centroids) ms of the last run
        with time_kernels=True."""
        return _stage_rows(self._lib.kmc_get_stage_times, self._h, self._check)

    def autotune(self):
        """Time every launch variant of the distance kernel on the current
        points and centroids and keep the fastest."""
        t = _Tuning()
        self._check(self._lib.kmc_autotune(self._h, ctypes.byref(t)), "autotune")
        n = t.nvariants
        return {"chosen": t.chosen, "ms": list(t.ms[:n]), "threads": list(t.threads[:n]),
                "tile": list(t.tile[:n])}

    def set_variant(self, v):
        self._check(self._lib.kmc_set_assign_variant(self._h, int(v)), "set_assign_variant")

    def set_prune_min(self, n):
        """All-k refinement queues of n or more points take the MFMA-pruned
        pass (1: always, 0: never; default 32768)."""
        self._check(self._lib.kmc_set_prune_min(self._h, int(n)), "set_prune_min")

    def set_short_prune(self, enable):
        """The all-k queue takes the single-sweep MFMA-pruned groups, in one
        launch with the two/three-candidate refinement (default on; K <= 512
        at D <= 64, K <= 256 at D <= 128); off: set_prune_min chooses between
        the two-sweep pruned pass and the exact pass."""
        self._check(self._lib.kmc_set_short_prune(self._h, int(bool(enable))),
                    "set_short_prune")

    def labels(self):
        out = np.empty(self.n, dtype=np.int32)
        self._check(self._lib.kmc_get_labels(self._h, out.ctypes.data), "get_labels")
        return out

    def centroids(self):
        out = np.empty((self.k, self.d), dtype=np.float32)
        self._check(self._lib.kmc_get_centroids(self._h, out.ctypes.data), "get_centroids")
        return out

    def assign(self, C):
        C = _f32(C, (self.k, self.d))
        out = np.empty(self.n, dtype=np.int32)
        r = _Result()
        self._check(self._lib.kmc_assign(self._h, C.ctypes.data, out.ctypes.data,
                                         ctypes.byref(r)), "assign")
        return out, r.as_dict()

    def update(self, labels, C_old):
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        C_old = _f32(C_old, (self.k, self.d))
        C_new = np.empty((self.k, self.d), dtype=np.float32)
        counts = np.empty(self.k, dtype=np.uint32)
        r = _Result()
        self._check(self._lib.kmc_update(self._h, lab.ctypes.data, C_old.ctypes.data,
                                         C_new.ctypes.data, counts.ctypes.data,
                                         ctypes.byref(r)), "update")
        return C_new, counts, r.as_dict()


def assign_static(x_ptr, n, d, ld, c_ptr, k, labels_ptr, stream=None):
    """kmc_assign_static on device pointers (allocation-free assignment;
    enqueues on `stream`, a hipStream_t as an int, or the null stream)."""
    st = load_library().kmc_assign_static(ctypes.c_void_p(x_ptr), int(n), int(d), int(ld),
                                           ctypes.c_void_p(c_ptr), int(k),
                                           ctypes.c_void_p(labels_ptr),
                                           ctypes.c_void_p(stream) if stream else None)
    if st != OK:
        raise KmcError(st, "kmc_assign_static(n=%d, d=%d, k=%d)" % (n, d, k))


def probe_mfma(precision, A, B):
    """kmc_probe_mfma: (out 32 x 32, Aq, Bq) for 32 x d operands A, B."""
    A = _f32(A)
    B = _f32(B)
    if A.shape != B.shape or A.shape[0] != 32:
        raise ValueError("A and B must both be 32 x d")
    d = A.shape[1]
    out = np.empty((32, 32), np.float32)
    Aq = np.empty_like(A)
    Bq = np.empty_like(B)
    st = load_library().kmc_probe_mfma(PRECISIONS[precision], d, A.ctypes.data, B.ctypes.data,
                                       out.ctypes.data, Aq.ctypes.data, Bq.ctypes.data)
    if st != OK:
        raise KmcError(st, "kmc_probe_mfma(%s, d=%d)" % (precision, d))
    return out, Aq, Bq


def fit(X, C0, maxiter, tol=1e-4, precision="fp32", check_converged=True, update="ordered"):
    """Run Lloyd from explicit centroids; returns (labels, centroids, result)."""
    X = _f32(X)
    n, d = X.shape
    C0 = _f32(C0)
    with Engine(n, d, C0.shape[0], tol=tol, precision=precision, update=update) as e:
        e.set_points(X)
        e.set_centroids(C0)
        res = e.run(maxiter, check_converged=check_converged)
        return e.labels(), e.centroids(), res


class KernelEngine:
    """Kernel k-means on one GPU (the reference's Popcorn path): the n x n
    kernel matrix K = f(X X^T) is built once on the GPU and streamed every
    iteration (src/kmeans.cu:71-305, 772-1122)."""

    def __init__(self, n, d, k, kernel="polynomial", store="fp32", gamma=0.0, tol=1e-4,
                 device=0, product="auto", build="fp32"):
        self._lib = load_library()
        self.n, self.d, self.k = int(n), int(d), int(k)
        cfg = _KKConfig(self.n, self.d, self.k, float(tol), KERNELS[kernel], STORES[store],
                        float(gamma), int(device), KK_PRODUCTS[product], KBUILDS[build])
        h = ctypes.c_void_p()
        st = self._lib.kmc_kk_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != OK:
            raise KmcError(st, "kmc_kk_create(n=%d, d=%d, k=%d, %s, %s)" % (n, d, k, kernel, store))
        self._h = h

    def _check(self, st, what):
        if st != OK:
            msg = self._lib.kmc_kk_last_error(self._h)
            raise KmcError(st, "%s: %s" % (what, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kmc_kk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_points(self, X):
        X = _f32(X, (self.n, self.d))
        self._check(self._lib.kmc_kk_set_points(self._h, X.ctypes.data, self.d), "kk_set_points")

    def set_labels(self, labels):
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        self._check(self._lib.kmc_kk_set_labels(self._h, lab.ctypes.data), "kk_set_labels")

    def run(self, maxiter, check_converged=False, stop_stable=False, time_kernels=False):
        flags = (RUN_CHECK_CONVERGED if check_converged else 0) | \
                (RUN_STOP_STABLE if stop_stable else 0) | (RUN_TIME_KERNELS if time_kernels else 0)
        r = _KKResult()
        self._check(self._lib.kmc_kk_run(self._h, int(maxiter), flags, ctypes.byref(r)), "kk_run")
        return r.as_dict()

    def setup_times(self):
        t = _SetupTimes()
        self._check(self._lib.kmc_kk_get_setup_times(self._h, ctypes.byref(t)), "setup_times")
        return {"memcpy_ms": t.memcpy_ms, "operands_ms": t.operands_ms, "init_ms": t.init_ms}

    def stage_times(self):
        return _stage_rows(self._lib.kmc_kk_get_stage_times, self._h, self._check)

    def labels(self):
        out = np.empty(self.n, dtype=np.int32)
        self._check(self._lib.kmc_kk_get_labels(self._h, out.ctypes.data), "kk_get_labels")
        return out

    def distances(self, labels):
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        D = np.empty((self.n, self.k), dtype=np.float32)
        self._check(self._lib.kmc_kk_distances(self._h, lab.ctypes.data, D.ctypes.data),
                    "kk_distances")
        return D

    def kernel_rows(self, row0, nrows):
        out = np.empty((nrows, self.n), dtype=np.float32)
        self._check(self._lib.kmc_kk_get_kernel_rows(self._h, row0, nrows, out.ctypes.data),
                    "kk_get_kernel_rows")
        return out


class Batch:
    """Many independent Lloyd problems of uneven n, d and k run together
    (kmc_batch_*): every stage of an iteration is one launch for the whole
    batch. problems: sequence of (X n x d, C0 k x d) float32 arrays. Each
    problem follows the reference CPU k-means with its own shift test; the
    centroids are exact fixed-point sums rounded once (as update="fixed")."""

    def __init__(self, problems, tol=1e-4, precision="fp32", device=0):
        self._lib = load_library()
        self._keep = []
        arr = (_Problem * len(problems))()
        self.shapes = []
        for i, (X, C0) in enumerate(problems):
            X = np.ascontiguousarray(X, dtype=np.float32)
            C0 = np.ascontiguousarray(C0, dtype=np.float32)
            if X.ndim != 2 or C0.ndim != 2 or C0.shape[1] != X.shape[1]:
                raise ValueError("problem %d: X n x d and C0 k x d expected" % i)
            self._keep += [X, C0]
            n, d = X.shape
            k = C0.shape[0]
            arr[i] = _Problem(n, d, k, X.ctypes.data, d, C0.ctypes.data)
            self.shapes.append((n, d, k))
        cfg = _BatchConfig(float(tol), PRECISIONS[precision], int(device), 0)
        h = ctypes.c_void_p()
        st = self._lib.kmc_batch_create(arr, len(problems), ctypes.byref(cfg), ctypes.byref(h))
        self._keep = []
        if st != OK:
            raise KmcError(st, "kmc_batch_create(%d problems)" % len(problems))
        self._h = h

    def _check(self, st, what):
        if st != OK:
            msg = self._lib.kmc_batch_last_error(self._h)
            raise KmcError(st, "%s: %s" % (what, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kmc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, maxiter, check_converged=True, graph=False):
        """Run every problem from its initial centroids; graph=True replays
        each chunk of iterations (all width groups) as one captured HIP graph."""
        res = (_BatchResult * len(self.shapes))()
        ms = ctypes.c_double(0.0)
        flags = (RUN_CHECK_CONVERGED if check_converged else 0) | (RUN_GRAPH if graph else 0)
        self._check(self._lib.kmc_batch_run(self._h, int(maxiter), flags, res,
                                            ctypes.byref(ms)), "batch_run")
        return [{"iterations": r.iterations, "converged_at": r.converged_at,
                 "inertia": r.inertia} for r in res], ms.value

    def labels(self, p):
        out = np.empty(self.shapes[p][0], dtype=np.int32)
        self._check(self._lib.kmc_batch_get_labels(self._h, p, out.ctypes.data), "batch_labels")
        return out

    def centroids(self, p):
        n, d, k = self.shapes[p]
        out = np.empty((k, d), dtype=np.float32)
        self._check(self._lib.kmc_batch_get_centroids(self._h, p, out.ctypes.data),
                    "batch_centroids")
        return out
