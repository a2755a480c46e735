"""kmc_assign_static (the allocation-free assignment) through the C ABI,
against kmc_assign on a handle: bit-identical labels on two calls in a row,
the second smaller, plus the reference CPU k-means' labels (oracle) on a
slice of the points, and the size checks. The caller's device buffers are
allocated here with the HIP runtime the library itself is linked to."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip(kmc_mod):
    # the library's own libamdhip64.so.7 (already loaded by kmc_mod)
    L = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p
    L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    L.hipFree.argtypes = [vp]
    L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class _Dev:
    """A device copy of a host array (freed on exit)."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(1, self.a.nbytes)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *e):
        self.hip.hipFree(self.p)


def _blobs(seed, n, d, k):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-10, 10, (k, d))
    sig = rng.uniform(0.5, 2.0, k)
    lab = rng.integers(0, k, n)
    X = (cen[lab] + sig[lab, None] * rng.standard_normal((n, d))).astype(np.float32)
    C = X[rng.choice(n, k, replace=False)].copy()
    return X, C


def _static_labels(kmc_mod, hip, X, C, ld):
    n, d = X.shape
    k = C.shape[0]
    Xp = np.zeros((n, ld), np.float32)
    Xp[:, :d] = X
    Cp = np.zeros((k, ld), np.float32)
    Cp[:, :d] = C
    with _Dev(hip, Xp) as dX, _Dev(hip, Cp) as dC, \
            _Dev(hip, np.full(n, -1, np.int32)) as dL:
        kmc_mod.assign_static(dX.p.value, n, d, ld, dC.p.value, k, dL.p.value)
        assert hip.hipDeviceSynchronize() == 0
        return dL.get()


def _handle_labels(kmc_mod, X, C):
    with kmc_mod.Engine(X.shape[0], X.shape[1], C.shape[0], precision="fp32") as e:
        e.set_points(X)
        lab, _ = e.assign(C)
    return lab


def test_static_assign_equals_handle_assign(kmc_mod, oracle, hip):
    cases = [(1, 300000, 100, 256, 128), (2, 60000, 30, 50, 32)]
    for seed, n, d, k, ld in cases:
        _say("kmc_assign_static n=%d d=%d k=%d ld=%d" % (n, d, k, ld))
        X, C = _blobs(seed, n, d, k)
        got = _static_labels(kmc_mod, hip, X, C, ld)
        want = _handle_labels(kmc_mod, X, C)
        assert np.array_equal(got, want), int((got != want).sum())
        ref, _ = oracle.argmin(oracle.distances(X[:3000], C))
        assert np.array_equal(got[:3000], ref.astype(np.int32))


def test_static_assign_limits(kmc_mod, hip):
    lib = kmc_mod.load_library()
    with _Dev(hip, np.zeros(64, np.float32)) as t, _Dev(hip, np.zeros(4, np.int32)) as lab:
        # ld must be d rounded up to 16/32/64/128/256
        assert lib.kmc_assign_static(t.p, 4, 10, 10, t.p, 2, lab.p, None) == kmc_mod.ERR_ARG
        # beyond the fixed scratch array: refused before any work is queued
        assert lib.kmc_assign_static(t.p, (1 << 19) + 1, 16, 16, t.p, 2, lab.p,
                                     None) == kmc_mod.ERR_UNSUPPORTED
        assert lib.kmc_assign_static(t.p, 4096, 16, 16, t.p, 1025, lab.p,
                                     None) == kmc_mod.ERR_UNSUPPORTED
        assert hip.hipDeviceSynchronize() == 0
