"""The C-ABI library loads on a host without a GPU and exports every entry
point that include/kmc.h declares; host-side argument checks and error paths
(no compute is launched)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kmc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kmc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for want in ("kmc_create", "kmc_set_points", "kmc_run", "kmc_get_labels",
                 "kmc_assign", "kmc_update", "kmc_destroy"):
        assert want in names


def test_library_exports_every_declared_symbol(kmc_mod):
    lib = ctypes.CDLL(kmc_mod.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings(kmc_mod):
    lib = kmc_mod.load_library()
    assert lib.kmc_status_string(0) == b"ok"
    assert lib.kmc_status_string(8) == b"unsupported configuration"
    assert lib.kmc_status_string(99) == b"unknown status"


def test_create_rejects_bad_config(kmc_mod):
    lib = kmc_mod.load_library()
    h = ctypes.c_void_p()
    assert lib.kmc_create(None, ctypes.byref(h)) == kmc_mod.ERR_ARG
    for n, d, k, prec in [(0, 4, 2, 0), (10, 0, 2, 0), (10, 4, 0, 0), (10, 4, 11, 0),
                          (10, 4, 2, 7)]:
        cfg = kmc_mod._Config(n, d, k, 1e-4, prec, 0, 0, 0)
        assert lib.kmc_create(ctypes.byref(cfg), ctypes.byref(h)) == kmc_mod.ERR_ARG
    cfg = kmc_mod._Config(10, 4, 2, -1.0, 0, 0, 0, 0)
    assert lib.kmc_create(ctypes.byref(cfg), ctypes.byref(h)) == kmc_mod.ERR_ARG
    cfg = kmc_mod._Config(100, 300, 2, 1e-4, 0, 0, 0, 0)  # d beyond the fused kernel
    assert lib.kmc_create(ctypes.byref(cfg), ctypes.byref(h)) == kmc_mod.ERR_UNSUPPORTED


def test_no_device_is_reported(kmc_mod):
    if kmc_mod.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(kmc_mod.KmcError) as ei:
        kmc_mod.Engine(100, 8, 4)
    assert ei.value.status == kmc_mod.ERR_NO_DEVICE


def test_batch_create_rejects_bad_problems(kmc_mod):
    lib = kmc_mod.load_library()
    X = np.zeros((100, 8), np.float32)
    C = np.zeros((4, 8), np.float32)
    P = kmc_mod._Problem
    cfg = kmc_mod._BatchConfig(1e-4, 0, 0, 0)
    h = ctypes.c_void_p()
    empty = (P * 1)()
    assert lib.kmc_batch_create(empty, 0, ctypes.byref(cfg), ctypes.byref(h)) == kmc_mod.ERR_ARG
    cases = [
        P(0, 8, 4, X.ctypes.data, 8, C.ctypes.data),      # n = 0
        P(100, 8, 0, X.ctypes.data, 8, C.ctypes.data),    # k = 0
        P(100, 8, 101, X.ctypes.data, 8, C.ctypes.data),  # k > n
        P(100, 8, 4, None, 8, C.ctypes.data),             # no points
        P(100, 8, 4, X.ctypes.data, 4, C.ctypes.data),    # ld < d
    ]
    for c in cases:
        arr = (P * 1)(c)
        assert lib.kmc_batch_create(arr, 1, ctypes.byref(cfg), ctypes.byref(h)) == kmc_mod.ERR_ARG
    big = (P * 1)(P(100, 300, 4, np.zeros((100, 300), np.float32).ctypes.data, 300,
                    np.zeros((4, 300), np.float32).ctypes.data))
    assert lib.kmc_batch_create(big, 1, ctypes.byref(cfg), ctypes.byref(h)) in (
        kmc_mod.ERR_UNSUPPORTED,)
    bad_prec = kmc_mod._BatchConfig(1e-4, 7, 0, 0)
    ok = (P * 1)(P(100, 8, 4, X.ctypes.data, 8, C.ctypes.data))
    assert lib.kmc_batch_create(ok, 1, ctypes.byref(bad_prec), ctypes.byref(h)) == kmc_mod.ERR_ARG


def test_batch_without_device(kmc_mod):
    if kmc_mod.device_count() > 0:
        pytest.skip("a device is visible")
    X = np.zeros((100, 8), np.float32)
    C = np.zeros((4, 8), np.float32)
    with pytest.raises(kmc_mod.KmcError) as ei:
        kmc_mod.Batch([(X, C)])
    assert ei.value.status == kmc_mod.ERR_NO_DEVICE


def test_kernel_kmeans_config_checks(kmc_mod):
    """Kernel k-means: product selection and the cluster limits of each
    product are checked before any device call."""
    lib = kmc_mod.load_library()
    h = ctypes.c_void_p()

    def cfg(k, product):
        return kmc_mod._KKConfig(5000, 8, k, 1e-4, 1, 0, 0.0, 0, product)

    assert lib.kmc_kk_create(ctypes.byref(cfg(10, 3)), ctypes.byref(h)) == kmc_mod.ERR_ARG
    assert lib.kmc_kk_create(ctypes.byref(cfg(65, kmc_mod.KK_PRODUCTS["dense"])),
                             ctypes.byref(h)) == kmc_mod.ERR_UNSUPPORTED
    assert lib.kmc_kk_create(ctypes.byref(cfg(1025, kmc_mod.KK_PRODUCTS["sorted"])),
                             ctypes.byref(h)) == kmc_mod.ERR_UNSUPPORTED


def test_batch_rejects_fp8(kmc_mod):
    lib = kmc_mod.load_library()
    X = np.zeros((100, 4), np.float32)
    C0 = np.zeros((2, 4), np.float32)
    arr = (kmc_mod._Problem * 1)()
    arr[0] = kmc_mod._Problem(100, 4, 2, X.ctypes.data, 4, C0.ctypes.data)
    bc = kmc_mod._BatchConfig(1e-4, kmc_mod.PRECISIONS["fp8"], 0, 0)
    h = ctypes.c_void_p()
    assert lib.kmc_batch_create(arr, 1, ctypes.byref(bc), ctypes.byref(h)) == \
        kmc_mod.ERR_UNSUPPORTED


def test_bench_variant_table_matches_the_kernel_source():
    """bench.py names the distance kernel instantiation it timed from its copy
    of the launch-variant table; the copy must match kVariants in the HIP
    source (same shapes, same order as kmc_autotune reports them)."""
    src = open(os.path.join(ROOT, "matrix-centric-k-means_amd", "csrc", "kmc_lloyd.hip")).read()
    body = re.search(r"kVariants\[\]\s*=\s*\{(.*?)\};", src, flags=re.S).group(1)
    shapes = [(int(a), int(b), c == "true", int(d), int(e), int(f), int(g) if g else 1)
              for a, b, c, d, e, f, g in re.findall(
                  r"\{\s*(\d+),\s*(\d+),\s*(true|false),\s*(\d+),\s*(\d+),\s*(\d+)"
                  r"(?:,\s*(\d+))?\s*\}", body)]
    import sys
    sys.path.insert(0, ROOT)
    import bench
    assert shapes == bench.ASSIGN_VARIANTS
    assert bench.assign_kernel_name(128, "bf16", 9) == "k_assign2<128, 2, 8, 3, false, true, 0, 2>"
    assert bench.assign_kernel_name(128, "bf16", 1) == "k_assign<128, 2, 8, 1, false, 0, false>"
    assert bench.assign_kernel_name(128, "bf16", 8) == "k_assign2<128, 2, 8, 4, false, true, 0, 1>"


def test_fixed_point_double_add_conversion():
    """The fixed-point update rounds v = x * scale (|v| < 2^50) to a 64-bit
    integer as the bits of the double v + 1.5 * 2^52 less those of 1.5 * 2^52
    (fx_int in kmc_lloyd.hip); it must equal round-half-even of v, the
    general conversion (__float2ll_rn) the moved-point atomics use."""
    rng = np.random.default_rng(11)
    e = rng.integers(-30, 50, size=200_000)
    v = (rng.standard_normal(200_000) * np.exp2(e.astype(np.float64))).astype(np.float32)
    v = v[np.abs(v.astype(np.float64)) < 2.0 ** 50]
    ties = np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.0, -0.0, 2.0 ** 49, -(2.0 ** 49),
                     8388607.5, -8388607.5, 2.0 ** 50 - 2.0 ** 27], dtype=np.float32)
    v = np.concatenate([v, ties])
    magic = np.float64(1.5 * 2.0 ** 52)
    got = (v.astype(np.float64) + magic).view(np.int64) - np.int64(0x4338000000000000)
    want = np.rint(v.astype(np.float64)).astype(np.int64)  # half to even
    assert np.array_equal(got, want)
