"""GPU parity at the benchmark configurations' own sizes.

  * kernel k-means at N=65 536, D=784 (sparse U[0,1] blobs, 70 % zeros), K in
    fp32 (17.2 GB) and bf16: with the linear kernel one iteration is a Lloyd
    step, checked against the oracle's FP64 Lloyd step; the polynomial and
    RBF kernels at N=8192, D=784 against the FP64 kernel-space distances
    (K64 from the FP64 Gram matrix and the kernels' closed forms), one
    iteration and a run to stability (same labels, same iteration count)
  * the 512-problem batch (18 M points): 16 problems against lone runs, every
    problem's FP64 inertia against the oracle
  * N=4M, D=256, K=4096: three iterations in every precision form agree bit
    for bit, and the first 100 000 labels are the reference CPU k-means'
  * N=1M, D=128, K=256: the exact fixed-point update against the
    reference's float summation order, iteration by iteration

Label bar: identical wherever the FP64 top-2 margin exceeds twice the stated
distance tolerance (kernel k-means: tests/test_gpu_kkmeans.py's bound), bit
for bit elsewhere (Lloyd forms).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dist_tol(n, store, scale):
    u_store = 0.0 if store == "fp32" else 2.0 ** -9
    return 4.0 * (n * 2.0 ** -24 + u_store) * scale


def _noisy_truth(truth, k, seed, frac=0.1):
    rng = np.random.default_rng(seed)
    lab = truth.astype(np.int32).copy()
    flip = rng.random(lab.size) < frac
    lab[flip] = rng.integers(0, k, size=int(flip.sum()))
    return lab


@pytest.fixture(scope="module")
def c3_data():
    from kmc import datagen
    X, truth = datagen.sparse_unit_blobs(65536, 784, 10)
    return X, _noisy_truth(truth, 10, 3)


@pytest.mark.parametrize("store", ["fp32", "bf16"])
def test_c3_linear_kernel_is_fp64_lloyd_step(kmc_mod, oracle, c3_data, store):
    X, lab0 = c3_data
    n, d, k = X.shape[0], X.shape[1], 10
    C, _ = oracle.f64_centroids(X, lab0.astype(np.uint32), k)
    want, margin, inert = oracle.f64_assign(X, C)
    scale = float((X.astype(np.float64) ** 2).sum(1).max())
    safe = margin > 2 * _dist_tol(n, store, scale)
    _say("C3 linear", store, "safe", safe.mean())
    # bf16 storage takes the bf16 hi+lo Gram build
    build = "fp32" if store == "fp32" else "bf16x3"
    with kmc_mod.KernelEngine(n, d, k, kernel="linear", store=store, build=build) as e:
        e.set_points(X)
        e.set_labels(lab0)
        res = e.run(1)
        lab = e.labels()
    assert safe.mean() > 0.95
    assert np.array_equal(lab[safe], want[safe].astype(np.int32)), int((lab != want)[safe].sum())
    if store == "fp32":
        # the objective of the new labels against the means of lab0 is the
        # FP64 inertia of that assignment (labels may part only on near-ties)
        assert abs(res["objective"] - inert) <= 1e-4 * inert, (res["objective"], inert)


def _kernel64(X, kind):
    X64 = X.astype(np.float64)
    G = X64 @ X64.T
    if kind == "polynomial":
        return (G + 1.0) ** 2
    nr = np.diag(G).copy()
    K = np.exp(-(1.0 / X.shape[1]) * np.maximum(nr[:, None] + nr[None, :] - 2.0 * G, 0.0))
    np.fill_diagonal(K, 1.0)
    return K


@pytest.fixture(scope="module")
def c3_small():
    from kmc import datagen
    X, truth = datagen.sparse_unit_blobs(8192, 784, 10)
    return X, _noisy_truth(truth, 10, 3)


@pytest.mark.parametrize("kernel", ["polynomial", "rbf"])
@pytest.mark.parametrize("store", ["fp32", "bf16"])
def test_c3_width_kernels_against_fp64(kmc_mod, oracle, c3_small, kernel, store):
    X, lab0 = c3_small
    n, d, k = X.shape[0], X.shape[1], 10
    K64 = _kernel64(X, kernel)
    scale = float(np.abs(K64).max())
    tol = _dist_tol(n, store, scale)
    # FP64 iterations to stability, noting whether every point of every
    # iteration is decided by more than twice the tolerance
    lab64 = lab0.astype(np.uint32).copy()
    all_safe, it64 = True, 0
    first = None
    while it64 < 50:
        D64 = oracle.kk_distances(K64, lab64, k, add_diag=True)
        srt = np.sort(D64, axis=1)
        safe = srt[:, 1] - srt[:, 0] > 2 * tol
        new = np.argmin(D64, axis=1).astype(np.uint32)
        if first is None:
            first = (new.copy(), safe.copy())
        all_safe &= bool(safe.all())
        it64 += 1
        changed = np.any(new != lab64)
        lab64 = new
        if not changed:
            break
    _say("C3 width", kernel, store, "safe first", first[1].mean(), "all safe", all_safe,
         "iterations", it64)
    with kmc_mod.KernelEngine(n, d, k, kernel=kernel, store=store,
                              gamma=1.0 / d if kernel == "rbf" else 0.0) as e:
        e.set_points(X)
        e.set_labels(lab0)
        e.run(1)
        lab1 = e.labels()
        e.set_labels(lab0)
        res = e.run(50, check_converged=True, stop_stable=True)
        lab = e.labels()
    want1, safe1 = first
    assert np.array_equal(lab1[safe1], want1[safe1].astype(np.int32))
    # the bf16 bound is the worst case of every stored entry's rounding: it
    # decides most polynomial points and a sixth of the RBF points
    floor = 0.99 if store == "fp32" else {"polynomial": 0.9, "rbf": 0.15}[kernel]
    assert safe1.mean() > floor, safe1.mean()
    if all_safe:
        assert res["iterations"] == it64, (res, it64)
        assert np.array_equal(lab, lab64.astype(np.int32))


@pytest.fixture(scope="module")
def c4_batch():
    from kmc import datagen
    return datagen.batch_problems(512)


def test_c4_batch_of_512(kmc_mod, oracle, c4_batch):
    probs = c4_batch
    steps = 30
    _say("C4 batch", len(probs), "problems", sum(X.shape[0] for X, _ in probs), "points")
    with kmc_mod.Batch(probs, tol=0.0, precision="bf16") as b:
        res, ms = b.run(steps, check_converged=False)
        _say("C4 batch device ms", ms)
        assert all(int(r["iterations"]) == steps for r in res)
        for p, (X, C0) in enumerate(probs):
            want = oracle.f64_inertia(X, b.centroids(p), b.labels(p).astype(np.uint32))
            assert abs(res[p]["inertia"] - want) <= 1e-9 * max(want, 1e-30), (p, res[p], want)
        for p in range(0, 512, 32):
            X, C0 = probs[p]
            with kmc_mod.Engine(X.shape[0], X.shape[1], C0.shape[0], tol=0.0, precision="bf16",
                                update="fixed") as e:
                e.set_points(X)
                e.set_centroids(C0)
                e.run(steps, check_converged=False)
                assert np.array_equal(b.labels(p), e.labels()), p
                assert np.array_equal(b.centroids(p), e.centroids()), p


def _reference_labels(oracle, X, C, rows=4096, margin_frac=1e-3):
    """Reference CPU k-means labels: FP64 argmin where the FP64 top-2 gap is
    above 1e-3 of the distance (the reference's float error is below 2e-5
    relative there), the oracle's float restatement elsewhere."""
    Cd = C.astype(np.float64)
    cn = (Cd * Cd).sum(1)
    out = np.empty(X.shape[0], np.int32)
    for s in range(0, X.shape[0], rows):
        Xs = X[s:s + rows]
        Xd = Xs.astype(np.float64)
        Dd = (Xd * Xd).sum(1)[:, None] + cn[None, :] - 2.0 * (Xd @ Cd.T)
        part = np.partition(Dd, 1, axis=1)
        lab = np.argmin(Dd, axis=1)
        close = np.nonzero(part[:, 1] - part[:, 0] <= margin_frac * np.maximum(part[:, 1], 1.0))[0]
        if close.size:
            lab[close], _ = oracle.argmin(oracle.distances(np.ascontiguousarray(Xs[close]), C))
        out[s:s + rows] = lab
    return out


def test_c5_full_size_forms_agree(kmc_mod, oracle):
    from kmc import datagen
    X, _, _ = datagen.huge_blobs()
    n, d, k = X.shape[0], X.shape[1], 4096
    C0 = datagen.init_rows(X, k, 7)
    out = {}
    for prec in ("fp32", "bf16x3", "bf16", "fp8"):
        _say("C5 full size", prec)
        with kmc_mod.Engine(n, d, k, tol=0.0, precision=prec, update="tree") as e:
            e.set_points(X)
            e.set_centroids(C0)
            if prec == "fp32":
                e.run(2, check_converged=False)
                C2 = e.centroids()
                e.set_centroids(C0)
            r = e.run(3, check_converged=False, graph=True)
            out[prec] = (e.labels(), e.centroids(), r)
    for prec in ("bf16x3", "bf16", "fp8"):
        assert np.array_equal(out[prec][0], out["fp32"][0]), prec
        assert np.array_equal(out[prec][1], out["fp32"][1]), prec
    m = 100_000
    want = _reference_labels(oracle, X[:m], C2)
    assert np.array_equal(out["fp32"][0][:m], want), int((out["fp32"][0][:m] != want).sum())


def test_c2_fixed_update_against_reference_order(kmc_mod, oracle):
    """The fixed-point update (exact sums rounded once: the bench's) against
    the ordered update (the reference's float sums in point order) at N=1M,
    D=128, K=256 from the same seeded rows.

    Step by step (both updates fed the same centroids each iteration, along
    the reference-order trajectory, iterations 1..20): identical labels and
    centroids within 1e-5 relative per centroid row. Free running (each
    update's own trajectory): the first iteration at which the labels part
    is reported, and every parting point must sit on a tie that the two
    centroid sets' difference can move."""
    from kmc import datagen
    X, _, _ = datagen.large_blobs()
    n, d, k = X.shape[0], X.shape[1], 256
    C0 = datagen.init_rows(X, k, 7)
    with kmc_mod.Engine(n, d, k, tol=0.0, precision="bf16", update="fixed") as ef, \
            kmc_mod.Engine(n, d, k, tol=0.0, precision="bf16", update="ordered") as eo:
        for e in (ef, eo):
            e.set_points(X)
        C = C0
        worst = 0.0
        for t in range(1, 21):
            for e in (ef, eo):
                e.set_centroids(C)
                e.run(1, check_converged=False)
            Cf, Co = ef.centroids(), eo.centroids()
            assert np.array_equal(ef.labels(), eo.labels()), t
            dC = np.linalg.norm(Cf.astype(np.float64) - Co, axis=1)
            row = float(np.max(dC / np.maximum(np.linalg.norm(Co.astype(np.float64), axis=1),
                                              1e-30)))
            worst = max(worst, row)
            assert row <= 1e-5, (t, row)
            C = Co
        _say("step by step: labels identical, max centroid row rel diff %.3g" % worst)

        Cf_prev = Co_prev = C0
        for t in range(1, 21):
            for e in (ef, eo):
                e.set_centroids(C0)
                e.run(t, check_converged=False, graph=True)
            lf, lo = ef.labels(), eo.labels()
            diff = np.nonzero(lf != lo)[0]
            if diff.size:
                # tie test: the FP64 margin against the ordered set is below
                # what the difference delta of the two sets can shift, summed
                # over the two candidates: 2 ||x - c|| delta_c + delta_c^2
                Xd = X[diff].astype(np.float64)
                delta = np.linalg.norm(Cf_prev.astype(np.float64) - Co_prev, axis=1)
                _, margin, _ = oracle.f64_assign(np.ascontiguousarray(X[diff]), Co_prev)
                bound = np.zeros(diff.size)
                for lab in (lf[diff], lo[diff]):
                    dist = np.linalg.norm(Xd - Co_prev[lab].astype(np.float64), axis=1)
                    bound += 2.0 * dist * delta[lab] + delta[lab] ** 2
                bound += 1e-6 * np.linalg.norm(Xd, axis=1) ** 2  # float distance rounding
                _say("free running: labels part at iteration", t, "points", diff[:20].tolist(),
                     "FP64 margins", margin[:20].tolist(), "bounds", bound[:20].tolist())
                assert diff.size <= 64, (t, diff.size)
                assert np.all(margin <= bound), (t, margin, bound)
                break
            Cf_prev, Co_prev = ef.centroids(), eo.centroids()
        else:
            _say("free running: labels identical for 20 iterations")
