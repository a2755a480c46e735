"""GPU parity of the linear Lloyd path through the C ABI (libkmc.so).

Bar: labels and centroids bit-identical to the reference CPU k-means
(comparisons/CPU_Kmeans/src/kmeans.cpp) for every precision form, because the
margin band sends every point whose order could differ to the reference's own
float arithmetic; inertia within 1e-6 relative of the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

PRECS = ("fp32", "bf16x3", "bf16", "fp8")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(kmc_mod, X, k, prec, tol=1e-4):
    e = kmc_mod.Engine(X.shape[0], X.shape[1], k, tol=tol, precision=prec)
    e.set_points(X)
    return e


@pytest.mark.parametrize("prec", PRECS)
def test_assign_update_golden_steps(kmc_mod, oracle, prec):
    g = load_golden("ref_steps.npz")
    for i in range(int(g["ncases"])):
        X, C = g["%d_X" % i], g["%d_C" % i]
        k = C.shape[0]
        _say("steps case", i, X.shape, "k", k, prec)
        with _engine(kmc_mod, X, k, prec) as e:
            lab, info = e.assign(C)
            assert np.array_equal(lab, g["%d_labels" % i].astype(np.int32)), (i, info)
            Cn, cnt, res = e.update(g["%d_labels" % i].astype(np.int32), C)
            assert np.array_equal(cnt, g["%d_counts" % i]), i
            assert np.array_equal(Cn, g["%d_Cnew" % i]), i
            ref_in = oracle.f64_inertia(X, C, g["%d_labels" % i])
            assert abs(res["inertia"] - ref_in) <= 1e-6 * max(ref_in, 1e-30), (res, ref_in)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", ["ref_c1_blobs.npz", "ref_digits.npz"])
def test_run_golden_full(kmc_mod, oracle, prec, name):
    g = load_golden(name)
    X, C0 = g["X"], g["C0"]
    _say("full run", name, prec)
    with _engine(kmc_mod, X, int(g["k"]), prec, tol=float(g["tol"])) as e:
        e.set_centroids(C0)
        res = e.run(int(g["maxiter"]), check_converged=True)
        assert res["converged_at"] == int(g["converged"])
        assert np.array_equal(e.labels(), g["labels"].astype(np.int32))
        assert np.array_equal(e.centroids(), g["C"])


@pytest.mark.parametrize("prec", PRECS)
def test_run_golden_small(kmc_mod, prec):
    g = load_golden("ref_runs.npz")
    for i in range(int(g["ncases"])):
        X, C0 = g["%d_X" % i], g["%d_C0" % i]
        _say("run case", i, X.shape, prec)
        with _engine(kmc_mod, X, C0.shape[0], prec, tol=float(g["%d_tol" % i])) as e:
            e.set_centroids(C0)
            res = e.run(int(g["%d_maxiter" % i]), check_converged=True)
            assert res["converged_at"] == int(g["%d_converged" % i]), (i, res)
            assert np.array_equal(e.labels(), g["%d_labels" % i].astype(np.int32)), i
            assert np.array_equal(e.centroids(), g["%d_C" % i]), i


@pytest.mark.parametrize("update", ["tree", "fixed"])
def test_update_tree_matches_fp64_means(kmc_mod, oracle, update):
    """Tree mode (FP64 sums rounded once) and fixed mode (exact fixed-point
    sums rounded once): within 1 ulp of the FP64 means; FP64 inertia."""
    g = load_golden("ref_steps.npz")
    for i in range(int(g["ncases"])):
        X, C = g["%d_X" % i], g["%d_C" % i]
        lab = g["%d_labels" % i].astype(np.int32)
        k = C.shape[0]
        _say(update, "update case", i, X.shape, k)
        with kmc_mod.Engine(X.shape[0], X.shape[1], k, update=update) as e:
            e.set_points(X)
            Cn, cnt, res = e.update(lab, C)
        C64, cnt64 = oracle.f64_centroids(X, lab.astype(np.uint32), k)
        assert np.array_equal(cnt, cnt64)
        assert np.allclose(Cn, C64, rtol=2.4e-7, atol=1e-30), np.max(np.abs(Cn - C64))
        ref_in = oracle.f64_inertia(X, C, lab.astype(np.uint32))
        assert abs(res["inertia"] - ref_in) <= 1e-9 * max(ref_in, 1e-30)


@pytest.mark.parametrize("update", ["tree", "fixed"])
@pytest.mark.parametrize("prec", PRECS)
def test_run_tree_mode_golden(kmc_mod, prec, update):
    g = load_golden("ref_c1_blobs.npz")
    with kmc_mod.Engine(g["X"].shape[0], 64, 10, tol=float(g["tol"]), precision=prec,
                        update=update) as e:
        e.set_points(g["X"])
        e.set_centroids(g["C0"])
        res = e.run(300, check_converged=True)
        assert abs(int(res["converged_at"]) - int(g["converged"])) <= 1
        assert np.array_equal(e.labels(), g["labels"].astype(np.int32))
        assert np.allclose(e.centroids(), g["C"], rtol=1e-6, atol=1e-6)


@pytest.fixture(scope="module")
def c2_data():
    from kmc import datagen
    X, _, _ = datagen.large_blobs()
    C0 = datagen.init_rows(X, 256, 7)
    return X, C0


def _reference_labels(oracle, X, C, margin_frac=1e-3):
    """Reference FP32 first-min labels for every point: FP64 distances decide
    wherever the FP64 top-2 gap is large; the rest are recomputed with the
    oracle's float restatement of the reference distance."""
    Xd, Cd = X.astype(np.float64), C.astype(np.float64)
    D = (Xd * Xd).sum(1)[:, None] + (Cd * Cd).sum(1)[None, :] - 2.0 * (Xd @ Cd.T)
    part = np.partition(D, 1, axis=1)
    lab = np.argmin(D, axis=1)
    gap = part[:, 1] - part[:, 0]
    close = np.nonzero(gap <= margin_frac * np.maximum(part[:, 1], 1.0))[0]
    if close.size:
        Dc = oracle.distances(X[close], C)
        lab[close], _ = oracle.argmin(Dc)
    return lab.astype(np.int32), close.size


@pytest.mark.parametrize("prec", PRECS)
def test_full_size_assign_exact(kmc_mod, oracle, c2_data, prec):
    """N=1M, D=128, K=256: one assignment equals the reference's FP32 labels."""
    X, C0 = c2_data
    # centroids after a few iterations: blobs split between centroids give
    # plenty of small margins
    with kmc_mod.Engine(X.shape[0], 128, 256, precision="fp32") as e:
        e.set_points(X)
        e.set_centroids(C0)
        e.run(3, check_converged=False)
        C = e.centroids()
    want, nclose = _reference_labels(oracle, X, C)
    _say("full-size assign", prec, "close points", nclose)
    with kmc_mod.Engine(X.shape[0], 128, 256, precision=prec) as e:
        e.set_points(X)
        lab, info = e.assign(C)
    _say("refined", info["refined_two"], info["refined_full"])
    assert np.array_equal(lab, want)


def test_full_size_update_exact(kmc_mod, oracle, c2_data):
    X, C0 = c2_data
    rng = np.random.default_rng(3)
    lab = rng.integers(0, 256, size=X.shape[0]).astype(np.int32)
    lab[lab == 17] = 18  # one empty cluster -> zero vector
    with kmc_mod.Engine(X.shape[0], 128, 256) as e:
        e.set_points(X)
        Cn, cnt, res = e.update(lab, C0)
    Co, cnto = oracle.centroids(X, lab.astype(np.uint32), 256)
    assert np.array_equal(cnt, cnto)
    assert np.array_equal(Cn, Co)
    assert not Cn[17].any()
    ref_in = oracle.f64_inertia(X, C0, lab.astype(np.uint32))
    assert abs(res["inertia"] - ref_in) <= 1e-9 * ref_in


@pytest.mark.parametrize("prec", PRECS)
def test_every_launch_variant_gives_the_same_labels(kmc_mod, oracle, c2_data, prec):
    """Every k_assign launch variant built for the problem (and the
    autotuner's choice) reproduces the reference's labels on a 200K-point
    slice."""
    X, C0 = c2_data
    Xs = np.ascontiguousarray(X[:200_000])
    want, _ = _reference_labels(oracle, Xs, C0)
    with kmc_mod.Engine(Xs.shape[0], 128, 256, precision=prec) as e:
        e.set_points(Xs)
        e.set_centroids(C0)
        t = e.autotune()
        _say("variants", prec, t)
        for v in range(len(t["ms"])):
            if t["ms"][v] < 0:  # variant not built for this (D, precision)
                continue
            e.set_variant(v)
            lab, _ = e.assign(C0)
            assert np.array_equal(lab, want), v


@pytest.mark.parametrize("prec", ("bf16", "fp8"))
@pytest.mark.parametrize("d,k", [(20, 40), (64, 130), (100, 300)])
def test_every_launch_variant_on_other_shapes(kmc_mod, oracle, prec, d, k):
    """Every launch variant offered for a shape -- the two-set kernel's
    ranking-list forms included, over padded widths (DP 32/64/128, fp8 at
    least 64) and centroid counts that are not a multiple of 64 -- gives the
    reference's labels; half the centroids are pulled towards a neighbour so
    that many points fall in the margin band."""
    from kmc import datagen
    X, _, _ = datagen.blobs(20_000, d, k, 20261204 + d, -4.0, 4.0, 0.6, 1.4)
    C = datagen.init_rows(X, k, 11).astype(np.float32)
    C[1::2] = (0.8 * C[1::2] + 0.2 * C[0::2]).astype(np.float32)
    want, nclose = _reference_labels(oracle, X, C)
    with kmc_mod.Engine(X.shape[0], d, k, precision=prec) as e:
        e.set_points(X)
        e.set_centroids(C)
        t = e.autotune()
        _say("shape", d, k, prec, "close", nclose, t["ms"])
        offered = [v for v in range(len(t["ms"])) if t["ms"][v] >= 0]
        assert offered
        for v in offered:
            e.set_variant(v)
            lab, info = e.assign(C)
            assert np.array_equal(lab, want), (v, int(np.sum(lab != want)))


@pytest.mark.parametrize("prec", PRECS)
def test_pruned_all_k_pass_gives_the_same_labels(kmc_mod, oracle, c2_data, prec):
    """The all-k refinement with and without its bf16x3 MFMA pruning stage
    (in both its forms) reproduces the reference's labels on a 200K-point slice; centroids
    perturbed towards each other put many points in the band."""
    X, C0 = c2_data
    Xs = np.ascontiguousarray(X[:200_000])
    C = C0.copy()
    C[1::2] = (0.9 * C0[1::2] + 0.1 * C0[0::2]).astype(np.float32)  # pairs closer together
    want, _ = _reference_labels(oracle, Xs, C)
    with kmc_mod.Engine(Xs.shape[0], 128, 256, precision=prec) as e:
        e.set_points(Xs)
        # the single-sweep groups (the default at this shape), then without
        # them: the two-sweep pruned pass always, from the default threshold,
        # never
        for short, pm in ((True, 32768), (False, 1), (False, 32768), (False, 0)):
            e.set_short_prune(short)
            e.set_prune_min(pm)
            lab, info = e.assign(C)
            _say("prune_min", pm, short, prec, "refined", info["refined_two"],
                 info["refined_full"])
            assert np.array_equal(lab, want), (short, pm)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_fixed_update_from_moved_points_equals_full_passes(kmc_mod, c2_data, prec, graph):
    """The fixed-point update keeps exact cluster totals across the
    iterations of a run and, once few points move, adds and subtracts only
    the moved points' rows. Single-iteration runs (each starts with a full
    pass) must give the same labels, centroid bits and inertia as one run
    over the same iterations. Only the first iteration of the run streams
    every point; an iteration after one in which more than n/64 points moved
    (kFxMovedDiv) takes the moved points through the workgroups' LDS
    accumulators, the others by global atomics -- both forms must occur
    here."""
    X, C0 = c2_data
    Xs = np.ascontiguousarray(X[:200_000])
    n = Xs.shape[0]
    iters = 12
    with kmc_mod.Engine(n, 128, 256, tol=0.0, precision=prec, update="fixed") as e:
        e.set_points(Xs)
        e.set_centroids(C0)
        r = e.run(iters, check_converged=False, graph=graph)
        C_run, L_run = e.centroids(), e.labels()
        e.set_centroids(C0)
        labs = []
        for _ in range(iters):
            r1 = e.run(1, check_converged=False)
            labs.append(e.labels())
        # moved[t]: points whose label changed from iteration t to t + 1; the
        # update of iteration t + 2 chooses its form from it
        moved = [int(np.count_nonzero(labs[t] != labs[t - 1])) for t in range(1, iters - 1)]
        _say("fixed", prec, graph, r["inertia"], r1["inertia"], "moved", moved,
             "full passes", r["full_passes"], "moved points", r["moved_points"])
        assert np.array_equal(e.labels(), L_run)
        assert np.array_equal(e.centroids(), C_run)
        assert r1["inertia"] == r["inertia"]
        assert r["full_passes"] == 1
        assert any(m > n // 64 for m in moved) and any(0 < m <= n // 64 for m in moved)


@pytest.mark.parametrize("n,d,k", [(1001, 20, 7), (4099, 70, 40), (30_003, 130, 300)])
def test_fixed_update_forms_on_ragged_shapes(kmc_mod, n, d, k):
    """The three forms of the fixed-point update (full pass, moved points
    through the LDS accumulators, moved points by global atomics) on shapes
    with n not a multiple of four (the label copy's tail), narrow and wide
    dimension slices and k beyond one accumulator block: a 10-iteration run
    equals ten single-iteration runs (each a full pass) bit for bit."""
    from kmc import datagen
    X, _, _ = datagen.blobs(n, d, k, 777 + n, -3.0, 3.0, 0.8, 1.6)
    C0 = datagen.init_rows(X, k, 5).astype(np.float32)
    iters = 10
    with kmc_mod.Engine(n, d, k, tol=0.0, precision="bf16", update="fixed") as e:
        e.set_points(X)
        e.set_centroids(C0)
        r = e.run(iters, check_converged=False)
        C_run, L_run = e.centroids(), e.labels()
        e.set_centroids(C0)
        for _ in range(iters):
            r1 = e.run(1, check_converged=False)
        _say("fixed forms", n, d, k, "full passes", r["full_passes"], "moved", r["moved_points"])
        assert r["full_passes"] == 1
        assert np.array_equal(e.labels(), L_run)
        assert np.array_equal(e.centroids(), C_run)
        assert r1["inertia"] == r["inertia"]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("update", ["ordered", "tree", "fixed"])
def test_graph_replays_match_eager(kmc_mod, prec, update):
    """Captured iteration chunks replayed several times, with eager runs and
    synchronisations in between, give the eager result every time."""
    g = load_golden("ref_c1_blobs.npz")
    X, C0 = g["X"], g["C0"]
    with kmc_mod.Engine(X.shape[0], 64, 10, tol=float(g["tol"]), precision=prec,
                        update=update) as e:
        e.set_points(X)
        e.set_centroids(C0)
        r_eager = e.run(300, check_converged=True)
        lab_e, C_e = e.labels(), e.centroids()
        for rep in range(3):
            _say("graph replay", prec, update, rep)
            e.set_centroids(C0)
            r_g = e.run(300, check_converged=True, graph=True)
            assert r_g["converged_at"] == r_eager["converged_at"]
            assert np.array_equal(e.labels(), lab_e)
            assert np.array_equal(e.centroids(), C_e)
            e.set_centroids(C0)
            r2 = e.run(7, check_converged=False)  # eager work between replays
            assert r2["iterations"] == 7
        # fixed-length graph (one chunk) against eager
        e.set_centroids(C0)
        e.run(12, check_converged=False)
        C12 = e.centroids()
        for rep in range(2):
            e.set_centroids(C0)
            e.run(12, check_converged=False, graph=True)
            assert np.array_equal(e.centroids(), C12)
    if update == "ordered":
        assert np.array_equal(lab_e, g["labels"].astype(np.int32))


def test_full_size_fixed_update_matches_fp64(kmc_mod, oracle, c2_data):
    """Fixed-point update at N=1M: counts exact, means within 1 ulp of the
    FP64 means, FP64 inertia within 1e-9."""
    X, C0 = c2_data
    rng = np.random.default_rng(4)
    lab = rng.integers(0, 256, size=X.shape[0]).astype(np.int32)
    lab[lab == 17] = 18  # one empty cluster -> zero vector
    with kmc_mod.Engine(X.shape[0], 128, 256, update="fixed") as e:
        e.set_points(X)
        Cn, cnt, res = e.update(lab, C0)
    C64, cnt64 = oracle.f64_centroids(X, lab.astype(np.uint32), 256)
    assert np.array_equal(cnt, cnt64)
    assert np.allclose(Cn, C64, rtol=2.4e-7, atol=1e-30), np.max(np.abs(Cn - C64))
    assert not Cn[17].any()
    ref_in = oracle.f64_inertia(X, C0, lab.astype(np.uint32))
    assert abs(res["inertia"] - ref_in) <= 1e-9 * ref_in


@pytest.mark.parametrize("n,d,k,seed", [(5000, 20, 16, 3), (70000, 64, 40, 11),
                                        (1797, 64, 10, 2026)])
def test_plusplus_matches_oracle(kmc_mod, oracle, n, d, k, seed):
    """k-means++ seeding on the GPU picks exactly the oracle's points (same
    uniforms, same FP64 summation order)."""
    from kmc import datagen
    from oracle import splitmix_uniforms
    X, _, _ = datagen.blobs(n, d, max(2, k // 2), seed, -8.0, 8.0, 0.5, 2.0)
    _say("k-means++", n, d, k, seed)
    with kmc_mod.Engine(n, d, k) as e:
        e.set_points(X)
        chosen = e.init_plusplus(seed)
        C = e.centroids()
    want, Cw = oracle.plusplus(X, k, splitmix_uniforms(seed, k))
    assert np.array_equal(chosen, want), (chosen, want)
    assert np.array_equal(C, X[want])
    assert len(set(chosen.tolist())) == k


@pytest.mark.parametrize("n,d,k,seed,trials", [(5000, 20, 16, 3, 0), (70000, 64, 40, 11, 4),
                                               (1797, 64, 10, 42, 0)])
def test_greedy_plusplus_matches_oracle(kmc_mod, oracle, n, d, k, seed, trials):
    """Greedy k-means++ (trials D^2 candidates per round, 0 = 2 + ceil(ln k))
    on the GPU picks exactly the oracle's points."""
    import math
    from kmc import datagen
    from oracle import splitmix_uniforms
    X, _, _ = datagen.blobs(n, d, max(2, k // 2), seed, -8.0, 8.0, 0.5, 2.0)
    s = trials if trials > 0 else 2 + math.ceil(math.log(k))
    _say("greedy k-means++", n, d, k, seed, s)
    with kmc_mod.Engine(n, d, k) as e:
        e.set_points(X)
        chosen = e.init_plusplus(seed, trials=trials)
        C = e.centroids()
    want, _ = oracle.plusplus(X, k, splitmix_uniforms(seed, 1 + (k - 1) * s), trials=s)
    assert np.array_equal(chosen, want), (chosen, want)
    assert np.array_equal(C, X[want])


@pytest.mark.parametrize("d,k", [(16, 512), (48, 400), (64, 100), (100, 200), (128, 256),
                                 (40, 600), (128, 300), (200, 64)])
@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_short_queue_pruned_groups_over_shapes(kmc_mod, oracle, prec, d, k):
    """The single-sweep MFMA-pruned all-k groups, run beside the
    two/three-candidate refinement in one launch, for each count of
    32-centroid blocks per wave (K up to 128, 256 and 512) and padded widths
    16 to 128, on a long queue (every point: the grid is split by estimated
    work and workgroups take several groups), against the two-sweep pruned
    pass and the exact pass over every centroid; shapes beyond them (K over
    512, over 256 at D <= 128, any K at D = 256) take the other two paths
    whatever the setting. The
    centroids come in groups of four near-duplicates around a quarter of the
    blob centres, so a point has four or more candidates inside the band
    (its own blob's group, or for the other blobs' points many far
    groups at similar distances)."""
    from kmc import datagen
    n = 60_000
    X, _, centres = datagen.blobs(n, d, k, 4000 + d + k, -10.0, 10.0, 0.5, 2.0)
    rng = np.random.default_rng(d * 1000 + k)
    C = np.repeat(centres[: k // 4], 4, axis=0)
    C = np.ascontiguousarray(C + rng.normal(0.0, 0.02, C.shape), dtype=np.float32)
    want, _ = _reference_labels(oracle, X, C)
    with kmc_mod.Engine(n, d, k, precision=prec) as e:
        e.set_points(X)
        for short, pm in ((True, 32768), (True, 0), (False, 1), (False, 0)):
            e.set_short_prune(short)
            e.set_prune_min(pm)
            lab, info = e.assign(C)
            _say("shape", d, k, prec, short, pm, "refined", info["refined_two"],
                 info["refined_full"])
            assert np.array_equal(lab, want), (d, k, prec, short, pm)
