"""Edge cases of the Lloyd path on the GPU, each against the oracle's
restatement of the reference CPU k-means: one cluster, fewer points than a
tile, one dimension, all points identical (every distance ties: the first
centroid wins), exact duplicate centroids, a width that needs padding, and
many tied candidates -- the tie-prone cases with each all-k refinement
path (prune 0: every centroid exactly; 1: the bf16x3 MFMA-pruned pass;
"short": its single-sweep form for short queues)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = ("fp32", "bf16x3", "bf16", "fp8")


def _check(kmc_mod, oracle, X, C0, prec, tol=1e-4, maxiter=50, update="ordered", prune=None):
    want_conv, want_lab, want_C, _, _ = oracle.lloyd(X, C0, tol, maxiter)
    with kmc_mod.Engine(X.shape[0], X.shape[1], C0.shape[0], tol=tol, precision=prec,
                        update=update) as e:
        if prune in (0, 1):
            e.set_short_prune(False)
        if prune is not None:
            e.set_prune_min(0 if prune == "short" else prune)
        e.set_points(X)
        e.set_centroids(C0)
        r = e.run(maxiter, check_converged=True)
        lab, C = e.labels(), e.centroids()
    assert np.array_equal(lab, want_lab.astype(np.int32)), (prec, lab, want_lab)
    assert int(r["converged_at"]) == want_conv
    if update == "ordered":
        assert np.array_equal(C, want_C)
    return lab


@pytest.mark.parametrize("prec", PRECS)
def test_single_cluster(kmc_mod, oracle, prec):
    rng = np.random.default_rng(1)
    X = rng.normal(0, 3, size=(777, 9)).astype(np.float32)
    lab = _check(kmc_mod, oracle, X, X[:1].copy(), prec)
    assert not lab.any()


@pytest.mark.parametrize("prec", PRECS)
def test_fewer_points_than_a_tile(kmc_mod, oracle, prec):
    X = np.array([[0, 0], [1, 0], [9, 9], [10, 10], [0, 1]], np.float32)
    _check(kmc_mod, oracle, X, X[[0, 2]].copy(), prec)


@pytest.mark.parametrize("prune", [0, 1, "short"])
@pytest.mark.parametrize("prec", PRECS)
def test_one_dimension(kmc_mod, oracle, prec, prune):
    rng = np.random.default_rng(2)
    X = np.concatenate([rng.normal(c, 0.5, size=(300, 1)) for c in (-5, 0, 5)]).astype(np.float32)
    _check(kmc_mod, oracle, X, X[[0, 300, 600]].copy(), prec, prune=prune)


@pytest.mark.parametrize("prune", [0, 1, "short"])
@pytest.mark.parametrize("prec", PRECS)
def test_identical_points_tie_to_first_centroid(kmc_mod, oracle, prec, prune):
    X = np.full((2000, 16), 3.25, np.float32)
    C0 = np.stack([np.full(16, 3.0, np.float32), np.full(16, 3.5, np.float32)])
    lab = _check(kmc_mod, oracle, X, C0, prec, prune=prune)
    assert not lab.any()  # equidistant: the first index wins (strict <)


@pytest.mark.parametrize("prune", [0, 1, "short"])
@pytest.mark.parametrize("prec", PRECS)
def test_duplicate_centroids(kmc_mod, oracle, prec, prune):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, size=(3000, 24)).astype(np.float32)
    C0 = np.concatenate([X[:5], X[:5]])  # clusters 5..9 duplicate 0..4
    lab = _check(kmc_mod, oracle, X, C0, prec, maxiter=1, prune=prune)
    assert lab.max() < 5  # a duplicate never beats its first copy


@pytest.mark.parametrize("prune", [0, 1, "short"])
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d", [3, 33, 130])
def test_padded_widths(kmc_mod, oracle, prec, d, prune):
    rng = np.random.default_rng(d)
    X = rng.normal(0, 2, size=(4000, d)).astype(np.float32)
    C0 = X[rng.choice(4000, 12, replace=False)]
    _check(kmc_mod, oracle, X, C0, prec, maxiter=20, prune=prune)
    _check(kmc_mod, oracle, X, C0, prec, maxiter=20, update="tree", prune=prune)
    _check(kmc_mod, oracle, X, C0, prec, maxiter=20, update="fixed", prune=prune)


@pytest.mark.parametrize("prune", [0, 1, "short"])
@pytest.mark.parametrize("prec", PRECS)
def test_many_tied_candidates(kmc_mod, oracle, prec, prune):
    """Twelve copies of one centroid among twenty others: points near the
    copies have twelve exactly tied candidates, which the all-k pass narrows
    and evaluates one by one (the first copy wins)."""
    rng = np.random.default_rng(11)
    X = rng.normal(0, 1, size=(4000, 48)).astype(np.float32)
    others = rng.normal(0, 3, size=(20, 48)).astype(np.float32)
    C0 = np.concatenate([others[:7], np.repeat(X[:1], 12, axis=0), others[7:]])
    lab = _check(kmc_mod, oracle, X, C0, prec, maxiter=1, prune=prune)
    assert not np.any((lab > 7) & (lab < 19))  # never a later copy


@pytest.mark.parametrize("prune", [0, 1, "short"])
@pytest.mark.parametrize("prec", PRECS)
def test_all_centroids_tied(kmc_mod, oracle, prec, prune):
    """Forty identical centroids: every point has forty tied candidates, more
    than the pruned all-k pass keeps, so every centroid is evaluated exactly
    and the first one wins."""
    rng = np.random.default_rng(12)
    X = rng.normal(0, 1, size=(3000, 40)).astype(np.float32)
    C0 = np.repeat(X[5:6], 40, axis=0)
    lab = _check(kmc_mod, oracle, X, C0, prec, maxiter=1, prune=prune)
    assert not lab.any()
