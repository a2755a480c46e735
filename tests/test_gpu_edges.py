"""Edge cases of the Lloyd path on the GPU, each against the oracle's
restatement of the reference CPU k-means: one cluster, fewer points than a
tile, one dimension, all points identical (every distance ties: the first
centroid wins), exact duplicate centroids, and a width that needs padding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PRECS = ("fp32", "bf16x3", "bf16", "fp8")


def _check(kmc_mod, oracle, X, C0, prec, tol=1e-4, maxiter=50, update="ordered"):
    want_conv, want_lab, want_C, _, _ = oracle.lloyd(X, C0, tol, maxiter)
    with kmc_mod.Engine(X.shape[0], X.shape[1], C0.shape[0], tol=tol, precision=prec,
                        update=update) as e:
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


@pytest.mark.parametrize("prec", PRECS)
def test_one_dimension(kmc_mod, oracle, prec):
    rng = np.random.default_rng(2)
    X = np.concatenate([rng.normal(c, 0.5, size=(300, 1)) for c in (-5, 0, 5)]).astype(np.float32)
    _check(kmc_mod, oracle, X, X[[0, 300, 600]].copy(), prec)


@pytest.mark.parametrize("prec", PRECS)
def test_identical_points_tie_to_first_centroid(kmc_mod, oracle, prec):
    X = np.full((2000, 16), 3.25, np.float32)
    C0 = np.stack([np.full(16, 3.0, np.float32), np.full(16, 3.5, np.float32)])
    lab = _check(kmc_mod, oracle, X, C0, prec)
    assert not lab.any()  # equidistant: the first index wins (strict <)


@pytest.mark.parametrize("prec", PRECS)
def test_duplicate_centroids(kmc_mod, oracle, prec):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, size=(3000, 24)).astype(np.float32)
    C0 = np.concatenate([X[:5], X[:5]])  # clusters 5..9 duplicate 0..4
    lab = _check(kmc_mod, oracle, X, C0, prec, maxiter=1)
    assert lab.max() < 5  # a duplicate never beats its first copy


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("d", [3, 33, 130])
def test_padded_widths(kmc_mod, oracle, prec, d):
    rng = np.random.default_rng(d)
    X = rng.normal(0, 2, size=(4000, d)).astype(np.float32)
    C0 = X[rng.choice(4000, 12, replace=False)]
    _check(kmc_mod, oracle, X, C0, prec, maxiter=20)
    _check(kmc_mod, oracle, X, C0, prec, maxiter=20, update="tree")
