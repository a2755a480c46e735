"""GPU parity of the batch engine (kmc_batch_*, the reference's `-r` loop and
many independent clusterings run as one batch).

Bar: every problem of a batch of uneven n, d and k ends exactly where a run of
that problem alone ends (Engine with update="fixed": identical labels,
identical centroids, identical iteration count and convergence iteration) —
the batch sums each cluster exactly, as 64-bit integers at the fixed-point
scale a lone run of the problem chooses (full passes over the sorted members,
then the moved points only), and integer sums do not depend on the order. A
batch member that is the C1 golden problem reproduces the reference CPU
k-means labels (its centroids are exact sums rounded once, so they match the
reference's float sums to 1e-6).
"""
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problems(seed, count=12):
    from kmc import datagen
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        n = int(rng.integers(300, 20000))
        d = int(rng.choice([3, 16, 20, 33, 64, 100, 128]))
        k = int(min(n, rng.integers(2, 129)))
        X, _, _ = datagen.blobs(n, d, max(2, k // 2 + 1), seed + i, -5.0, 5.0, 0.5, 1.5)
        probs.append((X, datagen.init_rows(X, k, seed + 100 + i)))
    return probs


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16"])
def test_batch_matches_lone_runs(kmc_mod, prec):
    probs = _problems(5)
    g = load_golden("ref_c1_blobs.npz")
    probs.insert(3, (g["X"], g["C0"]))
    tol, maxiter = float(g["tol"]), 60
    _say("batch", prec, [(X.shape[0], X.shape[1], C.shape[0]) for X, C in probs])
    with kmc_mod.Batch(probs, tol=tol, precision=prec) as b:
        res, ms = b.run(maxiter)
        _say("batch run ms", ms, [r["iterations"] for r in res])
        res2, _ = b.run(maxiter)  # every run starts from the initial centroids
        assert res2 == res
        for p, (X, C0) in enumerate(probs):
            with kmc_mod.Engine(X.shape[0], X.shape[1], C0.shape[0], tol=tol, precision=prec,
                                update="fixed") as e:
                e.set_points(X)
                e.set_centroids(C0)
                r = e.run(maxiter, check_converged=True)
                lab, C = e.labels(), e.centroids()
            assert res[p]["iterations"] == r["iterations"], (p, res[p], r)
            assert res[p]["converged_at"] == r["converged_at"], (p, res[p], r)
            assert np.array_equal(b.labels(p), lab), p
            assert np.array_equal(b.centroids(p), C), p
        lab_c1 = b.labels(3)
        assert np.array_equal(lab_c1, g["labels"].astype(np.int32))
        assert np.allclose(b.centroids(3), g["C"], rtol=1e-6, atol=1e-6)
        assert abs(int(res[3]["converged_at"]) - int(g["converged"])) <= 1


def test_batch_inertia_matches_fp64(kmc_mod, oracle):
    probs = _problems(9, count=6)
    with kmc_mod.Batch(probs, tol=0.0, precision="fp32") as b:
        res, _ = b.run(5, check_converged=False)
        for p, (X, C0) in enumerate(probs):
            assert res[p]["iterations"] == 5
            want = oracle.f64_inertia(X, b.centroids(p), b.labels(p).astype(np.uint32))
            assert abs(res[p]["inertia"] - want) <= 1e-9 * max(want, 1e-30), (p, res[p], want)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batch_graph_replays_equal_eager(kmc_mod, prec):
    """KMC_RUN_GRAPH: each chunk of batch iterations (every width group as a
    branch of one captured graph) ends exactly where the eager launches end;
    the graphs are replayed several times with eager runs in between, with
    and without the convergence test, and with a chunk shorter than 16."""
    probs = _problems(21, count=9)
    with kmc_mod.Batch(probs, tol=1e-4, precision=prec) as b:
        for maxiter, check in ((40, True), (19, False)):
            eager, _ = b.run(maxiter, check_converged=check)
            lab_e = [b.labels(p) for p in range(len(probs))]
            C_e = [b.centroids(p) for p in range(len(probs))]
            for rep in range(3):
                res, ms = b.run(maxiter, check_converged=check, graph=True)
                _say("graph", prec, maxiter, check, rep, ms, [r["iterations"] for r in res])
                assert res == eager
                for p in range(len(probs)):
                    assert np.array_equal(b.labels(p), lab_e[p]), (rep, p)
                    assert np.array_equal(b.centroids(p), C_e[p]), (rep, p)
                again, _ = b.run(maxiter, check_converged=check)  # eager between replays
                assert again == eager
