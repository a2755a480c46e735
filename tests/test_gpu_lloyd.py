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

PRECS = ("fp32", "bf16x3", "bf16")


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
