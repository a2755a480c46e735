"""The reference GPU engine's loop rules on the GPU: the defaults of the
drop-in Kmeans class and of gpukmeans without --cpu-semantics
(src/kmeans.cu:365-444 labels i % k, :1053-1073 float score test,
src/kernels/distances.cu:588-598 empty cluster -> +inf).

Checked against the oracle's restatement (orc_gpu_lloyd, pinned by hand-worked
cases in test_oracle.py): labels and centroids bit-identical for every
precision form (the margin band hands every point whose order could differ to
the reference's float arithmetic), the same convergence iteration, and the
score within 1e-6 relative (FP64 sums in another order, then float).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

PRECS = ("fp32", "bf16x3", "bf16", "fp8")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "matrix-centric-k-means_amd", "bin", "gpukmeans")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run_rules(kmc_mod, X, k, tol, maxiter, prec, labels0=None, graph=False):
    n, d = X.shape
    lab0 = (np.arange(n) % k).astype(np.int32) if labels0 is None else labels0
    with kmc_mod.Engine(n, d, k, tol=tol, precision=prec) as e:
        e.set_points(X)
        e.set_labels(lab0)
        res = e.run(maxiter, check_converged=True, converge="score", empty="exclude",
                    graph=graph)
        return e.labels(), e.centroids(), res


def _check(got, want, what):
    lab, C, res = got
    conv, lab_o, C_o, sc_o, ran_o = want
    assert res["converged_at"] == conv, (what, res["converged_at"], conv)
    assert res["iterations"] == ran_o, (what, res["iterations"], ran_o)
    assert np.array_equal(lab, lab_o.astype(np.int32)), (what, int((lab != lab_o).sum()))
    assert np.array_equal(C, C_o), what
    s = np.float32(res["score"])
    assert abs(float(s) - float(sc_o[-1])) <= 1e-6 * abs(float(sc_o[-1])) + 1e-6, \
        (what, float(s), float(sc_o[-1]))


@pytest.mark.parametrize("prec", PRECS)
def test_known_answers(kmc_mod, prec):
    # the hand-worked cases of test_oracle.py
    X = np.array([0, 1, 10, 11, 20, 21], np.float32).reshape(-1, 1)
    lab, C, res = _run_rules(kmc_mod, X, 2, 1e-4, 10, prec)
    assert lab.tolist() == [0, 0, 0, 1, 1, 1]
    assert res["converged_at"] == 3
    assert np.float32(res["score"]) == np.float32(-2825.0 / 3.0)
    assert _run_rules(kmc_mod, X, 2, 300.0, 10, prec)[2]["converged_at"] == 2
    X = np.array([0, 1, 2, 6, 7], np.float32).reshape(-1, 1)
    lab, C, res = _run_rules(kmc_mod, X, 3, 1e-4, 20, prec)
    assert lab.tolist() == [2, 2, 2, 1, 1]
    assert res["converged_at"] == 3
    assert C[:, 0].tolist() == [0.0, 6.5, 1.0]
    assert np.float32(res["score"]) == np.float32(-87.5)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", ["ref_c1_blobs.npz", "ref_digits.npz"])
def test_golden_data_from_modulo_labels(kmc_mod, oracle, prec, name):
    g = load_golden(name)
    X, k = g["X"], int(g["k"])
    tol, maxiter = 1e-4, 300
    _say("reference-GPU rules", name, prec)
    want = oracle.gpu_lloyd(X, np.arange(X.shape[0]) % k, k, tol, maxiter)
    _check(_run_rules(kmc_mod, X, k, tol, maxiter, prec), want, (name, prec))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("seed", [1, 2])
def test_emptied_clusters_stay_excluded(kmc_mod, oracle, prec, seed):
    """Labels i % k over three blobs with k = 24: most clusters empty in the
    first assignment and must never be chosen again."""
    rng = np.random.default_rng(seed)
    n, d, k = 3000, 4, 24
    cen = rng.uniform(-10, 10, (3, d))
    X = (cen[rng.integers(0, 3, n)] + rng.normal(0, 1.0, (n, d))).astype(np.float32)
    want = oracle.gpu_lloyd(X, np.arange(n) % k, k, 1e-3, 100)
    assert (np.bincount(want[1].astype(np.int64), minlength=k) == 0).sum() >= 10
    _say("emptied clusters, seed", seed, prec)
    for graph in (False, True):
        _check(_run_rules(kmc_mod, X, k, 1e-3, 100, prec, graph=graph), want,
               (seed, prec, graph))


def test_cli_default_rules(tmp_path, oracle):
    """gpukmeans without --cpu-semantics (the reference GPU engine's rules)
    on the C1 points: labels, convergence line and Score line."""
    g = load_golden("ref_c1_blobs.npz")
    X = g["X"]
    n, d = X.shape
    k, tol, maxiter = 10, 1e-4, 300
    inp = tmp_path / "in.csv"
    out = tmp_path / "out.csv"
    with open(inp, "w") as f:
        f.write("k," + ",".join("d%d" % j for j in range(d)) + "\n")
        for i in range(n):
            f.write("0," + ",".join("%.9g" % v for v in X[i]) + "\n")
    cmd = [EXE, "-d", str(d), "-n", str(n), "-k", str(k), "-m", str(maxiter), "-t", repr(tol),
           "-c", "1", "-b", "random", "-f", "linear", "-l", "2", "-r", "2", "-i", str(inp),
           "-o", str(out), "--label-col", "first"]
    _say("cli default rules")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    lab = np.loadtxt(out, delimiter=",", skiprows=1, usecols=0, dtype=np.int64)
    Xr = np.loadtxt(inp, delimiter=",", skiprows=1, dtype=np.float64)[:, 1:].astype(np.float32)
    conv, want, C, sc, ran = oracle.gpu_lloyd(Xr, np.arange(n) % k, k, tol, maxiter)
    assert np.array_equal(lab, want.astype(np.int64))
    if conv < maxiter:
        assert ("K-means converged at iteration %d" % conv) in r.stdout, r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("Score:")]
    assert len(line) == 1, r.stdout
    score = float(line[0].split(":")[1])
    assert abs(score - float(sc[-1])) <= 1e-6 * abs(float(sc[-1])) + 1e-5, (score, sc[-1])
