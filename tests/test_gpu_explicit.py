"""The explicit-Lloyd rules of the reference's RAFT benchmark
(experiments/raft-bench/src/kmeans.cu:28-265) on the GPU engine, and the
benchmark program that the reference's harness runs
(experiments/driver.py:240-311, 467-515): stop on the sum of squared
centroid changes and on the cost ratio, empty clusters keep their centroid.
Checked against the oracle's restatement of those rules over the exact
labels."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "matrix-centric-k-means_amd", "bin", "raftbench")


def _run_engine(kmc_mod, X, C0, tol, maxiter, check, prec="fp32", update="tree"):
    with kmc_mod.Engine(X.shape[0], X.shape[1], C0.shape[0], tol=tol, precision=prec,
                        update=update) as e:
        e.set_points(X)
        e.set_centroids(C0)
        r = e.run(maxiter, check_converged=check, converge="shift_sum", cost_ratio=True,
                  empty="keep")
        return r, e.labels(), e.centroids()


UPDATES = ("tree", "fixed")


@pytest.mark.parametrize("update", UPDATES)
@pytest.mark.parametrize("check", [True, False])
def test_known_answers(kmc_mod, check, update):
    X = np.array([[0.0], [1.0], [10.0], [11.0]], np.float32)
    C0 = np.array([[0.0], [1.0]], np.float32)
    r, lab, C = _run_engine(kmc_mod, X, C0, 1e-4, 50, check, update=update)
    assert int(r["converged_at"]) == (3 if check else 4)
    assert np.array_equal(lab, [0, 0, 1, 1])
    assert np.array_equal(C, np.array([[0.5], [10.5]], np.float32))


@pytest.mark.parametrize("update", UPDATES)
def test_empty_cluster_keeps_centroid(kmc_mod, update):
    X = np.array([[0.0], [1.0], [2.0]], np.float32)
    C0 = np.array([[1.0], [100.0]], np.float32)
    r, lab, C = _run_engine(kmc_mod, X, C0, 1e-4, 50, True, update=update)
    assert int(r["converged_at"]) == 1
    assert np.array_equal(lab, [0, 0, 0])
    assert np.array_equal(C, C0)


@pytest.mark.parametrize("update", UPDATES)
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("check", [True, False])
def test_matches_oracle_on_digits_like_data(kmc_mod, oracle, prec, check, update):
    """Integer-valued data: FP64 and fixed-point cluster sums are both exact,
    so every update form must give the oracle's centroid bits (the fixed-point
    one from totals kept across the run once few points move)."""
    g = load_golden("ref_c1_blobs.npz")
    X, C0 = g["X"], g["C0"]
    tol = 1e-3
    want_t, want_lab, want_C, costs = oracle.lloyd_explicit(X, C0, tol, 300, check)
    r, lab, C = _run_engine(kmc_mod, X, C0, tol, 300, check, prec, update)
    assert int(r["converged_at"]) == want_t
    assert np.array_equal(lab, want_lab.astype(np.int32))
    assert np.array_equal(C, want_C)
    assert abs(r["inertia"] - costs[-1]) <= 1e-9 * costs[-1]


def test_benchmark_program_report(tmp_path):
    """The report lines parse with the harness's own patterns; a seeded run
    repeats exactly."""
    g = load_golden("ref_c1_blobs.npz")
    X = g["X"]
    n, d = X.shape
    inp = tmp_path / "in.csv"
    with open(inp, "w") as f:
        f.write("k," + ",".join("d%d" % j for j in range(d)) + "\n")
        for i in range(n):
            f.write("0," + ",".join("%.9g" % v for v in X[i]) + "\n")
    cmd = [EXE, str(n), str(d), "10", "300", "1", "0.0001", "random", str(inp), "7"]
    outs = []
    for _ in range(2):
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout
        outs.append(r.stdout)
    out = outs[0]
    assert re.findall(r"n:(\d+) d:(\d+) k:(\d+)", out)[0] == (str(n), str(d), "10")
    t = re.escape("kmeans-time") + r": (\d+\.\d+)s"
    assert re.search(t, out)
    assert re.search(re.escape("init-time") + r": (\d+\.\d+)s", out)
    score = float(re.search(r"kmeans-score: (\d+\.\d+)", out).group(1))
    iters = int(re.search(r"kmeans-iterations: (\d+)", out).group(1))
    assert score > 0 and 0 <= iters <= 300
    per_trial = re.findall(r"^(\d+(?:\.\d+)?(?:e[+-]?\d+)?)$", out, re.M)
    assert len(per_trial) == 10
    # same seed, same trials
    assert per_trial == re.findall(r"^(\d+(?:\.\d+)?(?:e[+-]?\d+)?)$", outs[1], re.M)
