"""The gpukmeans command line on the GPU (the drop-in for src/main.cu):
with --cpu-semantics its labels and convergence iteration are the reference
CPU k-means' from the same seeded initial rows (random) or from the engine's
k-means++ seeding (plusplus), checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "matrix-centric-k-means_amd", "bin", "gpukmeans")


def _run_cli(tmp_path, X, k, init, seed, tol, maxiter, extra=(), runs=2):
    n, d = X.shape
    inp = tmp_path / "in.csv"
    out = tmp_path / "out.csv"
    with open(inp, "w") as f:
        f.write("k," + ",".join("d%d" % j for j in range(d)) + "\n")
        for i in range(n):
            f.write("0," + ",".join("%.9g" % v for v in X[i]) + "\n")
    cmd = [EXE, "-d", str(d), "-n", str(n), "-k", str(k), "-m", str(maxiter), "-t", repr(tol),
           "-c", "1", "-b", init, "-f", "linear", "-l", "2", "-s", str(seed), "-r", str(runs),
           "-i", str(inp), "-o", str(out), "--label-col", "first", "--cpu-semantics"]
    cmd += list(extra)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    lab = np.loadtxt(out, delimiter=",", skiprows=1, usecols=0, dtype=np.int64)
    # the points as the command read them: decimal text -> double -> float
    Xr = np.loadtxt(inp, delimiter=",", skiprows=1, dtype=np.float64)[:, 1:].astype(np.float32)
    return lab, Xr, r.stdout


@pytest.mark.parametrize("init", ["random", "plusplus"])
def test_cli_cpu_semantics_matches_reference(tmp_path, oracle, init):
    from oracle import splitmix_uniforms
    g = load_golden("ref_c1_blobs.npz")
    X = g["X"]
    k, seed, tol, maxiter = 10, 42, 1e-4, 300
    lab, X, text = _run_cli(tmp_path, X, k, init, seed, tol, maxiter)
    if init == "random":
        C0 = oracle.init_centroids(X, k, seed)
    else:
        # greedy k-means++ with 2 + ceil(ln k) trials per round
        import math
        st = 2 + math.ceil(math.log(k))
        chosen, C0 = oracle.plusplus(X, k, splitmix_uniforms(seed, 1 + (k - 1) * st), trials=st)
    conv, want, C, cnt, ran = oracle.lloyd(X, C0, tol, maxiter)
    assert np.array_equal(lab, want.astype(np.int64))
    assert ("K-means converged at iteration %d" % conv) in text, text
    assert "GPU_Kmeans:" in text and "Score:" in text


def test_cli_polynomial_kernel_runs_kernel_kmeans(tmp_path, kmc_mod):
    """-f polynomial runs kernel k-means from labels i % k (the reference's
    Popcorn path); its labels are those of the kernel k-means engine."""
    rng = np.random.default_rng(3)
    n, d, k = 600, 12, 4
    centres = rng.uniform(0, 1, size=(k, d))
    X = (centres[rng.integers(0, k, size=n)] + 0.05 * rng.standard_normal((n, d)))
    X = X.astype(np.float32)
    inp = tmp_path / "in.csv"
    out = tmp_path / "out.csv"
    with open(inp, "w") as f:
        f.write("k," + ",".join("d%d" % j for j in range(d)) + "\n")
        for i in range(n):
            f.write("0," + ",".join("%.9g" % v for v in X[i]) + "\n")
    Xr = np.loadtxt(inp, delimiter=",", skiprows=1, dtype=np.float64)[:, 1:].astype(np.float32)
    cmd = [EXE, "-d", str(d), "-n", str(n), "-k", str(k), "-m", "30", "-t", "0.0001", "-c", "1",
           "-b", "random", "-f", "polynomial", "-l", "2", "-r", "2", "-i", str(inp), "-o",
           str(out), "--label-col", "first"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    lab = np.loadtxt(out, delimiter=",", skiprows=1, usecols=0, dtype=np.int64)
    with kmc_mod.KernelEngine(n, d, k, kernel="polynomial", store="fp32", tol=1e-4) as e:
        e.set_points(Xr)
        e.set_labels((np.arange(n) % k).astype(np.int32))
        res = e.run(30, check_converged=True)
        want = e.labels()
    assert np.array_equal(lab, want.astype(np.int64))
    if res["converged_at"] < 30:
        assert ("K-means converged at iteration %d" % res["converged_at"]) in r.stdout


def _driver_fields(out, n_trials, maxiters):
    """The fields experiments/driver.py:55-141 extracts from gpukmeans'
    output, with its patterns and its averaging (first match skipped,
    statistics.mean / stdev of the rest)."""
    import re
    import statistics as stats

    def avg(pattern):
        lst = [float(t) for t in re.findall(pattern, out)[1:]]
        return stats.mean(lst), stats.stdev(lst)

    f = {
        "runtime": avg(r"Time: (\d+\.\d+)"),
        "mem": float(re.search(r"MEMORY FOOTPRINT: (\d+) MB", out).group(1)),
        "memcpy": avg(r"memcpy time: (\d+\.\d+)"),
        "argmin": avg(r"clusters_argmin_shfl time: (\d+\.\d+)"),
        "centroids": avg(r"compute_centroids time: (\d+\.\d+)"),
        "dist": avg(r"compute_distances time: (\d+\.\d+)"),
        "bmult": avg(r"b-mult time: (\d+\.\d+)"),
        "init": avg(r"init_centroids time: (\d+\.\d+)"),
        "score": float(re.search(r"Score: -?(\d+\.\d+)", out).group(1)),
    }
    iters = n = 0
    for line in out.splitlines():
        if "converged at iteration" in line:
            iters += int(line.split("iteration")[1].split("-")[0])
            n += 1
        elif "NOT" in line:
            iters += maxiters
            n += 1
    f["iterations"] = iters // n if n else 0
    return f


@pytest.mark.parametrize("kernel", ["linear", "polynomial"])
def test_cli_output_parses_with_harness_patterns(tmp_path, kernel):
    """gpukmeans -r 3 prints everything the reference harness parses: the
    six [PERFORMANCE] stage lines (three per run, three per iteration), the
    Time / MEMORY FOOTPRINT / Score lines and the convergence lines."""
    rng = np.random.default_rng(5)
    n, d, k, maxiter, runs = 4000, 16, 8, 12, 3
    cen = rng.uniform(-5, 5, (k, d))
    X = (cen[rng.integers(0, k, n)] + rng.normal(0, 1, (n, d))).astype(np.float32)
    inp = tmp_path / "in.csv"
    with open(inp, "w") as f:
        f.write("k," + ",".join("d%d" % j for j in range(d)) + "\n")
        for i in range(n):
            f.write("0," + ",".join("%.9g" % v for v in X[i]) + "\n")
    cmd = [EXE, "-d", str(d), "-n", str(n), "-k", str(k), "-m", str(maxiter), "-t", "0.0001",
           "-c", "0", "-b", "random", "-f", kernel, "-l", "2", "-r", str(runs), "-i", str(inp)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    out = r.stdout
    f = _driver_fields(out, runs, maxiter)
    assert f["iterations"] == maxiter
    assert out.count("memcpy time:") == runs
    assert out.count("b-mult time:") == runs
    assert out.count("init_centroids time:") == runs
    for stage in ("compute_distances", "clusters_argmin_shfl", "compute_centroids"):
        assert out.count(stage + " time:") == runs * maxiter, stage
    assert f["dist"][0] > 0 and f["centroids"][0] > 0 and f["mem"] > 0
    # stage lines carry seconds: a stage of this size takes well under 1 s
    assert f["dist"][0] < 1.0 and f["argmin"][0] < 1.0 and f["centroids"][0] < 1.0
    with_flag = subprocess.run(cmd + ["--no-stage-times"], stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=120)
    assert with_flag.returncode == 0 and "[PERFORMANCE]" not in with_flag.stdout


def test_cli_reads_libsvm(tmp_path, oracle):
    """A libsvm input (label, then 1-based idx:value pairs with zeros left
    out; src/include/input_parser.hpp:76-100) gives the same labels as the
    same points read from CSV, and the oracle's labels (CPU semantics)."""
    g = load_golden("ref_c1_blobs.npz")
    X = g["X"]
    n, d = X.shape
    k, seed, tol, maxiter = 10, 42, 1e-4, 300
    svm = tmp_path / "in.svm"
    out = tmp_path / "out.csv"
    with open(svm, "w") as f:
        for i in range(n):
            toks = ["%d:%.9g" % (j + 1, X[i, j]) for j in range(d) if X[i, j] != 0]
            f.write("1 " + " ".join(toks) + "\n")
    cmd = [EXE, "-d", str(d), "-n", str(n), "-k", str(k), "-m", str(maxiter), "-t", repr(tol),
           "-c", "1", "-b", "random", "-f", "linear", "-l", "2", "-s", str(seed), "-r", "2",
           "-i", str(svm), "-o", str(out), "--cpu-semantics", "--no-stage-times"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    lab = np.loadtxt(out, delimiter=",", skiprows=1, usecols=0, dtype=np.int64)
    Xw = np.loadtxt(out, delimiter=",", skiprows=1, dtype=np.float64)[:, 1:].astype(np.float32)
    assert np.array_equal(Xw, X)  # integer-valued points survive the text round trip
    C0 = oracle.init_centroids(X, k, seed)
    conv, want, C, cnt, ran = oracle.lloyd(X, C0, tol, maxiter)
    assert np.array_equal(lab, want.astype(np.int64))
    assert ("K-means converged at iteration %d" % conv) in r.stdout, r.stdout


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_cli_batch_runs_match_reference(tmp_path, oracle, prec):
    """--batch-runs -r 5: the five repetitions run as one batch (one launch
    per stage, graph replays) and each reports the reference CPU k-means'
    convergence iteration; the written labels are the reference's; every
    repetition's objective score equals the FP64 score of the oracle's final
    state to float rounding (the batch forms FP64 centroid sums)."""
    g = load_golden("ref_c1_blobs.npz")
    k, seed, tol, maxiter, runs = 10, 42, 1e-4, 300, 5
    lab, X, text = _run_cli(tmp_path, g["X"], k, "random", seed, tol, maxiter,
                            extra=["--batch-runs", "-p", prec], runs=runs)
    C0 = oracle.init_centroids(X, k, seed)
    conv, want, C, cnt, ran = oracle.lloyd(X, C0, tol, maxiter)
    assert np.array_equal(lab, want.astype(np.int64))
    assert text.count("K-means converged at iteration %d - " % conv) == runs, text
    assert ("GPU_Kmeans:" in text) and ("(%d runs)" % runs in text), text
    inertia = oracle.f64_inertia(X, C, want.astype(np.uint32))
    score = inertia - float(np.sum(X.astype(np.float64) ** 2))
    got = [float(line.split(":")[1]) for line in text.splitlines()
           if line.startswith("Objective score:")]
    assert len(got) == runs
    for v in got:
        assert abs(v - score) <= 1e-5 * abs(score) + 1e-3, (v, score)


def test_cli_batch_runs_refuses_gpu_semantics(tmp_path):
    """--batch-runs runs the batch engine, which follows the reference CPU
    k-means' rules: without --cpu-semantics it is refused."""
    cmd = [EXE, "-d", "4", "-n", "100", "-k", "3", "-m", "10", "-b", "random", "-f", "linear",
           "-l", "2", "-r", "2", "--batch-runs"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "--batch-runs needs" in r.stdout
