"""Lloyd-iteration throughput of the gfx950 k-means engine on one MI355X.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--prec bf16] [--update fixed]

Workload (the large plain-vs-BF16 configuration): N=1,000,000 points, D=128,
K=256 Gaussian blobs (centres U[-10,10]^128, per-blob sigma U[0.5,2.0]),
float32, seeded; 256 distinct seed-chosen rows as initial centroids; no
convergence test (tol=0, every iteration runs). A step is one Lloyd
iteration: fused MFMA distance/argmin + margin refinement + centroid update
+ shift, next-iteration operands and FP64 inertia.
Points are resident in HBM before the timed region; every timed pass runs
iterations 1..K from the initial centroids.

The headline form (``--prec``, default bf16: single-product bf16 MFMA with
exact recomputation of the margin-band points), the split form (bf16x3: hi+lo
bf16 products), the fp8 form (e4m3 with per-row power-of-two scales applied
by the matrix core) and the plain FP32 form run from the same initial
centroids;
the line reports the headline's iterations/s as ``value`` and the other forms
beside it, checks that every form ends with bit-identical centroids and
labels, and measures the inertia error of the headline form against an FP64
oracle. ``--update fixed`` (default) sums each cluster exactly in 64-bit
fixed point and keeps the totals across the run: the first iteration streams
every point, the others subtract and add the rows of the moved points only
(the totals equal those of a full pass, bit for bit); ``--update tree`` sorts the
members and sums each cluster in FP64 every iteration; ``--update ordered``
reproduces the reference's float summation order.

The roofline object covers the dominant kernel of the iteration (the one
with the larger mean launch time of the two large ones), the other is given
beside it: the fused distance kernel against the MFMA peak (achieved =
2*N*K*D algorithmic flops per launch / mean launch time) and the centroid
update against HBM (achieved = the bytes the update must move per iteration
-- N*D*4 + N*4 + K*D*4 for a full pass; labels, previous labels and the moved
points' rows for an incremental one, averaged over the run -- / mean
launch time). Both times are taken live with HIP events recorded around
every launch on the engine's stream, in a second timed pass of the same
steps; ``value`` and ``ms_per_step`` come from the first pass, which records
no events (``ms_per_step_with_kernel_events`` gives the second pass).

The CPU baseline is the reference CPU k-means (comparisons/CPU_Kmeans,
compiled from its own sources into oracle/_ref) timed single-threaded on this
host on a bounded sample (first 100,000 points, 3 iterations at N=1M), scaled to the
full N; falls back to the oracle restatement when oracle/_ref is absent.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "matrix-centric-k-means_amd"))

# workloads: c2 (default, the metric's configuration), c5 (large K), c3
# (kernel k-means, its own line)
WORKLOADS = {
    "c2": dict(n=1_000_000, d=128, k=256, seed=20261203, gen="large_blobs", steps=50,
               cpu_sample=100_000, cpu_iters=3, update="fixed",
               metric="Lloyd iter/s, N=1M D=128 K=256: BF16 MFMA vs FP32 at <=1e-5 inertia err",
               text="Gaussian blobs N=1000000 D=128 K=256"),
    # K = 4096: the fixed-point update's per-workgroup accumulators do not fit
    # the LDS, so the tree update runs
    "c5": dict(n=4_000_000, d=256, k=4096, seed=20261204, gen="huge_blobs", steps=20,
               cpu_sample=10_000, cpu_iters=1, update="tree",
               metric="Lloyd iter/s, N=4M D=256 K=4096: BF16 MFMA vs FP32",
               text="Gaussian blobs N=4000000 D=256 K=4096"),
}
N, D, K = 1_000_000, 128, 256
DATA_SEED, INIT_SEED = 20261203, 7
PEAK = {"fp32": 157.3e12, "bf16x3": 2.5e15, "bf16": 2.5e15, "fp8": 5.0e15}
DTYPE = {"bf16x3": "bf16", "bf16": "bf16", "fp32": "fp32", "fp8": "fp8"}
HBM_PEAK = 8.0e12


_T0 = time.perf_counter()


def hip_runtime():
    """The HIP runtime libkmc.so is linked against (loaded through it), for
    device-wide synchronisation and the plain device copy of --profile-copy:
    one runtime in the process, the library's own."""
    import ctypes
    import kmc
    kmc.load_library()
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipDeviceSynchronize.argtypes = []
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return L


def device_sync():
    """Wait for all work on the device (hipDeviceSynchronize)."""
    rt = hip_runtime()

    def sync():
        if rt.hipDeviceSynchronize() != 0:
            sys.exit("bench.py: hipDeviceSynchronize failed")
    return sync


def log(*a):
    print("[%7.1fs]" % (time.perf_counter() - _T0), *a, file=sys.stderr, flush=True)


def run_form(kmc, X, C0, prec, steps, warmup, sync, update, graph):
    with kmc.Engine(N, D, K, tol=0.0, precision=prec, update=update) as e:
        e.set_points(X)
        e.set_centroids(C0)
        if warmup:
            e.run(warmup, check_converged=False, graph=graph)
        # the variants are timed on the centroids after the warm-up
        # iterations: the first assignments from the seeded rows leave many
        # more band points to refine than the rest of a run does
        tuning = e.autotune()
        log("%s autotune: variant %d, ms %s" % (prec, tuning["chosen"],
                                               [round(v, 4) for v in tuning["ms"]]))
        # every timed pass runs the first `steps` iterations from the initial
        # centroids, as a run of the engine would. Capture the timed chunks
        # once (the capture is set-up, not a step): run each, then restore
        # the initial centroids and replay it timed. The step rate comes from
        # a pass without kernel events (an event record node in a graph
        # costs ~6 us of idle GPU); the per-kernel durations come from a
        # second pass of the same steps with events around every distance
        # and update launch.
        C_start = C0
        if graph:
            for tk in (False, True):
                e.set_centroids(C_start)
                e.run(steps, check_converged=False, time_kernels=tk, graph=True)
        e.set_centroids(C_start)
        sync()
        t0 = time.perf_counter()
        e.run(steps, check_converged=False, graph=graph)
        sync()
        wall = time.perf_counter() - t0
        e.set_centroids(C_start)
        sync()
        t1 = time.perf_counter()
        res = e.run(steps, check_converged=False, time_kernels=True, graph=graph)
        sync()
        wall_events = time.perf_counter() - t1
        C_end = e.centroids()
        lab_end = e.labels()
        # inertia of the final centroids through the engine's own assign+update
        lab_f, _ = e.assign(C_end)
        _, _, upd = e.update(lab_f, C_end)
    return {"wall_s": wall, "wall_events_s": wall_events, "res": res, "C_end": C_end,
            "labels_end": lab_end, "labels_final": lab_f, "inertia_gpu": upd["inertia"], "tuning": tuning}


def cpu_baseline(X, C0, n_sample=100_000, iters=2):
    n_sample = min(n_sample, X.shape[0])
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Reference
    Xs = np.ascontiguousarray(X[:n_sample])
    if Reference.available():
        secs = Reference().time_run(Xs, C0, iters)
        kind = "reference"
    else:
        O = Oracle()
        t0 = time.perf_counter()
        O.lloyd(Xs, C0, 0.0, iters)
        secs = time.perf_counter() - t0
        kind = "port"
    per_iter_full = secs / iters * (N / n_sample)
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"value": 1.0 / per_iter_full, "unit": "iter/s", "cores": 1, "kind": kind,
            "host_cpus": {"nproc": usable, "os_cpu_count": os.cpu_count()},
            "sample": "first %d of the %d points, D=%d, K=%d, %d iterations in %.2f s "
                      "(single thread, g++ -O2), scaled linearly in N; %.3f s per iteration "
                      "on the sample" % (n_sample, N, D, K, iters, secs, secs / iters)}


# Launch shapes of the distance kernel's variants in the autotuner's order:
# (waves per workgroup, point sets per wave, register prefetch, waves-per-SIMD
# target, two-set kernel, values kept per lane), the kVariants table of
# csrc/kmc_lloyd.hip (a CPU test compares them).
ASSIGN_VARIANTS = [(8, 1, True, 0, 0, 4, 1), (8, 1, False, 0, 0, 4, 1),
                   (12, 1, False, 0, 0, 4, 1), (16, 1, False, 0, 0, 4, 1),
                   (12, 1, False, 6, 0, 4, 1), (4, 2, False, 0, 0, 4, 1),
                   (8, 2, False, 0, 0, 4, 1), (8, 1, False, 8, 0, 4, 1),
                   (8, 2, False, 0, 1, 4, 1), (8, 2, False, 0, 1, 3, 2),
                   (4, 2, False, 0, 1, 4, 1), (4, 2, False, 0, 2, 4, 1),
                   (8, 2, False, 0, 1, 3, 4), (4, 2, False, 0, 2, 3, 2)]
FORM_CODE = {"fp32": 0, "bf16x3": 1, "bf16": 2, "fp8": 3}


def assign_kernel_name(dp, prec, variant):
    """The distance kernel's instantiation for a form and variant, as the
    profiler names it (without the return type and parameters)."""
    nw, ms, pf, wpe, two, topk, lists = ASSIGN_VARIANTS[variant]
    if two:
        return "k_assign2<%d, %d, %d, %d, %s, %s, %d, %d>" % (
            dp, FORM_CODE[prec], nw, topk, "true" if pf else "false",
            "true" if two == 1 else "false", wpe, lists)
    return "k_assign<%d, %d, %d, %d, %s, %d, false>" % (dp, FORM_CODE[prec], nw, ms,
                                                        "true" if pf else "false", wpe)


def traffic_from_profiles(role, form, workload="c2", kernel=None):
    """HBM bytes per launch of a kernel ("assign" per form, "update_tree",
    "update_ordered") from the committed rocprofv3 --pmc summary of the
    workload (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factors); for
    the distance kernel, the measured instantiation `kernel` when the summary
    holds it (the autotuner's choice can differ between runs)."""
    if workload != "c2":
        return None
    p = os.path.join(ROOT, "profiles", "lloyd_n1m_d128_k256_hbm_traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if role == "assign" and kernel and kernel in t.get("assign_variants", {}):
        return t["assign_variants"][kernel]
    e = t.get(role, {})
    if role == "assign":
        e = e.get(form, {})
    return e.get("hbm_bytes_per_launch")


def profile_copy(nbytes):
    """A plain device copy of known size (the runtime's blit kernel), for
    calibrating the PMC byte counters."""
    import ctypes
    rt = hip_runtime()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if rt.hipMalloc(ctypes.byref(a), nbytes) or rt.hipMalloc(ctypes.byref(b), nbytes):
        sys.exit("bench.py: hipMalloc failed")
    rt.hipMemset(a, 1, nbytes)
    for _ in range(3):
        rt.hipMemcpy(b, a, nbytes, 3)  # device to device
    rt.hipDeviceSynchronize()
    rt.hipFree(a)
    rt.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}))


def bench_kernel_kmeans(a):
    """Kernel k-means (the reference's Popcorn path) at N=65536, D=784, K=10:
    U[0,1] data with 70% zeros and 10-blob structure; polynomial (x.y+1)^2
    and RBF gamma=1/784 kernels, K stored fp32 (17.2 GB) and bf16 (8.6 GB).
    A step is one iteration: the stable sort of the points by label, the
    cluster-sorted S = K.H product streaming all of K (rows summed per
    cluster segment), the self terms, the distances and the argmin; the dense
    one-hot MFMA product is timed beside it. The kernel matrix is built once,
    outside the timed region (the reference builds it in the constructor)."""
    import kmc
    from kmc import datagen
    sync = device_sync()
    n, d, k = 65536, 784, 10
    steps = a.steps if a.steps is not None else 100
    X = datagen.sparse_unit_blobs(n, d, k)[0]
    rng = np.random.default_rng(7)
    lab0 = rng.integers(0, k, size=n).astype(np.int32)
    forms = {}
    cases = [(kern, store, "sorted") for kern in ("polynomial", "rbf") for store in ("fp32", "bf16")]
    cases += [("polynomial", store, "dense") for store in ("fp32", "bf16")]
    for kernel, store, product in cases:
        build = "fp32" if store == "fp32" else "bf16x3"
        with kmc.KernelEngine(n, d, k, kernel=kernel, store=store, product=product,
                              gamma=1.0 / d if kernel == "rbf" else 0.0, build=build) as e:
            t0 = time.perf_counter()
            e.set_points(X)
            sync()
            build_s = time.perf_counter() - t0
            build_dev_ms = e.setup_times()["operands_ms"]
            e.set_labels(lab0)
            e.run(a.warmup, check_converged=False)
            e.set_labels(lab0)
            sync()
            # step rate from a pass without kernel events, the product's
            # launch time from a second pass of the same steps with them
            t0 = time.perf_counter()
            e.run(steps, check_converged=False)
            sync()
            wall = time.perf_counter() - t0
            lab = e.labels()
            e.set_labels(lab0)
            sync()
            r = e.run(steps, check_converged=False, time_kernels=True)
            sync()
            assert np.array_equal(e.labels(), lab)
        its = int(r["iterations"])
        t_spmm = r["ms_spmm"] / 1e3 / max(1, int(r["spmm_launches"]))
        kbytes = n * n * (4 if store == "fp32" else 2)
        name = "%s_%s" % (kernel, store) + ("_dense" if product == "dense" else "")
        forms[name] = {
            "value": round(its / wall, 3), "ms_per_step": round(wall * 1e3 / max(1, its), 4),
            "build_s": round(build_s, 3), "build_device_ms": round(build_dev_ms, 2),
            "gram_form": build, "objective": r["objective"],
            "spmm_mean_launch_ms": round(t_spmm * 1e3, 4),
            "spmm_hbm_frac": round(kbytes / t_spmm / HBM_PEAK, 4),
            "cluster_sizes": np.bincount(lab, minlength=k).tolist()}
        log("kernel k-means", name, forms[name])
    head = forms["polynomial_bf16"]
    t_spmm = head["spmm_mean_launch_ms"] / 1e3
    kbytes = n * n * 2
    out = {"metric": "kernel k-means iter/s, N=65536 D=784 K=10 (polynomial, K in bf16)",
           "value": head["value"], "unit": "iter/s", "n_gpus": 1, "steps": steps,
           "warmup": a.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
           "config": {"workload": "kernel k-means, sparse U[0,1] blobs N=65536 D=784 K=10, "
                                  "%d iterations" % steps},
           "roofline": {"bound": "hbm",
                        "kernel": "k_kk_rowsum + k_kk_colsum (cluster-sorted S = K.H, K bf16)",
                        "achieved": round(kbytes / t_spmm / 1e9, 1), "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": round(kbytes / t_spmm / HBM_PEAK, 4),
                        "traffic": None, "algorithmic_bytes_per_launch": kbytes,
                        "mean_launch_ms": round(t_spmm * 1e3, 4)},
           "forms": forms}
    print(json.dumps(out))


def c4_problems(count=512, seed=20261205):
    """512 independent problems: N log-uniform in [1e3, 2e5], D in
    {16, 32, 64, 128}, K uniform in [4, 128], Gaussian blobs with per-instance
    seeds, K seed-chosen rows as initial centroids (kmc.datagen)."""
    from kmc import datagen
    return datagen.batch_problems(count, seed)


def bench_batch(a):
    """Batched Lloyd (512 problems of uneven n, d, k, 30 iterations each, no
    convergence test): every stage of an iteration is one launch for the
    whole batch. value = problem-iterations per second of the headline form;
    FP32 beside it; a sample of problems run one by one gives the
    unbatched rate for comparison."""
    import kmc
    sync = device_sync()
    steps = a.steps if a.steps is not None else 30
    log("generating 512 problems")
    probs = c4_problems()
    nprob = len(probs)
    npts = sum(X.shape[0] for X, _ in probs)
    flops = sum(2.0 * X.shape[0] * C.shape[0] * X.shape[1] for X, C in probs)
    forms = {}
    for prec in (a.prec, "fp32") if a.prec != "fp32" else ("fp32",):
        with kmc.Batch(probs, tol=0.0, precision=prec) as b:
            # both launch paths of kmc_batch_run, each after its own warm-up:
            # eager launches and graph replays (each chunk of iterations one
            # captured graph); the faster gives the form's rate
            walls = {}
            for graph in ((False,) if a.eager else (False, True)):
                b.run(min(a.warmup, steps), check_converged=False, graph=graph)
                sync()
                t0 = time.perf_counter()
                res, ms = b.run(steps, check_converged=False, graph=graph)
                sync()
                walls["graph" if graph else "eager"] = (time.perf_counter() - t0, ms, res)
            path = min(walls, key=lambda k: walls[k][0])
            wall, ms, res = walls[path]
            its = sum(int(r["iterations"]) for r in res)
            forms[prec] = {"problem_iters_per_s": round(its / wall, 1),
                           "batch_iters_per_s": round(steps / wall, 3),
                           "ms_per_batch_iteration": round(wall * 1e3 / steps, 4),
                           "device_ms": round(ms, 3),
                           "distance_tflops": round(flops * steps / wall / 1e12, 2),
                           "launch_path": path}
            for k, (w, _, _) in walls.items():
                forms[prec]["%s_ms_per_batch_iteration" % k] = round(w * 1e3 / steps, 4)
            log("batch", prec, forms[prec])
    # unbatched comparison: the first 16 problems, one engine each
    sample = probs[:16]
    t_lone = 0.0
    for X, C0 in sample:
        with kmc.Engine(X.shape[0], X.shape[1], C0.shape[0], tol=0.0, precision=a.prec,
                        update="tree") as e:
            e.set_points(X)
            e.set_centroids(C0)
            e.run(1, check_converged=False)
            sync()
            t0 = time.perf_counter()
            e.run(steps, check_converged=False)
            sync()
            t_lone += time.perf_counter() - t0
    lone_rate = len(sample) * steps / t_lone
    head = forms[a.prec]
    wall_head = head["ms_per_batch_iteration"] * steps / 1e3
    ubytes = sum(X.shape[0] * (X.shape[1] * 4 + 4) for X, _ in probs)
    out = {"metric": "batched Lloyd problem-iterations/s, 512 problems N~logU[1e3,2e5] "
                     "D in {16,32,64,128} K~U[4,128]",
           "value": head["problem_iters_per_s"], "unit": "problem-iter/s", "n_gpus": 1,
           "steps": steps, "warmup": a.warmup, "ms_per_step": head["ms_per_batch_iteration"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": DTYPE[a.prec],
           "data": "synthetic",
           "config": {"workload": "512 batched Lloyd problems, %d points in total, %d "
                                  "iterations each, tol=0" % (npts, steps),
                      "problems": nprob, "points": npts, "form": a.prec},
           "roofline": {"bound": "hbm", "kernel": "whole batch iteration",
                        "achieved": round(ubytes * steps / wall_head / 1e9, 1),
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": round(ubytes * steps / wall_head / HBM_PEAK, 4),
                        "traffic": None, "algorithmic_bytes_per_launch": ubytes,
                        "note": "algorithmic bytes of one batch iteration (one pass over "
                                "every problem's X, n*d*4 + n*4) / wall time of an "
                                "iteration; problems are packed per padded width "
                                "(16/32/64/128) and each to a multiple of 1024 points, "
                                "so the update streams sum(ceil(n/1024)*1024)*DP*4 bytes"},
           "distance_flops_per_iteration": flops,
           "forms": forms,
           "unbatched_sample": {"problems": len(sample),
                                "problem_iters_per_s": round(lone_rate, 1),
                                "batched_speedup": round(head["problem_iters_per_s"] /
                                                         lone_rate, 2)}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c2", "c5", "c3", "c4"],
                    help="c2: N=1M D=128 K=256 (default); c5: N=4M D=256 K=4096; "
                         "c3: kernel k-means N=65536 D=784 K=10; c4: 512 batched problems")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed iterations (default: 50 for c2, 20 for c5, 100 for c3)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prec", default="bf16", choices=["bf16x3", "bf16", "fp32", "fp8"])
    ap.add_argument("--update", default=None, choices=["ordered", "tree", "fixed"],
                    help="centroid update: exact fixed-point sums kept across the run "
                         "(moved points only once few move; the C2 default), FP64 tree "
                         "over the sorted members (the C5 default), or the reference's "
                         "float summation order")
    ap.add_argument("--no-x3", action="store_true", help="skip the bf16x3 comparison form")
    ap.add_argument("--no-fp8", action="store_true", help="skip the fp8 comparison form")
    ap.add_argument("--eager", action="store_true",
                    help="launch kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-copy", type=int, default=0,
                    help="only run a device copy of this many bytes (PMC calibration)")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU: this build has no multi-GPU path "
                 "(--gpus %d requested)" % a.gpus)
    if a.profile_copy:
        profile_copy(a.profile_copy)
        return
    if a.workload == "c3":
        bench_kernel_kmeans(a)
        return
    if a.workload == "c4":
        bench_batch(a)
        return
    global N, D, K, DATA_SEED
    W = WORKLOADS[a.workload]
    N, D, K, DATA_SEED = W["n"], W["d"], W["k"], W["seed"]
    if a.steps is None:
        a.steps = W["steps"]
    if a.update is None:
        a.update = W["update"]

    import kmc
    from kmc import datagen

    if kmc.device_count() < 1:
        sys.exit("bench.py: no GPU visible")
    sync = device_sync()

    log("generating N=%d D=%d K=%d blobs" % (N, D, K))
    X, _, _ = getattr(datagen, W["gen"])(N, D, K, DATA_SEED)
    C0 = datagen.init_rows(X, K, INIT_SEED)

    log("timing %s: warmup %d, steps %d" % (a.prec, a.warmup, a.steps))
    head = run_form(kmc, X, C0, a.prec, a.steps, a.warmup, sync, a.update, not a.eager)
    fp32 = None
    if not a.no_fp32 and a.prec != "fp32":
        log("timing fp32: warmup %d, steps %d" % (a.warmup, a.steps))
        fp32 = run_form(kmc, X, C0, "fp32", a.steps, a.warmup, sync, a.update, not a.eager)
    x3 = None
    if not a.no_x3 and a.prec != "bf16x3":
        log("timing bf16x3: warmup %d, steps %d" % (a.warmup, a.steps))
        x3 = run_form(kmc, X, C0, "bf16x3", a.steps, a.warmup, sync, a.update, not a.eager)
    f8 = None
    if not a.no_fp8 and a.prec != "fp8":
        log("timing fp8: warmup %d, steps %d" % (a.warmup, a.steps))
        f8 = run_form(kmc, X, C0, "fp8", a.steps, a.warmup, sync, a.update, not a.eager)

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    O = Oracle()
    C_end = head["C_end"]
    inert64 = O.f64_inertia(X, C_end, head["labels_final"].astype(np.uint32))
    inert_err = abs(head["inertia_gpu"] - inert64) / inert64

    r = head["res"]
    iters = int(r["iterations"])
    ips = iters / head["wall_s"]
    launches = int(r["assign_launches"])
    t_assign = r["ms_assign"] / 1e3 / max(1, launches)
    t_update = r["ms_update"] / 1e3 / max(1, int(r["update_launches"]))
    flops = 2.0 * N * K * D
    achieved = flops / t_assign
    # the two largest kernels of an iteration, each against its own roof
    dp = 128 if D <= 128 else 256
    kname = assign_kernel_name(dp, a.prec, head["tuning"]["chosen"])
    roof_assign = {"bound": "mfma", "kernel": kname,
                   "achieved": round(achieved / 1e12, 3), "peak": PEAK[a.prec] / 1e12,
                   "unit": "TFLOP/s", "frac": round(achieved / PEAK[a.prec], 4),
                   "traffic": traffic_from_profiles("assign", a.prec, a.workload, kname),
                   "algorithmic_flops_per_launch": flops,
                   "mean_launch_ms": round(t_assign * 1e3, 5), "launches_timed": launches}
    ubytes = N * D * 4 + N * 4 + K * D * 4  # X rows once, member ids, centroids
    uname = {"tree": "k_update_tree_rows<128, false>" if dp <= 128 else "k_update_tree<4, false>",
             "ordered": "k_update", "fixed": "k_update_fx<64, true>"}[a.update]
    update_note = None
    if a.update == "fixed" and int(r["full_passes"]) < iters:
        # incremental: full passes stream X, labels and the previous labels;
        # the other iterations read labels and previous labels and the rows
        # of the points that moved
        full = int(r["full_passes"])
        total = full * N * (D * 4 + 12) + (iters - full) * N * 8 + \
            int(r["moved_points"]) * D * 4
        ubytes = total / max(1, iters)
        uname = "k_update_fx<64, true> (incremental)"
        update_note = ("%d of %d iterations streamed all points; the others updated the "
                       "cluster totals from %d moved points in all" %
                       (full, iters, int(r["moved_points"])))
    roof_update = {"bound": "hbm", "kernel": uname,
                   "achieved": round(ubytes / t_update / 1e9, 1), "peak": HBM_PEAK / 1e9,
                   "unit": "GB/s", "frac": round(ubytes / t_update / HBM_PEAK, 4),
                   "traffic": traffic_from_profiles("update_" + a.update, a.prec, a.workload),
                   "algorithmic_bytes_per_launch": ubytes,
                   "mean_launch_ms": round(t_update * 1e3, 5),
                   "launches_timed": int(r["update_launches"])}
    if update_note:
        roof_update["incremental"] = update_note
    if t_update > t_assign:
        roof, roof_other = roof_update, roof_assign
    else:
        roof, roof_other = roof_assign, roof_update

    out = {
        "metric": W["metric"],
        "value": round(ips, 3),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["wall_s"] * 1e3 / max(1, iters), 5),
        "ms_per_step_with_kernel_events": round(head["wall_events_s"] * 1e3 / max(1, iters), 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[a.prec],
        "data": "synthetic",
        "config": {"workload": "Lloyd k-means, %s, maxiter=%d tol=0" % (W["text"], a.steps),
                   "n": N, "d": D, "k": K, "form": a.prec, "seed": DATA_SEED,
                   "init": "%d distinct seed-chosen rows (seed %d)" % (K, INIT_SEED),
                   "update": {"ordered": "ordered (reference summation order)",
                              "tree": "tree (FP64 chunk partials)",
                              "fixed": "fixed (exact fixed-point totals, moved points "
                                       "only after the first iteration)"}[a.update]},
        "roofline": roof,
        "roofline_second_kernel": roof_other,
        "assign_variant": {"chosen": head["tuning"]["chosen"],
                           "threads": head["tuning"]["threads"][head["tuning"]["chosen"]],
                           "autotune_ms": [round(v, 4) for v in head["tuning"]["ms"]]},
        "inertia_rel_err_vs_fp64": inert_err,
        "inertia_fp64": inert64,
        "refined_points_last_iter": {"two_candidates": int(r["refined_two"]),
                                     "all_k": int(r["refined_full"])},
    }
    if fp32 is not None:
        f = fp32["res"]
        f_ips = int(f["iterations"]) / fp32["wall_s"]
        f_t = f["ms_assign"] / 1e3 / max(1, int(f["assign_launches"]))
        out["fp32"] = {"value": round(f_ips, 3), "ms_per_step": round(fp32["wall_s"] * 1e3 /
                                                                      max(1, f["iterations"]), 5),
                       "assign_mean_launch_ms": round(f_t * 1e3, 5),
                       "assign_tflops": round(flops / f_t / 1e12, 3),
                       "assign_frac_fp32_peak": round(flops / f_t / PEAK["fp32"], 4)}
        out["speedup_vs_fp32"] = round(ips / f_ips, 3)
        out["identical_to_fp32"] = bool(np.array_equal(fp32["C_end"], C_end) and
                                        np.array_equal(fp32["labels_end"], head["labels_end"]))
    if x3 is not None:
        f = x3["res"]
        f_ips = int(f["iterations"]) / x3["wall_s"]
        f_t = f["ms_assign"] / 1e3 / max(1, int(f["assign_launches"]))
        out["bf16x3"] = {"value": round(f_ips, 3),
                         "ms_per_step": round(x3["wall_s"] * 1e3 / max(1, f["iterations"]), 5),
                         "assign_mean_launch_ms": round(f_t * 1e3, 5),
                         "assign_frac_bf16_peak": round(3 * flops / f_t / PEAK["bf16"], 4),
                         "identical": bool(np.array_equal(x3["C_end"], C_end) and
                                           np.array_equal(x3["labels_end"], head["labels_end"]))}
    if f8 is not None:
        f = f8["res"]
        f_ips = int(f["iterations"]) / f8["wall_s"]
        f_t = f["ms_assign"] / 1e3 / max(1, int(f["assign_launches"]))
        out["fp8"] = {"value": round(f_ips, 3),
                      "ms_per_step": round(f8["wall_s"] * 1e3 / max(1, f["iterations"]), 5),
                      "assign_mean_launch_ms": round(f_t * 1e3, 5),
                      "assign_frac_fp8_peak": round(flops / f_t / PEAK["fp8"], 4),
                      "refined_points_last_iter": {"two_candidates": int(f["refined_two"]),
                                                   "all_k": int(f["refined_full"])},
                      "identical": bool(np.array_equal(f8["C_end"], C_end) and
                                        np.array_equal(f8["labels_end"], head["labels_end"]))}
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(X, C0, W["cpu_sample"], W["cpu_iters"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
