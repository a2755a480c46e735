"""Pins the CPU oracle (oracle/kmc_oracle.cpp) before it is trusted as the
checker of the HIP path.

  * bit-exact against the golden fixtures written from the compiled reference
    CPU k-means (tests/golden/make_golden.py), and against the compiled
    reference itself (oracle/_ref) on fresh random inputs where it is built;
  * the worked distance examples of the reference's report
    (files/GPU_Computing_Report.pdf p.6);
  * FP64 parts against closed forms, brute force and Lloyd monotonicity;
  * kernel k-means against the reference unit test's O(n^3) host formula
    (tests/unit_kernels.cu:365-395) and the linear-kernel identity; its Lloyd
    loop against a hand-worked case and explicit Lloyd on finite feature maps;
  * the reference GPU engine's loop rules (orc_gpu_lloyd) against hand-worked
    1-D cases and run invariants.
"""
import numpy as np
import pytest

from conftest import load_golden


def test_report_known_answers(oracle):
    # files/GPU_Computing_Report.pdf p.6: 1-D c=7, p=3 -> 16, p=1 -> 36;
    # 2-D c=(1,4), p=(5,2) -> 20; 3-D c=(5,2,3), p=(4,3,2) -> 3
    cases = [([[3.0], [1.0]], [[7.0]], [16.0, 36.0]),
             ([[5.0, 2.0]], [[1.0, 4.0]], [20.0]),
             ([[4.0, 3.0, 2.0]], [[5.0, 2.0, 3.0]], [3.0])]
    for X, C, want in cases:
        X = np.array(X, np.float32)
        C = np.array(C, np.float32)
        assert oracle.distances(X, C)[:, 0].tolist() == want
        assert oracle.f64_distances(X, C)[:, 0].tolist() == want


def test_steps_match_golden(oracle):
    g = load_golden("ref_steps.npz")
    for i in range(int(g["ncases"])):
        X, C = g["%d_X" % i], g["%d_C" % i]
        D = oracle.distances(X, C)
        assert np.array_equal(D, g["%d_D" % i]), i
        lab, cnt = oracle.argmin(D)
        assert np.array_equal(lab, g["%d_labels" % i]), i
        assert np.array_equal(cnt, g["%d_counts" % i]), i
        Cn, _ = oracle.centroids(X, lab, C.shape[0])
        assert np.array_equal(Cn, g["%d_Cnew" % i]), i


def test_golden_pins_summation_order(oracle):
    """A plausible mistake -- pairwise (numpy) instead of sequential float
    sums -- must fail the golden distances, else the fixture pins nothing."""
    g = load_golden("ref_steps.npz")
    differs = 0
    for i in range(int(g["ncases"])):
        X, C = g["%d_X" % i], g["%d_C" % i]
        diff = X[:, None, :] - C[None, :, :]
        wrong = np.sum(diff * diff, axis=2, dtype=np.float32)
        differs += int(not np.array_equal(wrong, g["%d_D" % i]))
    assert differs > 0


@pytest.mark.parametrize("name", ["ref_c1_blobs.npz", "ref_digits.npz"])
def test_full_runs_match_golden(oracle, name):
    g = load_golden(name)
    X = g["X"]
    C0 = oracle.init_centroids(X, int(g["k"]), int(g["seed"]))
    assert np.array_equal(C0, g["C0"])
    conv, lab, C, _, _ = oracle.lloyd(X, C0, float(g["tol"]), int(g["maxiter"]))
    assert conv == int(g["converged"])
    assert np.array_equal(lab, g["labels"])
    assert np.array_equal(C, g["C"])


def test_runs_from_centroids_match_golden(oracle):
    g = load_golden("ref_runs.npz")
    for i in range(int(g["ncases"])):
        conv, lab, C, _, _ = oracle.lloyd(g["%d_X" % i], g["%d_C0" % i], float(g["%d_tol" % i]),
                                          int(g["%d_maxiter" % i]))
        assert conv == int(g["%d_converged" % i]), i
        assert np.array_equal(lab, g["%d_labels" % i]), i
        assert np.array_equal(C, g["%d_C" % i]), i


def test_against_compiled_reference(oracle):
    from oracle import Reference
    if not Reference.available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    R = Reference()
    rng = np.random.default_rng(7)
    for n, d, k in [(40, 3, 5), (300, 11, 9), (700, 40, 23)]:
        X = rng.normal(0, 10, size=(n, d)).astype(np.float32)
        C = X[rng.choice(n, k, replace=False)]
        D, lab, cnt, Cn = R.step(X, C)
        D2 = oracle.distances(X, C)
        lab2, cnt2 = oracle.argmin(D2)
        Cn2, _ = oracle.centroids(X, lab2, k)
        assert np.array_equal(D, D2) and np.array_equal(lab, lab2)
        assert np.array_equal(cnt, cnt2) and np.array_equal(Cn, Cn2)
        for tol in (0.0, 1e-3, 1.0):
            conv, Cr, lr = R.run_from(X, C, tol, 40)
            conv2, lo, Co, _, _ = oracle.lloyd(X, C, tol, 40)
            assert conv == conv2 and np.array_equal(lr, lo) and np.array_equal(Cr, Co)
        # shift test exactly at the tolerance
        C2 = C.copy()
        C2[0, 0] += np.float32(0.5)
        for tol in (0.5, np.nextafter(np.float32(0.5), np.float32(0)), 0.25):
            assert R.converged(X, C2, C, tol) == oracle.converged(C2, C, tol)


def test_empty_cluster_becomes_zero(oracle):
    X = np.array([[0, 0], [0, 1], [10, 10]], np.float32)
    C = np.array([[0, 0.5], [10, 10], [100, 100]], np.float32)
    lab, cnt = oracle.argmin(oracle.distances(X, C))
    Cn, _ = oracle.centroids(X, lab, 3)
    assert cnt.tolist() == [2, 1, 0]
    assert Cn[2].tolist() == [0.0, 0.0]


def test_f64_closed_forms(oracle):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 5, size=(500, 13)).astype(np.float32)
    C = X[:7] + 0.25
    D = oracle.f64_distances(X, C)
    # brute force argmin in numpy vs oracle assign
    lab, margin, inert = oracle.f64_assign(X, C)
    assert np.array_equal(lab, np.argmin(D, axis=1))
    srt = np.sort(D, axis=1)
    assert np.allclose(margin, srt[:, 1] - srt[:, 0], rtol=0, atol=1e-9)
    # inertia identity: direct form vs expanded GEMM form in FP64
    Xd, Cd = X.astype(np.float64), C.astype(np.float64)
    expanded = (Xd ** 2).sum() - 2 * np.sum(Xd * Cd[lab]) + (Cd[lab] ** 2).sum()
    assert abs(inert - expanded) <= 1e-9 * inert
    assert abs(oracle.f64_inertia(X, C, lab) - inert) <= 1e-12 * inert
    # means of the one-hot groups
    Cm, cnt = oracle.f64_centroids(X, lab, 7)
    for c in range(7):
        if cnt[c]:
            assert np.allclose(Cm[c], Xd[lab == c].mean(axis=0).astype(np.float32), rtol=1e-6,
                               atol=1e-6)


def test_f64_lloyd_inertia_monotone(oracle):
    from kmc import datagen
    X, _, _ = datagen.blobs(2000, 8, 12, 5, -3, 3, 0.5, 1.5)
    C = datagen.init_rows(X, 12, 9)
    last = np.inf
    for _ in range(15):
        lab, _, inert = oracle.f64_assign(X, C)
        assert inert <= last * (1 + 1e-12)
        last = inert
        C, _ = oracle.f64_centroids(X, lab, 12)


def _unit_test_formula(K, labels, k):
    """tests/unit_kernels.cu:365-395, restated: K_ii - 2 sum_{j in c} K_ij/|c|
    + sum_{j,l in c} K_jl / |c|^2 by explicit loops over clusters."""
    n = K.shape[0]
    cnt = np.bincount(labels, minlength=k).astype(np.float64)
    out = np.zeros((n, k))
    for c in range(k):
        idx = np.nonzero(labels == c)[0]
        cross = K[:, idx].sum(axis=1) / cnt[c]
        within = K[np.ix_(idx, idx)].sum() / cnt[c] ** 2
        out[:, c] = np.diag(K) - 2 * cross + within
    return out


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_kernel_kmeans_oracle(oracle, kind):
    rng = np.random.default_rng(11 + kind)
    n, d, k = 60, 5, 4
    X = rng.uniform(0, 1, size=(n, d)).astype(np.float32)
    gamma = 1.0 / d
    K = oracle.kernel_matrix(X, kind, gamma)
    G = X.astype(np.float64) @ X.astype(np.float64).T
    sq = np.maximum((G.diagonal()[:, None] + G.diagonal()[None, :] - 2 * G), 0)
    np.fill_diagonal(sq, 0)
    want = {0: G, 1: (G + 1) ** 2, 2: np.tanh(G + 1), 3: np.exp(-gamma * sq)}[kind]
    assert np.allclose(K, want, rtol=1e-12, atol=1e-12)
    labels = (np.arange(n) % k).astype(np.uint32)
    D = oracle.kk_distances(K, labels, k, add_diag=True)
    assert np.allclose(D, _unit_test_formula(K, labels, k), rtol=1e-10, atol=1e-10)
    # without the diagonal (the reference's form) the argmin is unchanged
    D0 = oracle.kk_distances(K, labels, k, add_diag=False)
    assert np.array_equal(np.argmin(D, 1), np.argmin(D0, 1))
    if kind == 0:
        # linear kernel: kernel-space distance == ||x - mean||^2
        Xd = X.astype(np.float64)
        mu = np.stack([Xd[labels == c].mean(0) for c in range(k)])
        direct = ((Xd[:, None, :] - mu[None]) ** 2).sum(-1)
        assert np.allclose(D, direct, rtol=1e-9, atol=1e-9)


def _plain_plusplus(X, k, u):
    """Textbook D^2 sampling (Arthur & Vassilvitskii 2007) with a plain
    cumulative sum: the first point whose running sum of squared distances
    exceeds u_r times the total."""
    Xd = X.astype(np.float64)
    n = X.shape[0]
    chosen = [min(int(u[0] * n), n - 1)]
    D2 = None
    for r in range(1, k):
        d = ((Xd - Xd[chosen[-1]]) ** 2).sum(1)
        D2 = d if D2 is None else np.minimum(D2, d)
        total = D2.sum()
        cs = np.cumsum(D2)
        hit = np.nonzero((cs > u[r] * total) & (D2 > 0))[0]
        chosen.append(int(hit[0]))
    return np.array(chosen)


def test_plusplus_matches_plain_sampler_on_exact_data(oracle):
    """On integer-valued data every summation order is exact, so the oracle's
    blocked sums must pick exactly what a plain cumulative sum picks."""
    from oracle import splitmix_uniforms
    rng = np.random.default_rng(3)
    for n, d, k, seed in [(700, 5, 12, 1), (5000, 8, 20, 99), (3100, 3, 7, 2026)]:
        X = rng.integers(-50, 51, size=(n, d)).astype(np.float32)
        u = splitmix_uniforms(seed, k)
        got, C = oracle.plusplus(X, k, u)
        want = _plain_plusplus(X, k, u)
        assert np.array_equal(got, want), (n, d, k, got, want)
        assert np.array_equal(C, X[got])
        assert len(set(got.tolist())) == k


def _plain_greedy_plusplus(X, k, u, s):
    """Greedy k-means++ (s D^2 trials per round, the candidate leaving the
    smallest potential kept, the first on ties) with plain cumulative sums."""
    Xd = X.astype(np.float64)
    n = X.shape[0]
    chosen = [min(int(u[0] * n), n - 1)]
    D2 = None
    for r in range(1, k):
        d = ((Xd - Xd[chosen[-1]]) ** 2).sum(1)
        D2 = d if D2 is None else np.minimum(D2, d)
        total = D2.sum()
        cs = np.cumsum(D2)
        best, best_pot = None, None
        for t in range(s):
            uu = u[1 + (r - 1) * s + t]
            cand = int(np.nonzero((cs > uu * total) & (D2 > 0))[0][0])
            pot = np.minimum(D2, ((Xd - Xd[cand]) ** 2).sum(1)).sum()
            if best is None or pot < best_pot:
                best, best_pot = cand, pot
        chosen.append(best)
    return np.array(chosen)


def test_greedy_plusplus_matches_plain_sampler_on_exact_data(oracle):
    """Greedy k-means++ (the reference's s = 2 + ceil(ln k) scheme,
    src/kmeans.cu:463-769): on integer-valued data the oracle's blocked sums
    must pick what plain cumulative sums pick; one trial is plain k-means++."""
    from oracle import splitmix_uniforms
    rng = np.random.default_rng(5)
    for n, d, k, s, seed in [(900, 4, 10, 5, 11), (3000, 6, 16, 3, 12)]:
        X = rng.integers(-40, 41, size=(n, d)).astype(np.float32)
        u = splitmix_uniforms(seed, 1 + (k - 1) * s)
        got, C = oracle.plusplus(X, k, u, trials=s)
        want = _plain_greedy_plusplus(X, k, u, s)
        assert np.array_equal(got, want), (n, d, k, s, got, want)
        assert np.array_equal(C, X[got])
    X = rng.standard_normal((1500, 8)).astype(np.float32)
    u = splitmix_uniforms(3, 12)
    assert np.array_equal(oracle.plusplus(X, 12, u, trials=1)[0], oracle.plusplus(X, 12, u)[0])


def test_plusplus_first_pick_and_distinct(oracle):
    from oracle import splitmix_uniforms
    rng = np.random.default_rng(4)
    X = rng.standard_normal((2000, 16)).astype(np.float32)
    u = splitmix_uniforms(7, 30)
    got, _ = oracle.plusplus(X, 30, u)
    assert got[0] == int(u[0] * 2000)
    assert len(set(got.tolist())) == 30
    # splitmix64 reference values (seed 0): first outputs of the published
    # generator, 0xe220a8397b1dcdaf and 0x6e789e6aa1b965f4
    v = splitmix_uniforms(0, 2)
    assert v[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert v[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53


def test_explicit_lloyd_known_answers(oracle):
    """Hand-worked runs of the explicit-Lloyd rules (stop on the sum of
    squared centroid changes, stop on the cost ratio, empty cluster keeps its
    centroid), experiments/raft-bench/src/kmeans.cu:158-212."""
    X = np.array([[0.0], [1.0], [10.0], [11.0]], np.float32)
    C0 = np.array([[0.0], [1.0]], np.float32)
    # t=1: labels 0,1,1,1 -> C = (0, 22/3), cost 181; t=2: labels 0,0,1,1 ->
    # C = (0.5, 10.5), cost 1 + (10 - 22/3)^2 + (11 - 22/3)^2; t=3: same labels, C
    # unchanged (shift sum 0), cost 1; t=4: cost 1 again, ratio 1
    t, lab, C, costs = oracle.lloyd_explicit(X, C0, 1e-4, 50, check=True)
    assert t == 3
    assert np.array_equal(lab, np.array([0, 0, 1, 1], np.uint32))
    assert np.array_equal(C, np.array([[0.5], [10.5]], np.float32))
    c1 = float(np.float32(22.0 / 3.0))  # the float centroid of iteration 1
    assert costs[0] == 181.0 and abs(costs[1] - (1 + (10 - c1) ** 2 + (11 - c1) ** 2)) < 1e-12
    assert costs[2] == 1.0
    t, _, _, costs = oracle.lloyd_explicit(X, C0, 1e-4, 50, check=False)
    assert t == 4 and costs[3] == costs[2] == 1.0
    # the second cluster is empty from the start and keeps its centroid
    X2 = np.array([[0.0], [1.0], [2.0]], np.float32)
    C2 = np.array([[1.0], [100.0]], np.float32)
    t, lab, C, _ = oracle.lloyd_explicit(X2, C2, 1e-4, 50, check=True)
    assert np.array_equal(lab, np.zeros(3, np.uint32))
    assert np.array_equal(C, np.array([[1.0], [100.0]], np.float32))
    assert t == 1  # nothing moved at iteration 1


# ---- the reference GPU engine's loop rules (orc_gpu_lloyd) -----------------
# Hand-worked 1-D cases. Scores are sum_i min_c ||x_i - mu_c||^2 - sum_i x_i^2.

def _x(vals):
    return np.array(vals, np.float32).reshape(-1, 1)


def test_gpu_rules_known_answers_two_clusters(oracle):
    # X = 0 1 10 11 20 21, labels i % 2 -> mu = (10, 11).
    # it 1: labels 0 0 0 1 1 1, sum min d = 100+81+0+0+81+100 = 362,
    #       score = 362 - 1063 = -701; mu = (11/3, 52/3)
    # it 2: labels unchanged, sum min d = 364/3, score = -2825/3
    # it 3: same score -> |delta| = 0
    X = _x([0, 1, 10, 11, 20, 21])
    lab0 = np.arange(6) % 2
    conv, lab, C, sc, ran = oracle.gpu_lloyd(X, lab0, 2, 1e-4, 10)
    assert lab.tolist() == [0, 0, 0, 1, 1, 1]
    assert (conv, ran) == (3, 3)
    assert sc.tolist() == [np.float32(-701.0), np.float32(-2825.0 / 3.0),
                           np.float32(-2825.0 / 3.0)]
    np.testing.assert_allclose(C[:, 0], [11 / 3, 52 / 3], rtol=1e-6)
    # |score_2 - score_1| = 240.67: tol 300 stops at 2, tol 240 does not
    assert oracle.gpu_lloyd(X, lab0, 2, 300.0, 10)[0] == 2
    assert oracle.gpu_lloyd(X, lab0, 2, 240.0, 10)[0] == 3
    # no test without check, and none at the last iteration
    conv, _, _, _, ran = oracle.gpu_lloyd(X, lab0, 2, 1e-4, 10, check=False)
    assert (conv, ran) == (10, 10)
    assert oracle.gpu_lloyd(X, lab0, 2, 300.0, 2)[:1] == (2,)
    conv, _, _, _, ran = oracle.gpu_lloyd(X, lab0, 2, 300.0, 1)
    assert (conv, ran) == (1, 1)


def test_gpu_rules_empty_cluster_is_never_chosen_again(oracle):
    # X = 0 1 2 6 7, k = 3, labels i % 3 -> mu = (3, 4, 2).
    # it 1: labels 2 2 2 1 1 (cluster 0 empties), sum min d = 4+1+0+4+9 = 18,
    #       score = 18 - 90 = -72; mu_1 = 6.5, mu_2 = 1
    # it 2: cluster 0 has distance +inf, so x = 0 stays with cluster 2:
    #       sum min d = 1+0+1+0.25+0.25 = 2.5, score = -87.5
    # The reference CPU rule (empty -> zero vector, still selectable) would
    # move x = 0 to cluster 0 in iteration 2.
    X = _x([0, 1, 2, 6, 7])
    lab0 = np.arange(5) % 3
    conv, lab, C, sc, ran = oracle.gpu_lloyd(X, lab0, 3, 1e-4, 20)
    assert lab.tolist() == [2, 2, 2, 1, 1]
    assert (conv, ran) == (3, 3)
    assert sc.tolist() == [-72.0, -87.5, -87.5]
    assert C[:, 0].tolist() == [0.0, 6.5, 1.0]
    assert oracle.gpu_lloyd(X, lab0, 3, 20.0, 20)[0] == 2
    # the CPU rule from the same means does re-pick the zero vector
    C1 = np.array([[0.0], [6.5], [1.0]], np.float32)
    lab_cpu, _ = oracle.argmin(oracle.distances(X, C1))
    assert lab_cpu.tolist() == [0, 2, 2, 1, 1]


def test_gpu_rules_invariants_on_blobs(oracle):
    """Invariants of a long run from labels i % k (check off): the scores never
    rise (Lloyd monotonicity; dropping an empty cluster cannot raise the
    cost), the last score is the FP64 inertia of the last labels against the
    means they were assigned to, minus sum ||x||^2, and the last labels are a
    fixed point: assigning against the final means of the clusters still in
    use (FP64, first minimum) reproduces them."""
    rng = np.random.default_rng(3)
    cen = rng.uniform(-20, 20, (5, 3))
    X = (cen[rng.integers(0, 5, 400)] + rng.normal(0, 1.0, (400, 3))).astype(np.float32)
    lab0 = np.arange(400) % 5
    conv, lab, C, sc, ran = oracle.gpu_lloyd(X, lab0, 5, 0.0, 40, check=False)
    assert ran == 40 and conv == 40
    s64 = sc.astype(np.float64)
    assert np.all(np.diff(s64) <= 1e-6 * np.abs(s64[:-1]))
    live = np.bincount(lab.astype(np.int64), minlength=5) > 0
    D = oracle.f64_distances(X, C)
    D[:, ~live] = np.inf
    assert np.array_equal(D.argmin(1), lab)
    sx = float((X.astype(np.float64) ** 2).sum())
    # the fixed point: the last assignment was made against these same means
    assert sc[-1] == np.float32(oracle.f64_inertia(X, C, lab) - sx)


# ---- kernel k-means Lloyd (orc_kk_lloyd) ------------------------------------

def test_kk_lloyd_known_answer(oracle):
    # linear kernel, X = 0 1 10 11 20 21, labels i % 2: iteration 1 moves to
    # 0 0 0 1 1 1, iteration 2 changes nothing -> 2 iterations
    X = _x([0, 1, 10, 11, 20, 21])
    K = oracle.kernel_matrix(X, 0)
    it, lab = oracle.kk_lloyd(K, np.arange(6) % 2, 2, 50)
    assert (it, lab.tolist()) == (2, [0, 0, 0, 1, 1, 1])
    it, lab = oracle.kk_lloyd(K, np.arange(6) % 2, 2, 1)
    assert (it, lab.tolist()) == (1, [0, 0, 0, 1, 1, 1])


def _explicit_lloyd_to_stable(F, lab0, k, maxiter):
    """FP64 Lloyd on explicit features F from labels: centroids are label
    means, first-minimum assignment over non-empty clusters, stop when no
    label changes. Returns (iterations, labels)."""
    lab = np.array(lab0, np.int64)
    it = 0
    while it < maxiter:
        it += 1
        mu = np.zeros((k, F.shape[1]))
        cnt = np.bincount(lab, minlength=k)
        np.add.at(mu, lab, F)
        live = cnt > 0
        mu[live] /= cnt[live][:, None]
        D = ((F[:, None, :] - mu[None, :, :]) ** 2).sum(2)
        D[:, ~live] = np.inf
        new = D.argmin(1)
        changed = np.any(new != lab)
        lab = new
        if not changed:
            break
    return it, lab


@pytest.mark.parametrize("kind", [0, 1])
def test_kk_lloyd_equals_explicit_lloyd_on_feature_map(oracle, kind):
    """Kernel k-means with a kernel whose feature map is finite is explicit
    Lloyd on the mapped points: linear phi(x) = x; polynomial (x.y + 1)^2
    phi(x) = (vec(x x^T), sqrt(2) x, 1). Same labels, same iteration count,
    on blob data whose FP64 top-2 margins are far above rounding."""
    rng = np.random.default_rng(11 + kind)
    n, d, k = 240, 3, 4
    cen = rng.uniform(-3, 3, (k, d))
    X = (cen[rng.integers(0, k, n)] + rng.normal(0, 0.6, (n, d))).astype(np.float32)
    X64 = X.astype(np.float64)
    if kind == 0:
        F = X64
    else:
        F = np.concatenate([(X64[:, :, None] * X64[:, None, :]).reshape(n, -1),
                            np.sqrt(2.0) * X64, np.ones((n, 1))], axis=1)
    lab0 = rng.integers(0, k, n)
    it_e, lab_e = _explicit_lloyd_to_stable(F, lab0, k, 100)
    K = oracle.kernel_matrix(X, kind)
    np.testing.assert_allclose(K, F @ F.T, rtol=1e-12, atol=1e-9)
    it_k, lab_k = oracle.kk_lloyd(K, lab0, k, 100)
    assert it_e < 100
    assert it_k == it_e
    assert np.array_equal(lab_k.astype(np.int64), lab_e)
