// kmc_oracle.cpp -- CPU oracle for the matrix-centric k-means hot path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this
// file: it is the checker that tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py compare the HIP path against.
//
// Three restatements live here, each citing the reference lines it follows:
//   1. orc_cpu_*  : the reference CPU k-means, float arithmetic in the same
//                   operation order (comparisons/CPU_Kmeans/src/kmeans.cpp).
//   2. orc_f64_*  : the same Lloyd algorithm in FP64 (the accuracy yardstick
//                   for inertia and centroids).
//   3. orc_kk_*   : kernel k-means in the Popcorn formulation of the GPU
//                   reference (src/kmeans.cu, src/kernels/distances.cu,
//                   src/kernels/kernels.cuh), computed in FP64.
//   3b. orc_gpu_lloyd: the GPU reference engine's loop rules on the linear
//                   path (labels i % k -> means, empty cluster -> +inf,
//                   float score test), over the float arithmetic of 1.
//   4. orc_plusplus(_trials): k-means++ (D^2) seeding, plain or greedy, with the engine's fixed
//                   summation order (the reference's own, src/kmeans.cu:
//                   447-769, is compiled out and broken, so the order is the
//                   engine's; the mathematics is Arthur & Vassilvitskii 2007).
//
// How each function is pinned (checks live in tests/test_oracle.py):
//   orc_cpu_init_centroids, orc_cpu_distances, orc_cpu_argmin,
//   orc_cpu_centroids, orc_cpu_converged, orc_cpu_lloyd
//       -> bit-exact against the compiled reference CPU k-means (oracle/_ref)
//          and the committed golden fixtures made from it; distances also
//          against the worked examples 16/36/20/3 of the reference's report.
//   orc_f64_*  -> closed forms (mean of one-hot groups, brute-force argmin on
//                 tiny inputs), the identity inertia = sum ||x||^2 - ..., and
//                 Lloyd monotonicity of inertia.
//   orc_kk_kernel_matrix -> numpy FP64 GEMM X X^T and the kernel closed forms.
//   orc_plusplus -> a plain cumulative-sum D^2 sampler on integer-valued data
//                 (every summation order is exact there), first-round index
//                 floor(u*n), distinct picks.
//   orc_kk_distances     -> the reference unit test's O(n^3) host formula
//                 (tests/unit_kernels.cu:365-395) restated independently in
//                 the test, and the linear-kernel identity
//                 d_K(x_i, mu_c) = ||x_i - mu_c||^2.
//   orc_kk_lloyd -> a hand-worked 1-D case (labels and iteration count), and
//                 explicit FP64 Lloyd on the kernel's finite feature map:
//                 phi(x) = x for the linear kernel, phi(x) = (vec(x x^T),
//                 sqrt(2) x, 1) for (x.y + 1)^2, run from the same labels to
//                 the same stable labels in the same number of iterations.
//   orc_gpu_lloyd -> hand-worked 1-D cases: the score sequence, the stop
//                 iteration on either side of tol, an emptied cluster that is
//                 never chosen again (the zero-vector rule would re-pick it),
//                 and the no-test-at-maxiter rule.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

extern "C" {

// ---------------------------------------------------------------------------
// 1. Reference CPU k-means, float semantics
// ---------------------------------------------------------------------------

// comparisons/CPU_Kmeans/src/kmeans.cpp:84-119 (and the seeding at :22-29):
// mt19937 seeded through seed_seq{seed}; uniform_int_distribution<int>(0,n-1)
// draws row indices; a row is accepted only if it differs by value from every
// row accepted so far. Returns 0, or -1 if fewer than k distinct rows exist.
int orc_cpu_init_centroids(const float* X, size_t n, uint32_t d, uint32_t k,
                           int seed, float* C0) {
  std::seed_seq ss{seed};
  std::mt19937 gen(ss);
  std::uniform_int_distribution<int> pick(0, (int)n - 1);
  std::vector<size_t> used;
  // Bound the number of draws so that a data set with < k distinct rows
  // cannot loop forever (the reference would).
  const uint64_t max_draws = 64ull * n + 1024ull * k;
  uint64_t draws = 0;
  while (used.size() < k) {
    if (draws++ > max_draws) return -1;
    size_t r = (size_t)pick(gen);
    bool dup = false;
    for (size_t u : used) {
      // element-wise float ==, as Point::operator== (src/include/point.hpp:66-72)
      bool eq = true;
      for (uint32_t j = 0; j < d; ++j)
        if (X[u * d + j] != X[r * d + j]) { eq = false; break; }
      if (eq) { dup = true; break; }
    }
    if (!dup) {
      std::memcpy(C0 + used.size() * d, X + r * d, sizeof(float) * d);
      used.push_back(r);
    }
  }
  return 0;
}

// comparisons/CPU_Kmeans/src/kmeans.cpp:121-131: direct form, float, the sum
// over dimensions runs in index order, one rounding per subtraction, product
// and addition (no fused multiply-add: the Makefile builds with
// -ffp-contract=off and without -ffast-math).
void orc_cpu_distances(const float* X, size_t n, uint32_t d, const float* C,
                       uint32_t k, float* D) {
  for (size_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) {
      float acc = 0.0f;
      for (uint32_t j = 0; j < d; ++j) {
        float t = X[i * d + j] - C[(size_t)c * d + j];
        float sq = t * t;
        acc = acc + sq;
      }
      D[i * k + c] = acc;
    }
}

// comparisons/CPU_Kmeans/src/kmeans.cpp:141-154: first index of the minimum
// (strict <, starting from +inf), plus the cluster histogram.
void orc_cpu_argmin(const float* D, size_t n, uint32_t k, uint32_t* labels,
                    uint32_t* counts) {
  std::memset(counts, 0, sizeof(uint32_t) * k);
  for (size_t i = 0; i < n; ++i) {
    float best = std::numeric_limits<float>::infinity();
    uint32_t idx = 0;
    for (uint32_t c = 0; c < k; ++c)
      if (D[i * k + c] < best) { best = D[i * k + c]; idx = c; }
    labels[i] = idx;
    counts[idx]++;
  }
}

// comparisons/CPU_Kmeans/src/kmeans.cpp:166-180: float sums in point order,
// then multiply by scale = (float)(1.0 / (double)max(count,1)). An empty
// cluster therefore becomes the zero vector.
void orc_cpu_centroids(const float* X, size_t n, uint32_t d,
                       const uint32_t* labels, const uint32_t* counts,
                       uint32_t k, float* C) {
  std::vector<float> acc((size_t)k * d, 0.0f);
  for (size_t i = 0; i < n; ++i) {
    float* a = acc.data() + (size_t)labels[i] * d;
    for (uint32_t j = 0; j < d; ++j) {
      a[j] = a[j] + X[i * d + j];
    }
  }
  for (uint32_t c = 0; c < k; ++c) {
    uint32_t cnt = counts[c] > 1 ? counts[c] : 1;
    float scale = (float)(1.0 / (double)cnt);
    for (uint32_t j = 0; j < d; ++j) {
      C[(size_t)c * d + j] = acc[(size_t)c * d + j] * scale;
    }
  }
}

// comparisons/CPU_Kmeans/src/kmeans.cpp:225-234: converged iff for every
// centroid sqrt(sum_j (c_j - last_j)^2) <= tol, the sum in float, in order.
int orc_cpu_converged(const float* C, const float* Clast, uint32_t k,
                      uint32_t d, float tol) {
  for (uint32_t c = 0; c < k; ++c) {
    float s = 0.0f;
    for (uint32_t j = 0; j < d; ++j) {
      float t = C[(size_t)c * d + j] - Clast[(size_t)c * d + j];
      float sq = t * t;
      s = s + sq;
    }
    if (std::sqrt((float)s) > tol) return 0;
  }
  return 1;
}

// comparisons/CPU_Kmeans/src/kmeans.cpp:192-223: the Lloyd loop. C holds the
// initial centroids on entry and the final ones on exit. Returns the
// iteration at which the shift test first passed (tested from iteration 2
// on), or maxiter. iters_run (optional) receives the number of iterations run.
uint64_t orc_cpu_lloyd(const float* X, size_t n, uint32_t d, uint32_t k,
                       float tol, uint64_t maxiter, float* C, uint32_t* labels,
                       uint32_t* counts, uint64_t* iters_run) {
  std::vector<float> D(n * (size_t)k);
  std::vector<float> last(C, C + (size_t)k * d);
  uint64_t converged = maxiter, iter = 0;
  while (iter++ < maxiter) {
    orc_cpu_distances(X, n, d, C, k, D.data());
    orc_cpu_argmin(D.data(), n, k, labels, counts);
    orc_cpu_centroids(X, n, d, labels, counts, k, C);
    if (iter > 1 && orc_cpu_converged(C, last.data(), k, d, tol)) {
      converged = iter;
      break;
    }
    std::memcpy(last.data(), C, sizeof(float) * (size_t)k * d);
  }
  if (iters_run) *iters_run = iter > maxiter ? maxiter : iter;
  return converged;
}

// ---------------------------------------------------------------------------
// 2. FP64 Lloyd (same assign -> update -> shift order)
// ---------------------------------------------------------------------------

// ||x_i||^2 in FP64.
void orc_f64_row_norms(const float* X, size_t n, uint32_t d, double* out) {
  for (size_t i = 0; i < n; ++i) {
    double s = 0;
    for (uint32_t j = 0; j < d; ++j) s += (double)X[i * d + j] * X[i * d + j];
    out[i] = s;
  }
}

// Exact-as-FP64 squared distances ||x_i - c||^2 (direct form).
void orc_f64_distances(const float* X, size_t n, uint32_t d, const float* C,
                       uint32_t k, double* D) {
  for (size_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) {
      double s = 0;
      for (uint32_t j = 0; j < d; ++j) {
        double t = (double)X[i * d + j] - (double)C[(size_t)c * d + j];
        s += t * t;
      }
      D[i * k + c] = s;
    }
}

// Inertia sum_i ||x_i - C[label_i]||^2 in FP64.
double orc_f64_inertia(const float* X, size_t n, uint32_t d, const float* C,
                       const uint32_t* labels) {
  double tot = 0;
  for (size_t i = 0; i < n; ++i) {
    const float* c = C + (size_t)labels[i] * d;
    double s = 0;
    for (uint32_t j = 0; j < d; ++j) {
      double t = (double)X[i * d + j] - (double)c[j];
      s += t * t;
    }
    tot += s;
  }
  return tot;
}

// Per-cluster means in FP64 (empty cluster -> zero vector, as the reference
// CPU k-means), rounded once to float.
void orc_f64_centroids(const float* X, size_t n, uint32_t d,
                       const uint32_t* labels, uint32_t k, float* C,
                       uint32_t* counts) {
  std::vector<double> acc((size_t)k * d, 0.0);
  std::vector<uint32_t> cnt(k, 0);
  for (size_t i = 0; i < n; ++i) {
    cnt[labels[i]]++;
    for (uint32_t j = 0; j < d; ++j)
      acc[(size_t)labels[i] * d + j] += X[i * d + j];
  }
  for (uint32_t c = 0; c < k; ++c) {
    double inv = 1.0 / (double)(cnt[c] > 1 ? cnt[c] : 1);
    for (uint32_t j = 0; j < d; ++j)
      C[(size_t)c * d + j] = (float)(acc[(size_t)c * d + j] * inv);
    if (counts) counts[c] = cnt[c];
  }
}

// One FP64 assignment: labels = first-min of exact FP64 distances; returns the
// FP64 inertia of that assignment. margin_out (optional) receives, per point,
// the gap between the second-smallest and the smallest FP64 distance.
double orc_f64_assign(const float* X, size_t n, uint32_t d, const float* C,
                      uint32_t k, uint32_t* labels, double* margin_out) {
  double tot = 0;
  for (size_t i = 0; i < n; ++i) {
    double b1 = std::numeric_limits<double>::infinity(), b2 = b1;
    uint32_t idx = 0;
    for (uint32_t c = 0; c < k; ++c) {
      double s = 0;
      for (uint32_t j = 0; j < d; ++j) {
        double t = (double)X[i * d + j] - (double)C[(size_t)c * d + j];
        s += t * t;
      }
      if (s < b1) { b2 = b1; b1 = s; idx = c; }
      else if (s < b2) { b2 = s; }
    }
    labels[i] = idx;
    if (margin_out) margin_out[i] = b2 - b1;
    tot += b1;
  }
  return tot;
}

// ---------------------------------------------------------------------------
// 3. Kernel k-means (Popcorn formulation), FP64
// ---------------------------------------------------------------------------

// Kernel functions of src/kernels/kernels.cuh:49-52 (linear), :74-75
// (sigmoid tanh(v+1)), :100-101 (polynomial (v+1)^2), before the reference's
// -2 scaling; kind 3 is the Gaussian RBF exp(-gamma*||x-y||^2), which the
// reference does not have. K is n x n row-major, FP64.
void orc_kk_kernel_matrix(const float* X, size_t n, uint32_t d, int kind,
                          double gamma, double* K) {
  std::vector<double> nrm(n);
  orc_f64_row_norms(X, n, d, nrm.data());
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j <= i; ++j) {
      double v = 0;
      for (uint32_t t = 0; t < d; ++t) v += (double)X[i * d + t] * X[j * d + t];
      double f;
      switch (kind) {
        case 0: f = v; break;
        case 1: f = (v + 1.0) * (v + 1.0); break;
        case 2: f = std::tanh(v + 1.0); break;
        default: {
          double sq = nrm[i] + nrm[j] - 2.0 * v;
          if (i == j) sq = 0.0;
          f = std::exp(-gamma * (sq < 0 ? 0 : sq));
        }
      }
      K[i * n + j] = f;
      K[j * n + i] = f;
    }
}

// Kernel-space distances for the labels given (src/kernels/distances.cu:895-1002
// with B = -2K): D[i,c] = ||mu_c||_K^2 - 2 <phi(x_i), mu_c>, where
// <phi(x_i), mu_c> = sum_{j in c} K[i,j] / |c| and
// ||mu_c||_K^2 = sum_{i,j in c} K[i,j] / |c|^2. An empty cluster (and, as in
// the reference's add_norm_mtx, a cluster whose norm is exactly 0) gets +inf.
// With add_diag != 0 the K[i,i] term is added, giving the full distance
// ||phi(x_i) - mu_c||^2 (the reference omits it; its argmin is unchanged).
void orc_kk_distances(const double* K, size_t n, const uint32_t* labels,
                      uint32_t k, int add_diag, double* D) {
  std::vector<double> cnt(k, 0.0), cn(k, 0.0);
  for (size_t i = 0; i < n; ++i) cnt[labels[i]] += 1.0;
  std::vector<double> S(n * (size_t)k, 0.0);  // S[i,c] = sum_{j in c} K[i,j]
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) S[i * k + labels[j]] += K[i * n + j];
  for (size_t i = 0; i < n; ++i) cn[labels[i]] += S[i * k + labels[i]];
  for (uint32_t c = 0; c < k; ++c)
    cn[c] = cnt[c] > 0 ? cn[c] / (cnt[c] * cnt[c]) : 0.0;
  const double inf = std::numeric_limits<double>::infinity();
  for (size_t i = 0; i < n; ++i)
    for (uint32_t c = 0; c < k; ++c) {
      if (cnt[c] == 0 || cn[c] == 0.0) { D[i * k + c] = inf; continue; }
      double v = cn[c] - 2.0 * S[i * k + c] / cnt[c];
      if (add_diag) v += K[i * n + i];
      D[i * k + c] = v;
    }
}

// Kernel k-means Lloyd in FP64 from initial labels: each iteration assigns
// every point to the first minimum of orc_kk_distances and stops early when
// no label changes. Returns the number of iterations run; labels in/out.
uint64_t orc_kk_lloyd(const double* K, size_t n, uint32_t k, uint64_t maxiter,
                      uint32_t* labels) {
  std::vector<double> D(n * (size_t)k);
  uint64_t it = 0;
  while (it < maxiter) {
    ++it;
    orc_kk_distances(K, n, labels, k, 1, D.data());
    bool changed = false;
    for (size_t i = 0; i < n; ++i) {
      double b = std::numeric_limits<double>::infinity();
      uint32_t idx = labels[i];
      for (uint32_t c = 0; c < k; ++c)
        if (D[i * k + c] < b) { b = D[i * k + c]; idx = c; }
      if (idx != labels[i]) changed = true;
      labels[i] = idx;
    }
    if (!changed) break;
  }
  return it;
}

// ---------------------------------------------------------------------------
// 3b. The reference GPU engine's Lloyd loop rules on the linear path
// ---------------------------------------------------------------------------

// The loop of Kmeans::run (src/kmeans.cu:772-1122) with the linear kernel,
// from initial labels (init_centroids_rand gives labels i % k,
// src/kmeans.cu:365-444). The engine keeps no explicit centroids: its
// distance D[i,c] = ||mu_c||^2 - 2 x_i . mu_c of the label means mu_c comes
// from the kernel matrix (src/kernels/distances.cu:895-1002) and ranks the
// centroids exactly as ||x_i - mu_c||^2 does. This restatement uses the
// reference CPU k-means' float arithmetic for that ranking (orc_cpu_distances:
// direct form, first minimum) and its float means (orc_cpu_centroids), so it
// shares labels with the rest of this oracle wherever the top-2 margin is
// above float rounding. Rules restated:
//   * a cluster empty under the current labels has c_tilde = 0 and distance
//     +inf (src/kernels/distances.cu:588-598): it is never chosen again;
//   * score_t = sum_i min_c D[i,c] = FP64 inertia of the assignment minus
//     sum_i ||x_i||^2 (the kernel form omits ||x||^2), held in float
//     (DATA_TYPE score, src/kmeans.cu:1053-1061);
//   * after iteration t: stop (without a test) when t == maxiter; with
//     check, stop when t > 1 and |score_t - score_{t-1}| < tol (float) and
//     return t (src/kmeans.cu:1064-1073); otherwise return maxiter.
// labels: in = initial labels, out = labels of the last assignment.
// C (k x d, out): means of those labels (empty cluster: zero vector).
// scores (optional, maxiter entries): float score of each iteration run.
// iters_run (optional): iterations run.
uint64_t orc_gpu_lloyd(const float* X, size_t n, uint32_t d, uint32_t k, float tol,
                       uint64_t maxiter, int check, uint32_t* labels, float* C,
                       float* scores, uint64_t* iters_run) {
  std::vector<uint32_t> cnt(k, 0);
  for (size_t i = 0; i < n; ++i) cnt[labels[i]]++;
  std::vector<char> alive(k);
  for (uint32_t c = 0; c < k; ++c) alive[c] = cnt[c] > 0;
  orc_cpu_centroids(X, n, d, labels, cnt.data(), k, C);
  std::vector<double> xn(n);
  orc_f64_row_norms(X, n, d, xn.data());
  double sum_x2 = 0.0;
  for (size_t i = 0; i < n; ++i) sum_x2 += xn[i];
  std::vector<float> D(n * (size_t)k);
  uint64_t converged = maxiter, iter = 0;
  float last_score = 0.0f;
  while (iter++ < maxiter) {
    orc_cpu_distances(X, n, d, C, k, D.data());
    std::fill(cnt.begin(), cnt.end(), 0u);
    for (size_t i = 0; i < n; ++i) {
      float best = std::numeric_limits<float>::infinity();
      uint32_t idx = 0;
      for (uint32_t c = 0; c < k; ++c)
        if (alive[c] && D[i * k + c] < best) { best = D[i * k + c]; idx = c; }
      labels[i] = idx;
      cnt[idx]++;
    }
    const double inert = orc_f64_inertia(X, n, d, C, labels);
    const float score = (float)(inert - sum_x2);
    if (scores) scores[iter - 1] = score;
    for (uint32_t c = 0; c < k; ++c)
      if (cnt[c] == 0) alive[c] = 0;
    orc_cpu_centroids(X, n, d, labels, cnt.data(), k, C);
    if (iter == maxiter) break;
    if (check && iter > 1 && std::fabs(score - last_score) < tol) {
      converged = iter;
      break;
    }
    last_score = score;
  }
  if (iters_run) *iters_run = iter > maxiter ? maxiter : iter;
  return converged;
}


// k-means++ seeding: K rounds of D^2 sampling with uniforms u[0..K).
// Per point, the FP64 squared distance to the newest centroid is summed over
// the dimensions in order (t = x - c in double, s = fma(t, t, s)) and kept as
// the minimum over the chosen set. Blocks of 1024 points: thread t of 256
// adds points t, t+256, t+512, t+768 in that order, then a pairwise tree
// (o = 128, 64, ..., 1: s[t] += s[t + o]); the total is the block sums added
// in block order. Round r > 0 picks the first block whose running prefix
// exceeds u_r * total (else the last block), then the first point of it with
// D^2 > 0 at which the running sum exceeds the target (else the block's last
// point with D^2 > 0); a zero total falls back to floor(u_r * n).
// Greedy k-means++ (kmc_init_plusplus_trials; src/kmeans.cu:463-769): round r
// >= 1 draws `trials` candidates by D^2 sampling from uniforms
// u[1 + (r-1)*trials + t], each candidate's potential sum_i min(D2_i,
// ||x_i - cand||^2) is summed in the same blocked order as D2, and the first
// smallest becomes centroid r. trials = 1 is plain k-means++ (u[r]).
void orc_plusplus_trials(const float* X, size_t n, uint32_t d, uint32_t k, uint32_t trials,
                         const double* u, int64_t* chosen, float* C) {
  const size_t nb = (n + 1023) / 1024;
  std::vector<double> D2(n, 0.0), bsum(nb, 0.0);
  auto pick_row = [&](uint32_t r, int64_t idx) {
    chosen[r] = idx;
    for (uint32_t j = 0; j < d; ++j) C[(size_t)r * d + j] = X[(size_t)idx * d + j];
  };
  auto dist = [&](size_t i, const float* c) {
    double acc = 0.0;
    for (uint32_t j = 0; j < d; ++j) {
      const double tt = (double)X[i * d + j] - (double)c[j];
      acc = std::fma(tt, tt, acc);
    }
    return acc;
  };
  // block b's sum of f(i) over its 1024 points: thread t adds points t,
  // t+256, t+512, t+768 in order, then a pairwise tree over the threads
  auto block_sum = [&](size_t b, auto&& f) {
    double part[256];
    for (int t = 0; t < 256; ++t) {
      double ps = 0.0;
      for (int q = 0; q < 4; ++q) {
        const size_t i = b * 1024 + (size_t)t + 256 * (size_t)q;
        if (i < n) ps += f(i);
      }
      part[t] = ps;
    }
    for (int o = 128; o > 0; o >>= 1)
      for (int t = 0; t < o; ++t) part[t] += part[t + o];
    return part[0];
  };
  auto sample = [&](double uu) {
    int64_t idx = (int64_t)(uu * (double)n);
    if (idx > (int64_t)n - 1) idx = (int64_t)n - 1;
    double total = 0.0;
    for (size_t b = 0; b < nb; ++b) total += bsum[b];
    if (total > 0.0) {
      const double target = uu * total;
      double run = 0.0;
      size_t b = 0;
      for (; b + 1 < nb; ++b) {
        if (run + bsum[b] > target) break;
        run += bsum[b];
      }
      const size_t lo = b * 1024, hi = std::min(n, lo + 1024);
      int64_t pick = -1, last_pos = -1;
      for (size_t i = lo; i < hi; ++i) {
        const double dv = D2[i];
        if (dv > 0.0) last_pos = (int64_t)i;
        run += dv;
        if (run > target && dv > 0.0) {
          pick = (int64_t)i;
          break;
        }
      }
      if (pick < 0) pick = last_pos;
      if (pick >= 0) idx = pick;
    }
    return idx;
  };
  int64_t first = (int64_t)(u[0] * (double)n);
  if (first > (int64_t)n - 1) first = (int64_t)n - 1;
  pick_row(0, first);
  const uint32_t s = trials < 1 ? 1 : trials;
  for (uint32_t r = 1; r < k; ++r) {
    const float* c = C + (size_t)(r - 1) * d;
    for (size_t b = 0; b < nb; ++b)
      bsum[b] = block_sum(b, [&](size_t i) {
        const double acc = dist(i, c);
        const double dv = (r == 1) ? acc : std::fmin(D2[i], acc);
        D2[i] = dv;
        return dv;
      });
    if (s == 1) {
      pick_row(r, sample(u[r]));
      continue;
    }
    int64_t best_idx = -1;
    double best = 0.0;
    for (uint32_t t = 0; t < s; ++t) {
      const int64_t cand = sample(u[1 + (size_t)(r - 1) * s + t]);
      const float* cr = X + (size_t)cand * d;
      double pot = 0.0;
      for (size_t b = 0; b < nb; ++b)
        pot += block_sum(b, [&](size_t i) { return std::fmin(D2[i], dist(i, cr)); });
      if (best_idx < 0 || pot < best) {
        best = pot;
        best_idx = cand;
      }
    }
    pick_row(r, best_idx);
  }
}

void orc_plusplus(const float* X, size_t n, uint32_t d, uint32_t k, const double* u,
                  int64_t* chosen, float* C) {
  orc_plusplus_trials(X, n, d, k, 1, u, chosen, C);
}

}  // extern "C"
