// kmeans.cpp -- the drop-in Kmeans class (include/kmeans.hpp) over the C ABI.
#include "kmeans.hpp"

#include <cstdio>
#include <random>
#include <string>
#include <vector>

namespace kmc {

bool query_device(int ordinal, DeviceInfo* info) {
  const int count = kmc_device_count();
  if (info) {
    info->ordinal = ordinal;
    info->count = count;
  }
  return ordinal >= 0 && ordinal < count;
}

}  // namespace kmc

namespace {

[[noreturn]] void raise(kmc_handle* h, kmc_status s, const char* what) {
  std::string msg = std::string(what) + ": " + kmc_status_string(s);
  if (h) {
    const char* e = kmc_last_error(h);
    if (e && *e) msg += std::string(" (") + e + ")";
  }
  throw kmc::Error(s, msg);
}

void check(kmc_handle* h, kmc_status s, const char* what) {
  if (s != KMC_OK) raise(h, s, what);
}

void kk_check(const kmc_kk_handle* h, kmc_status s, const char* what) {
  if (s == KMC_OK) return;
  std::string msg = std::string(what) + ": " + kmc_status_string(s);
  if (h) {
    const char* e = kmc_kk_last_error(h);
    if (e && *e) msg += std::string(" (") + e + ")";
  }
  throw kmc::Error(s, msg);
}

}  // namespace

Kmeans::Kmeans(const size_t _n, const uint32_t _d, const uint32_t _k, const float _tol,
               const int* seed, Point<DATA_TYPE>** _points, kmc::DeviceInfo* _deviceProps,
               InitMethod _initMethod, Kernel _kernel, int _level)
    : Kmeans(_n, _d, _k, _tol, seed, _points, _deviceProps, _initMethod, _kernel, _level,
             Options()) {}

Kmeans::Kmeans(const size_t _n, const uint32_t _d, const uint32_t _k, const float _tol,
               const int* seed, Point<DATA_TYPE>** _points, kmc::DeviceInfo* _deviceProps,
               InitMethod _initMethod, Kernel _kernel, int _level, const Options& _opts)
    : n(_n), d(_d), k(_k), tol(_tol), points(_points), deviceProps(_deviceProps),
      initMethod(_initMethod), kernel(_kernel), level(_level), opts(_opts) {
  init(seed);
  if (opts.report_stage_times) print_setup_times();
}

namespace {
const char* kCyan = "\033[36m";
const char* kReset = "\033[0m";
void perf_line(const char* stage, double ms) {
  std::printf("%s[PERFORMANCE]%s %s time: %.8f\n", kCyan, kReset, stage, ms / 1000.0);
}
}  // namespace

void Kmeans::print_setup_times() const {
  kmc_setup_times t{};
  const kmc_status s = kh ? kmc_kk_get_setup_times(kh, &t) : kmc_get_setup_times(h, &t);
  if (s != KMC_OK) return;
  perf_line("memcpy", t.memcpy_ms);
  perf_line("b-mult", t.operands_ms);
  perf_line("init_centroids", t.init_ms);
}

void Kmeans::print_stage_times() {
  uint64_t cnt = 0;
  kmc_status s = kh ? kmc_kk_get_stage_times(kh, nullptr, 0, &cnt)
                    : kmc_get_stage_times(h, nullptr, 0, &cnt);
  if (s != KMC_OK || cnt == 0) return;
  std::vector<kmc_stage_times> t(cnt);
  s = kh ? kmc_kk_get_stage_times(kh, t.data(), cnt, &cnt)
         : kmc_get_stage_times(h, t.data(), cnt, &cnt);
  if (s != KMC_OK) return;
  for (const kmc_stage_times& it : t) {
    perf_line("compute_distances", it.distances_ms);
    perf_line("clusters_argmin_shfl", it.argmin_ms);
    perf_line("compute_centroids", it.centroids_ms);
  }
}

void Kmeans::init(const int* seed) {
  if (!points || n == 0 || d == 0 || k == 0 || k > n)
    throw kmc::Error(KMC_ERR_ARG, "Kmeans: invalid n/d/k or points");
  if (level < 0 || level > 4) throw kmc::Error(KMC_ERR_ARG, "Kmeans: level must be 0..4");
  if (kernel != Kernel::linear) {
    // kernel k-means: K = f(X X^T) built on the device, labels i % k
    // (src/kmeans.cu:143-171, 372-374); the bf16 forms store K in bf16
    kmc_kk_config kc{};
    kc.n = n;
    kc.d = d;
    kc.k = k;
    kc.tol = tol;
    kc.kernel = kernel == Kernel::polynomial ? KMC_KERNEL_POLYNOMIAL
                : kernel == Kernel::sigmoid  ? KMC_KERNEL_SIGMOID
                                             : KMC_KERNEL_RBF;
    kc.store = opts.precision == Precision::fp32 ? KMC_STORE_FP32 : KMC_STORE_BF16;
    kc.gamma = 0.0f;
    kc.device = deviceProps ? deviceProps->ordinal : 0;
    kc.product = KMC_KK_AUTO;
    kc.build = opts.precision == Precision::fp32     ? KMC_KBUILD_FP32
               : opts.precision == Precision::bf16x3 ? KMC_KBUILD_BF16X3
                                                     : KMC_KBUILD_BF16;
    const kmc_status s = kmc_kk_create(&kc, &kh);
    if (s != KMC_OK) throw kmc::Error(s, std::string("kmc_kk_create: ") + kmc_status_string(s));
    std::vector<float> X(n * (size_t)d);
    for (size_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j < d; ++j) X[i * d + j] = points[i]->get(j);
    kk_check(kh, kmc_kk_set_points(kh, X.data(), d), "kmc_kk_set_points");
    std::vector<int32_t> lab(n);
    for (size_t i = 0; i < n; ++i) lab[i] = (int32_t)(i % k);
    kk_check(kh, kmc_kk_set_labels(kh, lab.data()), "kmc_kk_set_labels");
    return;
  }

  kmc_config cfg{};
  cfg.n = n;
  cfg.d = d;
  cfg.k = k;
  cfg.tol = tol;
  cfg.precision = opts.precision == Precision::fp32     ? KMC_PREC_FP32
                  : opts.precision == Precision::bf16x3 ? KMC_PREC_BF16X3
                  : opts.precision == Precision::fp8    ? KMC_PREC_FP8
                                                        : KMC_PREC_BF16;
  cfg.update = KMC_UPDATE_ORDERED;
  cfg.device = deviceProps ? deviceProps->ordinal : 0;
  check(nullptr, kmc_create(&cfg, &h), "kmc_create");

  // Pack the borrowed points (src/kmeans.cu:94-99).
  std::vector<float> X(n * (size_t)d);
  for (size_t i = 0; i < n; ++i)
    for (uint32_t j = 0; j < d; ++j) X[i * d + j] = points[i]->get(j);
  check(h, kmc_set_points(h, X.data(), d), "kmc_set_points");

  if (initMethod == InitMethod::plus_plus) {
    // k-means++ (D^2) seeding on the device; the reference's
    // init_centroids_plus_plus (src/kmeans.cu:447-769) is compiled out
    uint64_t s = 0;
    if (seed) {
      s = (uint64_t)(uint32_t)*seed;
    } else {
      std::random_device rd;
      s = ((uint64_t)rd() << 32) ^ rd();
    }
    // greedy, 2 + ceil(ln k) trials per round: the scheme the reference's
    // version implements (src/kmeans.cu:463-769)
    check(h, kmc_init_plusplus_trials(h, s, 0, nullptr), "kmc_init_plusplus_trials");
  } else if (opts.reference_gpu) {
    // init_centroids_rand (src/kmeans.cu:365-444): labels i % k, the
    // initial centroids are the label means.
    std::vector<int32_t> lab(n);
    for (size_t i = 0; i < n; ++i) lab[i] = (int32_t)(i % k);
    check(h, kmc_set_labels(h, lab.data()), "kmc_set_labels");
  } else {
    // k distinct rows drawn as the reference CPU k-means does
    const std::vector<float> C0 = kmc::seeded_rows(points, X, n, d, k, seed);
    check(h, kmc_set_centroids(h, C0.data()), "kmc_set_centroids");
  }
}

std::vector<float> kmc::seeded_rows(Point<DATA_TYPE>** points, const std::vector<float>& X,
                                    size_t n, uint32_t d, uint32_t k, const int* seed) {
  // mt19937(seed_seq{seed}) and distinct rows, as
  // comparisons/CPU_Kmeans/src/kmeans.cpp:57-64,84-119
  std::mt19937 gen;
  if (seed) {
    std::seed_seq ss{*seed};
    gen.seed(ss);
  } else {
    std::random_device rd;
    gen.seed(rd());
  }
  std::uniform_int_distribution<int> pick(0, (int)n - 1);
  std::vector<size_t> used;
  std::vector<float> C0((size_t)k * d);
  uint64_t draws = 0;
  const uint64_t max_draws = 64ull * n + 1024ull * k;
  while (used.size() < k) {
    if (draws++ > max_draws) throw kmc::Error(KMC_ERR_ARG, "Kmeans: fewer than k distinct points");
    const size_t r = (size_t)pick(gen);
    bool dup = false;
    for (size_t u : used)
      if (*points[u] == *points[r]) {
        dup = true;
        break;
      }
    if (dup) continue;
    for (uint32_t j = 0; j < d; ++j) C0[used.size() * d + j] = X[r * d + j];
    used.push_back(r);
  }
  return C0;
}

Kmeans::~Kmeans() {
  if (h) kmc_destroy(h);
  if (kh) kmc_kk_destroy(kh);
}

uint64_t Kmeans::run(uint64_t maxiter, bool check_converged) {
  if (kh) {
    kmc_kk_result res{};
    uint32_t kflags = check_converged ? KMC_RUN_CHECK_CONVERGED : 0u;
    if (opts.report_stage_times) kflags |= KMC_RUN_TIME_KERNELS;
    kk_check(kh, kmc_kk_run(kh, maxiter, kflags, &res), "kmc_kk_run");
    if (opts.report_stage_times) print_stage_times();
    score = (float)res.score;
    inertia = res.objective;
    iterations = res.iterations;
    std::vector<int32_t> lab(n);
    kk_check(kh, kmc_kk_get_labels(kh, lab.data()), "kmc_kk_get_labels");
    for (size_t i = 0; i < n; ++i) points[i]->setCluster((unsigned int)lab[i]);
    return res.converged_at;
  }
  uint32_t flags = check_converged ? KMC_RUN_CHECK_CONVERGED : 0u;
  if (opts.reference_gpu) flags |= KMC_RUN_CONVERGE_SCORE | KMC_RUN_EXCLUDE_EMPTY;
  if (opts.report_stage_times) flags |= KMC_RUN_TIME_KERNELS;
  kmc_result res{};
  check(h, kmc_run(h, maxiter, flags, &res), "kmc_run");
  if (opts.report_stage_times) print_stage_times();
  score = (float)res.score;
  inertia = res.inertia;
  iterations = res.iterations;
  std::vector<int32_t> lab(n);
  check(h, kmc_get_labels(h, lab.data()), "kmc_get_labels");
  for (size_t i = 0; i < n; ++i) points[i]->setCluster((unsigned int)lab[i]);
  return res.converged_at;
}

std::vector<float> Kmeans::get_centroids() const {
  if (kh)
    throw kmc::Error(KMC_ERR_UNSUPPORTED,
                     "Kmeans: kernel k-means centroids live in the kernel's feature space");
  std::vector<float> C((size_t)k * d);
  check(h, kmc_get_centroids(h, C.data()), "kmc_get_centroids");
  return C;
}
