// raftbench.cpp -- the explicit-Lloyd benchmark program of the reference's
// harness (experiments/raft-bench/src/kmeans.cu:470-491, run by the
// "raft-kmeans" trial of experiments/driver.py:467-515) on the gfx950 engine.
//
//   raftbench n d k n_iters check tol init [infile] [seed]
//
// Same arguments, data sources and report lines as the reference program:
// points from a CSV (header line skipped, first column skipped as the class
// label), a libsvm file ("idx:value", 1-based), or, with no file (or "-1"),
// uniform values in [-1e5, 1e5); 10 trials, the first a warm-up; per trial
// the initial centroids (random: k distinct rows; plusplus: the engine's
// greedy k-means++), one explicit Lloyd run with the reference program's rules --
// cost-ratio stop (cost_t / cost_{t-1} > 1 - tol from iteration 2), with
// check = 1 also sum_c ||c_new - c_old||^2 < tol, empty clusters keep their
// centroid -- and the final clustering cost. Printed at the end, as the
// reference does: kmeans-time (mean over trials 1..9), init-time,
// kmeans-score (mean over all trials) and kmeans-iterations (the
// reference's post-decremented iteration counter, kmeans.cu:215, averaged).
//
// Differences: labels are the exact first-min of the direct squared distance
// (the engine's margin band), not RAFT's expanded-form float argmin; the
// cost is FP64; the random data and the random initial rows come from
// mt19937 (seed 1234 for the data, the optional tenth argument or
// std::random_device for the initialisation), not RAFT's generator.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "kmc.h"

namespace {

void die(const char* what, kmc_status s, const kmc_handle* h) {
  std::cerr << what << ": " << kmc_status_string(s);
  if (h) std::cerr << " (" << kmc_last_error(h) << ")";
  std::cerr << std::endl;
  std::exit(1);
}

// experiments/raft-bench/src/kmeans.cu:272-298
void read_svm(std::istream& in, uint32_t n, uint32_t d, std::vector<float>& X) {
  std::string line;
  uint32_t i = 0;
  while (i < n && std::getline(in, line)) {
    std::istringstream ls(line);
    std::string tok;
    while (std::getline(ls, tok, ' ')) {
      const size_t c = tok.find(':');
      if (c == std::string::npos) continue;
      const long idx = std::atol(tok.substr(0, c).c_str()) - 1;
      if (idx >= 0 && idx < (long)d) X[(size_t)i * d + idx] = (float)std::atof(tok.substr(c + 1).c_str());
    }
    ++i;
  }
}

// experiments/raft-bench/src/kmeans.cu:302-344
void read_csv(std::istream& in, uint32_t n, uint32_t d, std::vector<float>& X) {
  std::string line;
  std::getline(in, line);  // header
  uint32_t i = 0;
  while (i < n && std::getline(in, line)) {
    std::istringstream ls(line);
    std::string tok;
    uint32_t j = 0;
    while (std::getline(ls, tok, ',')) {
      if (j > 0 && j <= d) X[(size_t)i * d + (j - 1)] = (float)std::atof(tok.c_str());
      ++j;
    }
    ++i;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 8) {
    std::cerr << "usage: raftbench n d k n_iters check tol init [infile] [seed]" << std::endl;
    return 1;
  }
  const uint32_t n = (uint32_t)std::atol(argv[1]);
  const uint32_t d = (uint32_t)std::atol(argv[2]);
  const uint32_t k = (uint32_t)std::atol(argv[3]);
  const uint64_t n_iters = (uint64_t)std::atoll(argv[4]);
  const bool check = std::atoi(argv[5]) != 0;
  const float tol = (float)std::atof(argv[6]);
  const std::string init = argv[7];
  const std::string infile = argc > 8 ? argv[8] : "-1";
  const bool seeded = argc > 9;
  const unsigned long long seed0 = seeded ? std::strtoull(argv[9], nullptr, 10) : 0ull;
  if (n == 0 || d == 0 || k == 0 || k > n) {
    std::cerr << "invalid n, d or k" << std::endl;
    return 1;
  }
  if (init != "random" && init != "plusplus") {
    std::cerr << "Invalid init method: " << init << std::endl;
    return 1;
  }

  std::vector<float> X((size_t)n * d, 0.0f);
  if (infile == "-1") {
    std::mt19937 eng(1234u);
    std::uniform_real_distribution<float> u(-1e5f, 1e5f);
    for (float& v : X) v = u(eng);
  } else {
    std::ifstream in(infile);
    if (!in) {
      std::cerr << "Invalid input file" << std::endl;
      return 1;
    }
    if (infile.find("svm") != std::string::npos) read_svm(in, n, d, X);
    else if (infile.find("csv") != std::string::npos) read_csv(in, n, d, X);
  }

  kmc_config cfg{};
  cfg.n = n;
  cfg.d = d;
  cfg.k = k;
  cfg.tol = tol;
  cfg.precision = KMC_PREC_FP32;
  cfg.update = KMC_UPDATE_TREE;
  cfg.device = 0;
  kmc_handle* h = nullptr;
  kmc_status st = kmc_create(&cfg, &h);
  if (st != KMC_OK) die("kmc_create", st, nullptr);
  st = kmc_set_points(h, X.data(), d);
  if (st != KMC_OK) die("kmc_set_points", st, h);

  std::mt19937 gen(seeded ? (unsigned)seed0 : std::random_device{}());
  const int n_trials = 10;  // experiments/raft-bench/src/kmeans.cu:23
  std::cout << "Running kmeans" << std::endl;
  std::cout << "n:" << n << " d:" << d << " k:" << k << std::endl;
  const uint32_t flags = KMC_RUN_COST_RATIO | KMC_RUN_KEEP_EMPTY |
                         (check ? (KMC_RUN_CHECK_CONVERGED | KMC_RUN_CONVERGE_SHIFT_SUM) : 0u);
  std::vector<float> C0((size_t)k * d), C((size_t)k * d), Cn((size_t)k * d);
  std::vector<int32_t> lab(n);
  std::vector<uint32_t> cnt(k);
  std::vector<uint32_t> idx(n);
  double score = 0.0, kmeans_time = 0.0, init_time = 0.0;
  uint64_t iters = 0;
  for (int t = 0; t < n_trials; ++t) {
    const unsigned trial_seed = gen();
    auto s0 = std::chrono::system_clock::now();
    if (init == "random") {
      // k distinct rows drawn uniformly (partial Fisher-Yates)
      std::mt19937 r(trial_seed);
      std::iota(idx.begin(), idx.end(), 0u);
      for (uint32_t c = 0; c < k; ++c) {
        std::uniform_int_distribution<uint32_t> pick(c, n - 1);
        std::swap(idx[c], idx[pick(r)]);
        std::copy(X.begin() + (size_t)idx[c] * d, X.begin() + (size_t)(idx[c] + 1) * d,
                  C0.begin() + (size_t)c * d);
      }
      st = kmc_set_centroids(h, C0.data());
      if (st != KMC_OK) die("kmc_set_centroids", st, h);
    } else {
      // greedy k-means++ with 2 + ceil(ln k) trials, as RAFT's init_plus_plus
      st = kmc_init_plusplus_trials(h, trial_seed, 0, nullptr);
      if (st != KMC_OK) die("kmc_init_plusplus_trials", st, h);
    }
    auto s1 = std::chrono::system_clock::now();
    if (t > 0) init_time += std::chrono::duration<double>(s1 - s0).count();

    kmc_result res{};
    s0 = std::chrono::system_clock::now();
    st = kmc_run(h, n_iters, flags, &res);
    if (st != KMC_OK) die("kmc_run", st, h);
    // the final clustering cost: one more assignment against the final
    // centroids (kmeans.cu:222-250)
    st = kmc_get_centroids(h, C.data());
    if (st == KMC_OK) st = kmc_assign(h, C.data(), lab.data(), nullptr);
    kmc_result fin{};
    if (st == KMC_OK) st = kmc_update(h, lab.data(), C.data(), Cn.data(), cnt.data(), &fin);
    if (st != KMC_OK) die("final cost", st, h);
    s1 = std::chrono::system_clock::now();
    if (t > 0) kmeans_time += std::chrono::duration<double>(s1 - s0).count();
    // the reference's counter: stop at iteration t leaves t - 1; a run to
    // the end leaves n_iters (kmeans.cu:96, 215)
    const uint64_t n_iter = res.converged_at < n_iters ? res.converged_at - 1 : n_iters;
    score += fin.inertia;
    iters += n_iter;
    std::cout << fin.inertia << std::endl;
  }
  kmeans_time /= (n_trials - 1);
  init_time /= (n_trials - 1);
  score /= n_trials;
  iters /= n_trials;
  std::cout << std::fixed << "kmeans-time: " << kmeans_time << "s" << std::endl;
  std::cout << std::fixed << "init-time: " << init_time << "s" << std::endl;
  std::cout << std::fixed << "kmeans-score: " << score << std::endl;
  std::cout << std::fixed << "kmeans-iterations: " << iters << std::endl;
  kmc_destroy(h);
  return 0;
}
