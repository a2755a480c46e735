// kmeans.hpp -- drop-in replacement for the reference engine class Kmeans
// (src/kmeans.cuh:48-210), running on the gfx950 engine through the C ABI of
// kmc.h. Host C++ only: no HIP or CUDA headers.
//
// Kept from the reference:
//   Kmeans(n, d, k, tol, seed, points, deviceProps, InitMethod, Kernel, level)
//   uint64_t run(uint64_t maxiter, bool check_converged)  -> convergence
//            iteration, or maxiter; writes each point's label through
//            Point::setCluster (src/kmeans.cu:1101-1103)
//   float get_score() const                          (src/kmeans.cuh:138)
//   enums InitMethod{random, plus_plus}, Kernel{linear, polynomial, sigmoid}
//   a non-linear kernel runs kernel k-means on the n x n kernel matrix (the
//   reference's Popcorn path, src/kmeans.cu:143-171,772-1122): initial labels
//   i % k, score test with check_converged; the linear kernel runs Lloyd
// Changed:
//   the device-properties argument is a DeviceInfo* (kmc_query_device), not a
//   cudaDeviceProp*; errors throw kmc::Error instead of exiting the process
//   (src/cuda_utils.cu:10-41 calls exit()).
// Added (defaulted, so reference call sites compile unchanged):
//   Kernel::rbf, a Precision argument and Options.
#ifndef KMC_KMEANS_HPP
#define KMC_KMEANS_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kmc.h"
#include "point.hpp"

#ifndef DATA_TYPE
#define DATA_TYPE float
#endif

namespace kmc {

class Error : public std::runtime_error {
 public:
  Error(kmc_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  kmc_status status;
};

// Replaces the cudaDeviceProp the reference passes around (src/main.cu:54-55).
struct DeviceInfo {
  int ordinal = 0;
  int count = 0;
};

// Fills *info for device `ordinal`; returns false when no device exists.
bool query_device(int ordinal, DeviceInfo* info);

// The reference CPU k-means' initial centroids: k distinct rows of the
// points drawn with mt19937(seed_seq{*seed}) (random_device when seed is
// null), comparisons/CPU_Kmeans/src/kmeans.cpp:57-64,84-119. X is the n x d
// row-major copy of the points; returns k x d floats.
std::vector<float> seeded_rows(Point<DATA_TYPE>** points, const std::vector<float>& X, size_t n,
                               uint32_t d, uint32_t k, const int* seed);

}  // namespace kmc

class Kmeans {
 public:
  enum class InitMethod { random, plus_plus };
  enum class Kernel { linear, polynomial, sigmoid, rbf };
  enum class Precision { fp32, bf16x3, bf16, fp8 };

  struct Options {
    // The reference GPU engine's semantics (src/kmeans.cu:772-1122): initial
    // labels i % k, convergence when |score_t - score_{t-1}| < tol, a cluster
    // left empty is never chosen again. With reference_gpu = false the engine
    // follows the reference CPU k-means instead (k distinct seed-chosen rows,
    // per-centroid shift test, empty cluster -> zero vector).
    bool reference_gpu = true;
    Precision precision = Precision::fp32;
    // Print the reference's "[PERFORMANCE] <stage> time: <s>" lines: memcpy,
    // b-mult and init_centroids from the constructor, compute_distances,
    // clusters_argmin_shfl and compute_centroids per iteration from run()
    // (src/kmeans.cu:136,205,297,904,958,1044, compiled in there by the
    // PERFORMANCES_* macros), timed with HIP events on the device.
    bool report_stage_times = false;
  };

  Kmeans(const size_t n, const uint32_t d, const uint32_t k, const float tol, const int* seed,
         Point<DATA_TYPE>** points, kmc::DeviceInfo* deviceProps,
         InitMethod _initMethod = InitMethod::random, Kernel _kernel = Kernel::linear,
         int _level = 3);
  Kmeans(const size_t n, const uint32_t d, const uint32_t k, const float tol, const int* seed,
         Point<DATA_TYPE>** points, kmc::DeviceInfo* deviceProps, InitMethod _initMethod,
         Kernel _kernel, int _level, const Options& opts);
  ~Kmeans();
  Kmeans(const Kmeans&) = delete;
  Kmeans& operator=(const Kmeans&) = delete;

  uint64_t run(uint64_t maxiter, bool check_converged);

  inline float get_score() const { return score; }

  // Extras, not in the reference class.
  double get_inertia() const { return inertia; }
  std::vector<float> get_centroids() const;
  uint64_t iterations_run() const { return iterations; }

 private:
  void init(const int* seed);
  void print_setup_times() const;
  void print_stage_times();
  const size_t n;
  const uint32_t d, k;
  const float tol;
  Point<DATA_TYPE>** points;
  kmc::DeviceInfo* deviceProps;
  InitMethod initMethod;
  Kernel kernel;
  int level;
  Options opts;
  kmc_handle* h = nullptr;
  kmc_kk_handle* kh = nullptr;  // kernel k-means (non-linear kernels)
  float score = 0.0f;
  double inertia = 0.0;
  uint64_t iterations = 0;
};

#endif  // KMC_KMEANS_HPP
