// ref_driver.cpp -- C entry points around the reference CPU k-means class.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
// the reference's comparisons/CPU_Kmeans/src/kmeans.cpp, read where it lies
// under the reference checkout, into oracle/_ref/libkmc_ref.so. The result is
// used to pin the oracle restatement (oracle/kmc_oracle.cpp), to write the
// golden fixtures under tests/golden/, and as the timed CPU baseline.
//
// The reference class keeps its state private (comparisons/CPU_Kmeans/include/
// kmeans.h:9-34); the driver opens it for reading and for setting explicit
// centroids so that single stages can be called on chosen inputs.
#include <cstdint>
#include <cstring>
#include <vector>
// Standard headers first, so that the access override below reaches only the
// reference class.
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>

#define private public
#include "kmeans.h"
#undef private

namespace {
struct Points {
  std::vector<Point<float>*> p;
  Points(const float* X, size_t n, uint32_t d) : p(n) {
    for (size_t i = 0; i < n; ++i) p[i] = new Point<float>(const_cast<float*>(X + i * d), d);
  }
  ~Points() {
    for (auto* q : p) delete q;
  }
};
}  // namespace

extern "C" {

// Full reference run: ctor (random init from seed, kmeans.cpp:51-74,84-119),
// run(maxiter) (kmeans.cpp:192-223). C0 receives the initial centroids, C the
// final ones, labels the Point clusters written back by run().
uint64_t kref_run(const float* X, size_t n, uint32_t d, uint32_t k, float tol,
                  int seed, uint64_t maxiter, float* C0, float* C,
                  uint32_t* labels) {
  Points pts(X, n, d);
  Kmeans km(n, d, k, tol, &seed, pts.p.data());
  if (C0) std::memcpy(C0, km.h_centroids, sizeof(float) * (size_t)k * d);
  uint64_t conv = km.run(maxiter);
  if (C) std::memcpy(C, km.h_centroids, sizeof(float) * (size_t)k * d);
  for (size_t i = 0; i < n; ++i) labels[i] = (uint32_t)pts.p[i]->getCluster();
  return conv;
}

// Same as kref_run but starting from caller-given centroids C (in/out).
uint64_t kref_run_from(const float* X, size_t n, uint32_t d, uint32_t k,
                       float tol, uint64_t maxiter, float* C,
                       uint32_t* labels) {
  Points pts(X, n, d);
  int seed = 1;
  Kmeans km(n, d, k, tol, &seed, pts.p.data());
  std::memcpy(km.h_centroids, C, sizeof(float) * (size_t)k * d);
  std::memcpy(km.h_last_centroids, C, sizeof(float) * (size_t)k * d);
  uint64_t conv = km.run(maxiter);
  std::memcpy(C, km.h_centroids, sizeof(float) * (size_t)k * d);
  for (size_t i = 0; i < n; ++i) labels[i] = (uint32_t)pts.p[i]->getCluster();
  return conv;
}

// One assign + update stage on caller-given centroids: distances
// (kmeans.cpp:121-131), argmin + counts (:141-154), centroids (:166-180).
void kref_step(const float* X, size_t n, uint32_t d, uint32_t k,
               const float* C_in, float* D, uint32_t* labels, uint32_t* counts,
               float* C_out) {
  Points pts(X, n, d);
  int seed = 1;
  Kmeans km(n, d, k, 0.0f, &seed, pts.p.data());
  std::memcpy(km.h_centroids, C_in, sizeof(float) * (size_t)k * d);
  km.points_clusters = new uint32_t[n];
  km.compute_distances(D);
  km.clusters_argmin(D, km.points_clusters, counts);
  std::memcpy(labels, km.points_clusters, sizeof(uint32_t) * n);
  km.compute_centroids(counts);
  std::memcpy(C_out, km.h_centroids, sizeof(float) * (size_t)k * d);
}

// The shift test (kmeans.cpp:225-234) on caller-given centroid pairs.
int kref_converged(const float* X, size_t n, uint32_t d, uint32_t k, float tol,
                   const float* C, const float* Clast) {
  Points pts(X, n, d);
  int seed = 1;
  Kmeans km(n, d, k, tol, &seed, pts.p.data());
  std::memcpy(km.h_centroids, C, sizeof(float) * (size_t)k * d);
  std::memcpy(km.h_last_centroids, Clast, sizeof(float) * (size_t)k * d);
  km.points_clusters = new uint32_t[1];
  return km.cmp_centroids() ? 1 : 0;
}

}  // extern "C"

#include <chrono>

extern "C" {

// Times Kmeans::run(maxiter) alone (as comparisons/CPU_Kmeans/src/main.cpp:
// 284-290 does), starting from caller-given centroids; returns seconds.
double kref_time_run(const float* X, size_t n, uint32_t d, uint32_t k,
                     uint64_t maxiter, const float* C0, uint64_t* converged) {
  Points pts(X, n, d);
  int seed = 1;
  Kmeans km(n, d, k, 0.0f, &seed, pts.p.data());
  std::memcpy(km.h_centroids, C0, sizeof(float) * (size_t)k * d);
  std::memcpy(km.h_last_centroids, C0, sizeof(float) * (size_t)k * d);
  const auto t0 = std::chrono::high_resolution_clock::now();
  const uint64_t conv = km.run(maxiter);
  const auto t1 = std::chrono::high_resolution_clock::now();
  if (converged) *converged = conv;
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
