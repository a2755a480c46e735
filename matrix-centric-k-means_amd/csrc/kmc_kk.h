// kmc_kk.h -- parameter block and launchers of the kernel k-means kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kmc_kernels.h"

namespace kmc {

enum { KK_LINEAR = 0, KK_POLY = 1, KK_SIGMOID = 2, KK_RBF = 3 };
enum { KS_FP32 = 0, KS_BF16 = 1 };
// arithmetic of the Gram product X X^T in the kernel-matrix build
enum { KB_FP32 = 0, KB_BF16X3 = 1, KB_BF16 = 2 };
// S = K H products: dense one-hot MFMA product, or cluster-sorted row sums
enum { KK_DENSE = 0, KK_SORTED = 1 };
constexpr int KK_MAX_K = 1024;       // clusters (cluster-sorted product)
constexpr int KK_DENSE_MAX_K = 64;   // clusters of the dense product
constexpr int kKKRows = 256;         // sorted rows per chunk of the sorted product

struct KKState {
  unsigned long long iter, maxiter, converged_at;
  int done, pad;
  double objective;        // FP64 sum_i ||phi(x_i) - mu_{l_i}||^2 (last assign)
  unsigned long long changed;  // labels changed by the last assign
  double sum_diag;         // FP64 sum_i K_ii
  double last_score;
};

struct KKParams {
  // set-up
  const float* X;
  size_t ldx;
  long long n;
  int d;
  long long Np;      // padded order of K (multiple of 64)
  void* Kmat;        // Np x Np, fp32 or bf16
  int store, kind;
  int build;         // KB_*
  // bf16 builds: X split once into bf16 hi (and lo) planes, n x dpad
  // (dpad = d rounded up to 32, zero pad), the operands the build stages
  const uint16_t* Xh;
  const uint16_t* Xl;
  int dpad;
  float gamma;
  float* xn2;        // n
  float* kdiag;      // n
  // iteration
  void* HT;          // Kp x Np one-hot (transposed), fp32 or bf16
  int* labels;       // n
  int* count;        // Kp
  float* S;          // dense: n x Kp (S[i*Kp + c]); sorted: Kp x Np (S[c*Np + i])
  long long sS_i, sS_c;  // strides of S along points and clusters
  int K, Kp;
  int product;       // KK_DENSE or KK_SORTED
  // cluster-sorted product: the stable sort of the points by label, chunk
  // partial sums of K rows (row c + q for cluster c in chunk q), per-tile
  // self-term partials
  int* bcount;
  int* start;
  int* members;
  int CH, nblk, lbits;
  float* tpart;      // (nq + K) x Np
  int nq;
  double* gtile;     // K x ntile
  int ntile;
  float* gpart;      // ceil(n/16) x Kp
  double* g;         // Kp
  double* opart;     // nassign
  int* cpart;        // nassign
  int nassign;
  float* Dout;       // optional n x K distance output
  KKState* state;
  const int* done;
  int mode;
  float tol;
};

hipError_t kk_launch_setup(const KKParams& P, hipStream_t s);
hipError_t kk_launch_iter(const KKParams& P, hipEvent_t e0, hipEvent_t e1, hipStream_t s);
hipError_t kk_launch_distances(const KKParams& P, hipStream_t s);

}  // namespace kmc
