// kmc_kk.h -- parameter block and launchers of the kernel k-means kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "kmc_kernels.h"

namespace kmc {

enum { KK_LINEAR = 0, KK_POLY = 1, KK_SIGMOID = 2, KK_RBF = 3 };
enum { KS_FP32 = 0, KS_BF16 = 1 };
constexpr int KK_MAX_K = 64;

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
  float gamma;
  float* xn2;        // n
  float* kdiag;      // n
  // iteration
  void* HT;          // Kp x Np one-hot (transposed), fp32 or bf16
  int* labels;       // n
  int* count;        // Kp
  float* S;          // n x Kp
  int K, Kp;
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
