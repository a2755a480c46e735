// kmc_device.h -- device-side types and helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __attribute__((ext_vector_type(8))) __bf16 kmc_bf16x8;
typedef __attribute__((ext_vector_type(16))) float kmc_f32x16;
typedef __attribute__((ext_vector_type(4))) float kmc_f32x4;
typedef __attribute__((ext_vector_type(8))) int kmc_i32x8;

namespace kmc {

constexpr int kWave = 64;

// Round-to-nearest-even float -> bf16 bit pattern (NaN kept quiet).
__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
    return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
  return __uint_as_float(((uint32_t)b) << 16);
}

// fp8 (OCP e4m3: 4-bit significand, exponent bias 7, largest finite 448 =
// 1.75 * 2^8, subnormal spacing 2^-9) with one power-of-two scale per row.
// The scale exponent e is the smallest with amax * 2^-e <= 448, so the row's
// largest entry lands in the top binade and no entry saturates; a zero row
// takes e = 0.
__device__ __forceinline__ int fp8_scale_exp(float amax) {
  if (!(amax > 0.0f)) return 0;
  int ex;
  const float f = frexpf(amax, &ex);  // amax = f * 2^ex, f in [0.5, 1)
  const int e = f <= 0.875f ? ex - 9 : ex - 8;
  return e < -127 ? -127 : (e > 127 ? 127 : e);
}

// Round to nearest even onto the e4m3 grid, for |v| <= 448: with E the
// binade of |v| (at least -6, the subnormal range), the grid spacing is
// 2^(E-3), m = rint(|v| / 2^(E-3)) lies in [0, 16] and the code is
// ((E + 7) << 3) + m - 8, which also covers the subnormals (E = -6, m < 8)
// and the carry into the next binade (m = 16).
__device__ __forceinline__ uint32_t fp8_e4m3_rne(float v) {
  const float a = fabsf(v);
  int ex;
  (void)frexpf(a, &ex);
  const int E = a > 0.0f ? max(ex - 1, -6) : -6;
  const float m = __builtin_rintf(ldexpf(a, 3 - E));
  const uint32_t code = (uint32_t)(((E + 7) << 3) + (int)m - 8);
  return code | (v < 0.0f ? 0x80u : 0u);
}

// The value of an e4m3 code (inverse of fp8_e4m3_rne).
__device__ __forceinline__ float fp8_e4m3_value(uint32_t code) {
  const int e = (int)((code >> 3) & 15u), m = (int)(code & 7u);
  const float a = e ? ldexpf(1.0f + 0.125f * (float)m, e - 7) : ldexpf((float)m, -9);
  return (code & 0x80u) ? -a : a;
}

// sqrt of an FP64 sum of squares as a float no smaller than the true norm
__device__ __forceinline__ float norm_up(double s2) {
  return (float)(sqrt(s2) * (1.0 + 1e-6)) * (1.0f + 1e-6f);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Rank of this lane among the set lanes of m below it.
__device__ __forceinline__ unsigned mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Deterministic wave sum (butterfly, fixed order for every lane).
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// Squared distance in the reference's float arithmetic: in index order,
// one rounding for each subtraction, product and addition, no fused
// multiply-add (comparisons/CPU_Kmeans/src/kmeans.cpp:121-131).
__device__ __forceinline__ float direct_dist_f32(const float* __restrict__ x,
                                                 const float* __restrict__ c,
                                                 int d) {
#pragma clang fp contract(off)
  float acc = 0.0f;
  int j = 0;
  // 16-byte loads; the sum itself stays strictly sequential.
  if ((((uintptr_t)x | (uintptr_t)c) & 15u) == 0) {
    for (; j + 4 <= d; j += 4) {
      const float4 xv = *reinterpret_cast<const float4*>(x + j);
      const float4 cv = *reinterpret_cast<const float4*>(c + j);
      float t, s;
      t = xv.x - cv.x; s = t * t; acc = acc + s;
      t = xv.y - cv.y; s = t * t; acc = acc + s;
      t = xv.z - cv.z; s = t * t; acc = acc + s;
      t = xv.w - cv.w; s = t * t; acc = acc + s;
    }
  }
  for (; j < d; ++j) {
    float t = x[j] - c[j];
    float s = t * t;
    acc = acc + s;
  }
  return acc;
}

// Running top-3 of values (b1 <= b2 <= b3) with the indices of the first
// two. Strict '<' keeps the earlier index on equal values.
struct Top3 {
  float b1, b2, b3;
  int i1, i2;
  __device__ __forceinline__ void init() {
    b1 = b2 = b3 = __builtin_inff();
    i1 = i2 = 0x7fffffff;
  }
  __device__ __forceinline__ void push(float v, int ci) {
    const float n3 = __builtin_amdgcn_fmed3f(b2, b3, v);
    const float n2 = __builtin_amdgcn_fmed3f(b1, b2, v);
    const bool lt1 = v < b1;
    const bool lt2 = v < b2;
    i2 = lt1 ? i1 : (lt2 ? ci : i2);
    i1 = lt1 ? ci : i1;
    b1 = fminf(b1, v);
    b2 = n2;
    b3 = n3;
  }
};

}  // namespace kmc
