// kmc_lloyd.hip -- gfx950 kernels of one Lloyd iteration (linear kernel).
//
// Per iteration, in stream order:
//   k_prep       centroid operands (-2c in fp32 / bf16 hi+lo), ||c||^2 (first
//                iteration of a run; later ones get them from k_finish)
//   k_assign     fused distance GEMM on MFMA (||c||^2 - 2 x.c accumulated
//                from the norm), running top-4 per point, margin band test;
//                writes labels and queues the points inside the band
//   k_refine2    band points with two or three candidates: their distances
//                in the reference's float arithmetic, first-min rule
//   k_refine_all band points with four or more candidates: all k distances
//   k_refine_prune
//                long all-k queues instead: bf16x3 MFMA sweeps narrow each
//                point's centroids before the exact distances
//   k_hist / k_scan_blocks / k_scatter
//                stable counting sort of point ids by label (the one-hot
//                V matrix in CSR form: row offsets + column ids)
//   k_update     C = V X with per-cluster float sums in point order, plus the
//                FP64 inertia of the assignment
//   k_finish     per-centroid shift test, C <- C_next, the next iteration's
//                operands; its last workgroup runs the FP64 inertia / shift
//                reductions, the iteration counter and the stop flag
//
// Every kernel returns at once when the device stop flag is set, so the host
// may queue iterations ahead of the convergence decision.
#include <algorithm>
#include <map>
#include <mutex>

#include "kmc.h"
#include "kmc_device.h"
#include "kmc_kernels.h"

namespace kmc {

// ---------------------------------------------------------------------------
// Distance operands of centroid row c, computed by one wave: -2c (fp32, or
// bf16 hi / hi+lo), the transposed copy for the all-k refinement and
// ||c||^2 (kFar for padding rows and excluded clusters). get(j) yields the
// row's value at dimension j (zero past D). Both k_prep and k_finish call
// this, so the operands are the same bits whichever kernel forms them.
// ---------------------------------------------------------------------------
template <class Get>
__device__ __forceinline__ void prep_row(const PrepParams& P, int c, int lane, bool alive,
                                         Get get) {
  const size_t row = (size_t)c * P.DP;
  if (P.prec == PREC_FP8) {
    // -2c quantised with the row's own power-of-two scale (lane owns
    // dimensions lane + 64 t)
    float w[kMaxDP / 64];
    float s = 0.0f, amax = 0.0f;
#pragma unroll
    for (int t = 0; t < kMaxDP / 64; ++t) {
      const int j = lane + 64 * t;
      const float v = (j < P.DP && c < P.K) ? get(j) : 0.0f;
      w[t] = -2.0f * v;  // exact scaling
      s = fmaf(v, v, s);
      amax = fmaxf(amax, fabsf(w[t]));
      if (j < P.DP) {
        P.CT[(size_t)j * P.Kp + c] = v;
        if (P.Phi) {
          const uint16_t hb = bf16_rne(w[t]);
          P.Phi[row + j] = hb;
          P.Plo[row + j] = bf16_rne(w[t] - bf16_to_f32(hb));
        }
      }
    }
    amax = wave_max_f32(amax);
    const int e = fp8_scale_exp(amax);
    uint8_t* q = P.Cq + (size_t)c * (P.DP + kFp8RowPad);
    double d2 = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxDP / 64; ++t) {
      const int j = lane + 64 * t;
      if (j < P.DP) {
        const uint32_t code = fp8_e4m3_rne(ldexpf(w[t], -e));
        q[j] = (uint8_t)code;
        const double dv = (double)ldexpf(fp8_e4m3_value(code), e) - (double)w[t];
        d2 = fma(dv, dv, d2);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    d2 = wave_sum_f64(d2);
    if (lane == 0) {
      *reinterpret_cast<int*>(q + P.DP) = e + 127;
      P.cn[c] = alive ? s : kFar;
      if (P.cdn) P.cdn[c] = norm_up(d2);
    }
    return;
  }
  float s = 0.0f;
  double d2 = 0.0;
  for (int j = lane; j < P.DP; j += 64) {
    const float v = (c < P.K) ? get(j) : 0.0f;
    const float w = -2.0f * v;  // exact scaling
    s = fmaf(v, v, s);
    P.CT[(size_t)j * P.Kp + c] = v;
    if (P.Phi) {
      const uint16_t hb = bf16_rne(w);
      P.Phi[row + j] = hb;
      P.Plo[row + j] = bf16_rne(w - bf16_to_f32(hb));
    }
    if (P.prec == PREC_FP32) {
      P.Cw[row + j] = w;
    } else {
      const uint16_t hi = bf16_rne(w);
      P.Chi[row + j] = hi;
      if (P.prec == PREC_BF16X3) P.Clo[row + j] = bf16_rne(w - bf16_to_f32(hi));
      const double dv = (double)bf16_to_f32(hi) - (double)w;
      d2 = fma(dv, dv, d2);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (P.cdn) d2 = wave_sum_f64(d2);
  if (lane == 0) {
    P.cn[c] = alive ? s : kFar;
    if (P.cdn) P.cdn[c] = norm_up(d2);
  }
}

// k_prep: one wave per (padded) centroid row.
__global__ __launch_bounds__(256) void k_prep(PrepParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= P.Kp) return;
  const float* cr = P.C + (size_t)c * P.DP;
  prep_row(P, c, lane, c < P.K && P.alive[c], [&](int j) { return cr[j]; });
}

// ---------------------------------------------------------------------------
// k_assign
//
// Work split: a workgroup of 4 waves owns a tile of 4*32 points; each wave
// owns 32 points whose operand fragments stay in registers for the whole
// sweep over the centroids. Centroid operands are staged in LDS (all of them
// when they fit, else chunk by chunk) and shared by the 4 waves.
//
// MFMA mapping (32x32 blocks, wave64): A = 32 centroids x k, B = k x 32
// points, D[centroid][point]. Lane l supplies A row (l & 31) and B column
// (l & 31) for the k-slice selected by h = l >> 5; the dimension order inside
// the dot product is permuted so that lanes h=0,1 of one point read adjacent
// 16-byte pieces of the row:
//   fp32  step s : dim = 8*(s>>2) + 4*h + (s&3)          (32x32x2 f32)
//   bf16  step s : dims 16*s + 8*h + [0,8)               (32x32x16 bf16)
//   fp8   step s : dims 64*s + 32*h + [0,32)             (32x32x64 e4m3, scaled)
// Accumulator register r of lane l holds centroid 8*(r>>2) + 4*h + (r&3) of
// the block against point (l & 31). The accumulator starts at ||c||^2 and the
// A operand is -2c, so the MFMA result is ||c||^2 - 2 x.c directly.
// ---------------------------------------------------------------------------
template <int DP, int PREC>
struct AssignTraits;

template <int DP>
struct AssignTraits<DP, PREC_FP32> {
  static constexpr int kRowElems = DP + 4;  // floats per LDS row (padded)
  static constexpr int kRowBytes = kRowElems * 4;
  static constexpr int kPlanes = 1;
};
template <int DP>
struct AssignTraits<DP, PREC_BF16X3> {
  static constexpr int kRowElems = DP + 8;  // bf16 per LDS row (padded)
  static constexpr int kRowBytes = kRowElems * 2;
  static constexpr int kPlanes = 2;  // hi, lo
};
template <int DP>
struct AssignTraits<DP, PREC_BF16> {
  static constexpr int kRowElems = DP + 8;
  static constexpr int kRowBytes = kRowElems * 2;
  static constexpr int kPlanes = 1;
};

template <int DP>
struct AssignTraits<DP, PREC_FP8> {
  static constexpr int kRowElems = DP + kFp8RowPad;  // bytes: e4m3 row, then its scale
  static constexpr int kRowBytes = kRowElems;
  static constexpr int kPlanes = 1;
};

template <int DP, int PREC>
size_t assign_lds_bytes(int kchunk) {
  using T = AssignTraits<DP, PREC>;
  return (size_t)kchunk * 4 + (size_t)kchunk * T::kRowBytes * T::kPlanes + 64;
}

// Copy centroid rows [c0, c0+cnt) of the prepared operands into LDS.
template <int DP, int PREC>
__device__ __forceinline__ void stage_centroids(const AssignParams& P,
                                                unsigned char* smem, int c0,
                                                int cnt) {
  using T = AssignTraits<DP, PREC>;
  float* cn_s = reinterpret_cast<float*>(smem);
  unsigned char* rows = smem + (size_t)cnt * 4;
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) cn_s[i] = P.cn[c0 + i];
  constexpr int kVecPerRow = (PREC == PREC_FP32)  ? DP / 4
                             : (PREC == PREC_FP8) ? (DP + kFp8RowPad) / 16
                                                  : DP / 8;  // 16-B units
  const int total = cnt * kVecPerRow;
  for (int plane = 0; plane < T::kPlanes; ++plane) {
    const uint4* src;
    if (PREC == PREC_FP32)
      src = reinterpret_cast<const uint4*>(P.Cw + (size_t)c0 * DP);
    else if (PREC == PREC_FP8)
      src = reinterpret_cast<const uint4*>(P.Cq + (size_t)c0 * (DP + kFp8RowPad));
    else
      src = reinterpret_cast<const uint4*>((plane == 0 ? P.Chi : P.Clo) + (size_t)c0 * DP);
    unsigned char* dst = rows + (size_t)plane * cnt * T::kRowBytes;
    for (int v = threadIdx.x; v < total; v += blockDim.x) {
      const int r = v / kVecPerRow, q = v - r * kVecPerRow;
      *reinterpret_cast<uint4*>(dst + (size_t)r * T::kRowBytes + q * 16) = src[v];
    }
  }
}

// Register-resident B-operand fragments of one point (this lane's k-slice).
// fp8: 32 bytes (two uint4) per 64-dimension step, plus the row's scale.
template <int DP, int PREC>
struct XFrag {
  static constexpr int NF = (PREC == PREC_FP32) ? DP / 2 : 1;
  static constexpr int NH = (PREC == PREC_BF16 || PREC == PREC_BF16X3) ? DP / 16 : 1;
  static constexpr int NL = (PREC == PREC_BF16X3) ? DP / 16 : 1;
  static constexpr int NQ = (PREC == PREC_FP8) ? DP / 32 : 1;
  float f[NF];
  uint4 hi[NH];
  uint4 lo[NL];
  uint4 q[NQ];
  int sc;
};

template <int DP, int PREC>
__device__ __forceinline__ void load_xfrag(const AssignParams& P, long long prow, bool valid,
                                           int h, XFrag<DP, PREC>& x) {
  const size_t row = (size_t)(valid ? prow : 0) * DP;
  if constexpr (PREC == PREC_FP8) {
    // step s: bytes 64 s + 32 h + [0, 32) of the row
    const uint8_t* xr = P.Xq + row;
#pragma unroll
    for (int s = 0; s < DP / 64; ++s) {
      x.q[2 * s] = valid ? *reinterpret_cast<const uint4*>(xr + 64 * s + 32 * h)
                         : make_uint4(0u, 0u, 0u, 0u);
      x.q[2 * s + 1] = valid ? *reinterpret_cast<const uint4*>(xr + 64 * s + 32 * h + 16)
                             : make_uint4(0u, 0u, 0u, 0u);
    }
    x.sc = valid ? P.xsc[prow] : 127;
  } else if (PREC == PREC_FP32) {
    const float* xr = P.X + row;
#pragma unroll
    for (int q = 0; q < DP / 8; ++q) {
      const float4 v = valid ? *reinterpret_cast<const float4*>(xr + 8 * q + 4 * h)
                             : make_float4(0.f, 0.f, 0.f, 0.f);
      x.f[4 * q + 0] = v.x;
      x.f[4 * q + 1] = v.y;
      x.f[4 * q + 2] = v.z;
      x.f[4 * q + 3] = v.w;
    }
  } else {
    const uint16_t* xr = P.Xhi + row;
#pragma unroll
    for (int s = 0; s < DP / 16; ++s)
      x.hi[s] = valid ? *reinterpret_cast<const uint4*>(xr + 16 * s + 8 * h)
                      : make_uint4(0u, 0u, 0u, 0u);
    if (PREC == PREC_BF16X3) {
      const uint16_t* xq = P.Xlo + row;
#pragma unroll
      for (int s = 0; s < DP / 16; ++s)
        x.lo[s] = valid ? *reinterpret_cast<const uint4*>(xq + 16 * s + 8 * h)
                        : make_uint4(0u, 0u, 0u, 0u);
    }
  }
}

// The same for the single-product forms without the zero fill: the row is
// clamped to row 0 instead, so the loads carry no lane mask and need no
// cleared destination (the values of an invalid point are never used).
template <int DP, int PREC>
__device__ __forceinline__ void load_xfrag_any(const AssignParams& P, long long prow, bool valid,
                                               int h, XFrag<DP, PREC>& x) {
  static_assert(PREC == PREC_FP8 || PREC == PREC_BF16, "single-product forms only");
  const long long r = valid ? prow : 0;
  if constexpr (PREC == PREC_FP8) {
    const uint8_t* xr = P.Xq + (size_t)r * DP;
#pragma unroll
    for (int s = 0; s < DP / 64; ++s) {
      x.q[2 * s] = *reinterpret_cast<const uint4*>(xr + 64 * s + 32 * h);
      x.q[2 * s + 1] = *reinterpret_cast<const uint4*>(xr + 64 * s + 32 * h + 16);
    }
    x.sc = P.xsc[r];
  } else {
    const uint16_t* xr = P.Xhi + (size_t)r * DP;
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) x.hi[s] = *reinterpret_cast<const uint4*>(xr + 16 * s + 8 * h);
  }
}

// Registers per lane of one point set's fragments.
template <int DP, int PREC>
constexpr int frag_regs() {
  return PREC == PREC_FP32     ? DP / 2
         : PREC == PREC_BF16X3 ? DP / 2
         : PREC == PREC_FP8    ? DP / 8 + 1
                               : DP / 4;
}

// Register-resident A-operand fragments of one 32-centroid block (lane l:
// centroid row cb + (l & 31), k-slice h = l >> 5, same permuted dimension
// order as the point fragments).
template <int DP, int PREC>
struct AFrag {
  static constexpr int NF = (PREC == PREC_FP32) ? DP / 8 : 1;
  static constexpr int NH = (PREC == PREC_BF16 || PREC == PREC_BF16X3) ? DP / 16 : 1;
  static constexpr int NL = (PREC == PREC_BF16X3) ? DP / 16 : 1;
  static constexpr int NQ = (PREC == PREC_FP8) ? DP / 32 : 1;
  float4 f[NF];
  uint4 hi[NH];
  uint4 lo[NL];
  uint4 q[NQ];
  int sc;
};

template <int DP, int PREC>
__device__ __forceinline__ void load_afrag(const unsigned char* rows, int cnt, int cb, int p,
                                           int h, AFrag<DP, PREC>& a) {
  using T = AssignTraits<DP, PREC>;
  const unsigned char* arow = rows + (size_t)(cb + p) * T::kRowBytes;
  if constexpr (PREC == PREC_FP8) {
#pragma unroll
    for (int s = 0; s < DP / 64; ++s) {
      a.q[2 * s] = *reinterpret_cast<const uint4*>(arow + 64 * s + 32 * h);
      a.q[2 * s + 1] = *reinterpret_cast<const uint4*>(arow + 64 * s + 32 * h + 16);
    }
    a.sc = *reinterpret_cast<const int*>(arow + DP);
  } else if (PREC == PREC_FP32) {
    const float* af = reinterpret_cast<const float*>(arow);
#pragma unroll
    for (int q = 0; q < DP / 8; ++q) a.f[q] = *reinterpret_cast<const float4*>(af + 8 * q + 4 * h);
  } else {
    const uint16_t* ah = reinterpret_cast<const uint16_t*>(arow);
    const uint16_t* al = reinterpret_cast<const uint16_t*>(
        rows + (size_t)cnt * T::kRowBytes + (size_t)(cb + p) * T::kRowBytes);
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) {
      a.hi[s] = *reinterpret_cast<const uint4*>(ah + 16 * s + 8 * h);
      if (PREC == PREC_BF16X3) a.lo[s] = *reinterpret_cast<const uint4*>(al + 16 * s + 8 * h);
    }
  }
}

// Accumulator start for the block at cb: ||c||^2, plus ||x||^2 when kAddX2.
template <bool kAddX2>
__device__ __forceinline__ kmc_f32x16 init_acc(const float* cn_s, int cb, int h, float x2) {
  kmc_f32x16 acc;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 n4 = *reinterpret_cast<const float4*>(cn_s + cb + 8 * g + 4 * h);
    acc[4 * g + 0] = kAddX2 ? n4.x + x2 : n4.x;
    acc[4 * g + 1] = kAddX2 ? n4.y + x2 : n4.y;
    acc[4 * g + 2] = kAddX2 ? n4.z + x2 : n4.z;
    acc[4 * g + 3] = kAddX2 ? n4.w + x2 : n4.w;
  }
  return acc;
}

// The MFMA chain of one block against one point set: acc += (-2c) . x
// (bf16x3: hi.hi + hi.lo + lo.hi).
// One scaled fp8 step: 32x32x64 e4m3 (format code 0 for both operands), the
// per-lane E8M0 scales of the centroid row and of the point applied by the
// matrix core, so the product is (2^ea * a) . (2^eb * b).
__device__ __forceinline__ kmc_f32x16 mfma_fp8_step(uint4 a0, uint4 a1, int sa, uint4 b0,
                                                    uint4 b1, int sb, kmc_f32x16 acc) {
  const kmc_i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                        (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
  const kmc_i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                        (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, sa, 0, sb);
}

template <int DP, int PREC>
__device__ __forceinline__ kmc_f32x16 mfma_chain(const AFrag<DP, PREC>& a,
                                                 const XFrag<DP, PREC>& x, kmc_f32x16 acc) {
  if constexpr (PREC == PREC_FP8) {
#pragma unroll
    for (int s = 0; s < DP / 64; ++s)
      acc = mfma_fp8_step(a.q[2 * s], a.q[2 * s + 1], a.sc, x.q[2 * s], x.q[2 * s + 1], x.sc,
                          acc);
  } else if (PREC == PREC_FP32) {
#pragma unroll
    for (int q = 0; q < DP / 8; ++q) {
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.f[q].x, x.f[4 * q + 0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.f[q].y, x.f[4 * q + 1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.f[q].z, x.f[4 * q + 2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.f[q].w, x.f[4 * q + 3], acc, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a.hi[s]),
                                                    __builtin_bit_cast(kmc_bf16x8, x.hi[s]),
                                                    acc, 0, 0, 0);
      if (PREC == PREC_BF16X3) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a.hi[s]),
                                                      __builtin_bit_cast(kmc_bf16x8, x.lo[s]),
                                                      acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a.lo[s]),
                                                      __builtin_bit_cast(kmc_bf16x8, x.hi[s]),
                                                      acc, 0, 0, 0);
      }
    }
  }
  return acc;
}

// The same chain with each A fragment read from LDS just before its MFMA
// (fewer live registers).
template <int DP, int PREC>
__device__ __forceinline__ kmc_f32x16 mfma_chain_lds(const unsigned char* rows, int cnt, int cb,
                                                     int p, int h, const XFrag<DP, PREC>& x,
                                                     kmc_f32x16 acc) {
  using T = AssignTraits<DP, PREC>;
  const unsigned char* arow = rows + (size_t)(cb + p) * T::kRowBytes;
  if constexpr (PREC == PREC_FP8) {
    const int sa = *reinterpret_cast<const int*>(arow + DP);
#pragma unroll
    for (int s = 0; s < DP / 64; ++s) {
      const uint4 a0 = *reinterpret_cast<const uint4*>(arow + 64 * s + 32 * h);
      const uint4 a1 = *reinterpret_cast<const uint4*>(arow + 64 * s + 32 * h + 16);
      acc = mfma_fp8_step(a0, a1, sa, x.q[2 * s], x.q[2 * s + 1], x.sc, acc);
    }
  } else if (PREC == PREC_FP32) {
    const float* af = reinterpret_cast<const float*>(arow);
#pragma unroll
    for (int q = 0; q < DP / 8; ++q) {
      const float4 a = *reinterpret_cast<const float4*>(af + 8 * q + 4 * h);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, x.f[4 * q + 0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, x.f[4 * q + 1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, x.f[4 * q + 2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, x.f[4 * q + 3], acc, 0, 0, 0);
    }
  } else {
    const uint16_t* ah = reinterpret_cast<const uint16_t*>(arow);
    const uint16_t* al = reinterpret_cast<const uint16_t*>(
        rows + (size_t)cnt * T::kRowBytes + (size_t)(cb + p) * T::kRowBytes);
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) {
      const uint4 a_hi = *reinterpret_cast<const uint4*>(ah + 16 * s + 8 * h);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a_hi),
                                                    __builtin_bit_cast(kmc_bf16x8, x.hi[s]),
                                                    acc, 0, 0, 0);
      if (PREC == PREC_BF16X3) {
        const uint4 a_lo = *reinterpret_cast<const uint4*>(al + 16 * s + 8 * h);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a_hi),
                                                      __builtin_bit_cast(kmc_bf16x8, x.lo[s]),
                                                      acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a_lo),
                                                      __builtin_bit_cast(kmc_bf16x8, x.hi[s]),
                                                      acc, 0, 0, 0);
      }
    }
  }
  return acc;
}

// Top-4 update over one block's 16 accumulator values. Each value carries
// its centroid index in the low mantissa bits ((v & ~mask) | idx), so the
// running top-4 is four v_med3_f32 per value and the indices travel with the
// values. The packed index is the wave-uniform part base + 8*(r>>2) + (r&3)
// (bit 2 always clear); the lane's k-half bit 4h is or-ed in once, when the
// top-4 is read out. The perturbation of the low bits is part of the
// margin-band bound.
// Insert v into the running top-4 b1 <= b2 <= b3 <= b4 (the k-th smallest of
// {b1..b4, v} is med3(b_{k-1}, b_k, v)).
// ninf is -inf held in a vector register the compiler cannot see through,
// so the first insertion stays one v_med3_f32 (a min would be preceded by a
// canonicalisation of the bit-built value).
__device__ __forceinline__ void top4_push(float v, float (&b)[4], float ninf) {
  b[3] = __builtin_amdgcn_fmed3f(b[2], b[3], v);
  b[2] = __builtin_amdgcn_fmed3f(b[1], b[2], v);
  b[1] = __builtin_amdgcn_fmed3f(b[0], b[1], v);
  b[0] = __builtin_amdgcn_fmed3f(ninf, b[0], v);
}

// keepv is the mask in a vector register (the wave-uniform index base is a
// scalar; one v_and_or_b32 may read only one scalar operand).
__device__ __forceinline__ void push_block(const kmc_f32x16& acc, uint32_t base, uint32_t keepv,
                                           float (&b)[4], float ninf) {
#pragma unroll
  for (int r = 0; r < 16; ++r)
    top4_push(__uint_as_float((__float_as_uint(acc[r]) & keepv) |
                              (base + (uint32_t)(8 * (r >> 2) + (r & 3)))),
              b, ninf);
}

// A copy of a wave-uniform value in a vector register, opaque to the
// compiler's uniformity analysis.
__device__ __forceinline__ uint32_t opaque_vgpr(uint32_t s) {
  uint32_t v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// Work split: a workgroup of NW waves shares one LDS copy of the centroid
// operands (all of them when they fit, else chunk by chunk); each wave owns
// MS sets of 32 points per tile, keeps their operand fragments in registers
// for the whole sweep over the centroids and, with PF, prefetches the
// fragments of its next tile while sweeping. Every A fragment read from LDS
// feeds MS MFMA chains (MS = 2 halves the LDS operand traffic), and the
// chains of one block run while the previous block's epilogue executes.
// The grid is persistent. (NW, MS, PF) are chosen per problem by the
// autotuner (kmc_autotune); the LDS copy limits a CU to one workgroup when
// all centroids are resident, so NW sets the waves per SIMD.
//
// Single-product bf16 and fp8 accumulate ||c||^2 - 2 x.c without ||x||^2
// (the ranking does not depend on it) and add ||x||^2 to the four candidates
// after the sweep: one VALU op less per value. The packing perturbation is
// then relative to | ||c||^2 - 2 x.c | <= d^2 + ||x||^2, which the band
// covers through the xpk * ||x||^2 term.
template <int DP, int PREC, int NW, int MS, bool PF, int WPE, bool BATCH = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_assign(AssignParams P) {
  if (*P.done) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ float red_s[NW];
  __shared__ unsigned q2c_s, qac_s;  // this workgroup's refinement queue fill

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p = lane & 31;
  const int h = lane >> 5;
  constexpr int kTile = 32 * MS * NW;  // points per workgroup tile
  constexpr bool kAddX2 = PREC != PREC_BF16 && PREC != PREC_FP8;
  constexpr bool kLean = WPE > 4;  // occupancy above four waves per SIMD
  if (threadIdx.x == 0) {
    q2c_s = 0u;
    qac_s = 0u;
  }
  // each workgroup appends to its own queue region (no global atomics)
  int* const q2 = P.q2 + 4 * (size_t)blockIdx.x * P.qstride;
  int* const qa = P.qa + (size_t)blockIdx.x * P.qstride;

  // max ||c|| over the live operand rows [lo, hi) (for the error bound)
  auto cnorm_max = [&](int lo, int hi) {
    float cm = 0.0f;
    for (int c = lo + threadIdx.x; c < hi; c += blockDim.x) {
      const float v = P.cn[c];
      if (v < kFarTest) cm = fmaxf(cm, v);  // excluded clusters carry kFar
    }
    cm = wave_max_f32(cm);
    __syncthreads();  // red_s free
    if (lane == 0) red_s[wave] = cm;
    __syncthreads();
    float cm_all = 0.0f;
#pragma unroll
    for (int w = 0; w < NW; ++w) cm_all = fmaxf(cm_all, red_s[w]);
    return sqrtf(cm_all);
  };
  float cmax = 0.0f, Ecm = 0.0f, cdmax = 0.0f;
  if constexpr (!BATCH) {
    cmax = cnorm_max(0, P.K);
    Ecm = P.beta * cmax * cmax;
    if (P.hop > 0.0f) {
      // max measured operand error of the centroid rows (all rows: padding
      // rows carry 0, excluded clusters only make the bound larger)
      float cd = 0.0f;
      for (int c = threadIdx.x; c < P.Kp; c += blockDim.x) cd = fmaxf(cd, P.cdn[c]);
      cd = wave_max_f32(cd);
      __syncthreads();
      if (lane == 0) red_s[wave] = cd;
      __syncthreads();
#pragma unroll
      for (int w = 0; w < NW; ++w) cdmax = fmaxf(cdmax, red_s[w]);
    }
  }
  const uint32_t keepv = opaque_vgpr(~P.idx_mask);
  const float ninf = __uint_as_float(opaque_vgpr(0xff800000u));

  const bool resident = BATCH || P.kchunk >= P.Kp;
  if (!BATCH && resident) {
    stage_centroids<DP, PREC>(P, smem, 0, P.Kp);
    __syncthreads();
  }

  // tiles: a single problem strides over its tiles; a batch walks a
  // contiguous run of the packed (problem, tile) queue, so consecutive tiles
  // mostly share a problem and its staged centroids
  long long t_begin, t_end, t_step;
  if constexpr (BATCH) {
    const long long per = ((long long)P.nitems + gridDim.x - 1) / gridDim.x;
    t_begin = (long long)blockIdx.x * per;
    t_end = min((long long)P.nitems, t_begin + per);
    t_step = 1;
  } else {
    t_begin = blockIdx.x;
    t_end = (P.n + kTile - 1) / kTile;
    t_step = gridDim.x;
  }
  // prefetch the next tile only while two fragment sets fit the registers
  constexpr int kFragRegs = frag_regs<DP, PREC>();
  constexpr bool kPrefetch = PF && MS == 1 && kFragRegs <= 64 && !BATCH;
  // point row of set m of tile t and its bound (single problem)
  auto row_of = [&](long long t, int m) {
    return t * kTile + (long long)(wave * MS + m) * 32 + p;
  };
  XFrag<DP, PREC> cur[MS], nxt;
  if (!BATCH && t_begin < t_end) {
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const long long r = row_of(t_begin, m);
      load_xfrag<DP, PREC>(P, r, r < P.n, h, cur[m]);
    }
  }
  int staged = -1;  // batch: problem whose centroids are in LDS
  for (long long tile = t_begin; tile < t_end; tile += t_step) {
    const long long ntile = tile + t_step;
    // this tile's rows: [rbase, rbase + rlim) of the packed point space
    long long rbase = 0, rlim = P.n;
    int Kp = P.Kp, Kv = P.K, cbase = 0;
    long long lt = tile;
    if constexpr (BATCH) {
      const int pb = P.item_prob[tile];
      if (P.pdone[pb]) continue;  // converged problem: frozen
      const BatchProblem bpr = P.bp[pb];
      lt = tile - bpr.tile0;
      rbase = bpr.poff;
      rlim = bpr.n;
      Kp = bpr.kpad;
      Kv = bpr.k;
      cbase = bpr.coff;
      if (pb != staged) {
        __syncthreads();  // every wave is done with the previous problem
        stage_centroids<DP, PREC>(P, smem, bpr.cpoff, Kp);
        cmax = cnorm_max(bpr.cpoff, bpr.cpoff + Kp);  // synchronises
        Ecm = P.beta * cmax * cmax;
        staged = pb;
      }
#pragma unroll
      for (int m = 0; m < MS; ++m) {
        const long long r = row_of(lt, m);
        load_xfrag<DP, PREC>(P, rbase + r, r < rlim, h, cur[m]);
      }
    }
    if (kPrefetch && ntile < t_end) {
      const long long r = row_of(ntile, 0);
      load_xfrag<DP, PREC>(P, r, r < P.n, h, nxt);
    }

    float b[MS][4], x2[MS], xd[MS];
#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const long long r = row_of(lt, m);
      x2[m] = r < rlim ? P.xn2[rbase + r] : 0.0f;
      // the measured operand error, needed after the sweep: loaded now so
      // that its latency hides behind the sweep
      xd[m] = (!BATCH && P.hop > 0.0f && r < rlim) ? P.xdn[rbase + r] : 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) b[m][k] = kFar;
    }
    const int kchunk = BATCH ? Kp : P.kchunk;
    for (int c0 = 0; c0 < Kp; c0 += kchunk) {
      const int cnt = min(kchunk, Kp - c0);
      if (!resident) {
        __syncthreads();
        stage_centroids<DP, PREC>(P, smem, c0, cnt);
        __syncthreads();
      }
      const float* cn_s = reinterpret_cast<const float*>(smem);
      const unsigned char* rows = smem + (size_t)cnt * 4;
      // Software pipeline over the 32-centroid blocks of the chunk: the
      // MFMA chains of block s are issued before the top-4 epilogue of block
      // s-1 (independent registers), so one wave keeps both the matrix
      // pipe and the VALU busy. The two accumulator sets alternate.
      auto chains = [&](int blk, kmc_f32x16(&acc)[MS]) {
        AFrag<DP, PREC> a;
        load_afrag<DP, PREC>(rows, cnt, 32 * blk, p, h, a);
#pragma unroll
        for (int m = 0; m < MS; ++m)
          acc[m] = mfma_chain<DP, PREC>(a, cur[m], init_acc<kAddX2>(cn_s, 32 * blk, h, x2[m]));
      };
      auto epilogue = [&](int blk, const kmc_f32x16(&acc)[MS]) {
#pragma unroll
        for (int m = 0; m < MS; ++m)
          push_block(acc[m], (uint32_t)(c0 + 32 * blk), keepv, b[m], ninf);
      };
      const int nblk = cnt / 32;  // even: chunks are multiples of 64
      if constexpr (kLean) {
        // register-lean form for high occupancy: one accumulator set, the
        // A fragments read from LDS step by step
#pragma unroll 1
        for (int blk = 0; blk < nblk; ++blk) {
          kmc_f32x16 acc[MS];
#pragma unroll
          for (int m = 0; m < MS; ++m)
            acc[m] = mfma_chain_lds<DP, PREC>(rows, cnt, 32 * blk, p, h, cur[m],
                                              init_acc<kAddX2>(cn_s, 32 * blk, h, x2[m]));
          epilogue(blk, acc);
        }
      } else {
        kmc_f32x16 accA[MS], accB[MS];
        chains(0, accA);
        int blk = 1;
        for (; blk + 1 < nblk; blk += 2) {
          chains(blk, accB);
          epilogue(blk - 1, accA);
          chains(blk + 1, accA);
          epilogue(blk, accB);
        }
        chains(blk, accB);  // blk == nblk - 1
        epilogue(blk - 1, accA);
        epilogue(blk, accB);
      }
    }

#pragma unroll
    for (int m = 0; m < MS; ++m) {
      const long long lrow = row_of(lt, m);
      const bool valid = lrow < rlim;
      const long long prow = rbase + lrow;
      float b1, b2, b3, b4;
      {
        // this lane's k-half into bit 2 of the packed indices, then merge
        // the two halves' top-4
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          v[k] = __uint_as_float(__float_as_uint(b[m][k]) | (uint32_t)(4 * h));
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = __shfl_xor(v[k], 32, 64);
#pragma unroll
        for (int k = 0; k < 4; ++k) top4_push(o[k], v, ninf);
        b1 = v[0];
        b2 = v[1];
        b3 = v[2];
        b4 = v[3];
      }
      const int i1 = (int)(__float_as_uint(b1) & P.idx_mask);
      const int i2 = (int)(__float_as_uint(b2) & P.idx_mask);
      const int i3 = (int)(__float_as_uint(b3) & P.idx_mask);
      if (!kAddX2) {
        b1 += x2[m];
        b2 += x2[m];
        b3 += x2[m];
        b4 += x2[m];
      }

      // ---- margin band test and output ----
      // b1 <= b2 <= b3 <= b4 are full squared distances (plus the packed
      // index bits); E bounds the MFMA error of each, gref the reference's
      // own float error plus the packing perturbation, both relative to the
      // distances.
      bool flag2 = false, flagA = false, few3 = false;
      if (h == 0 && valid) {
        const float xn = sqrtf(x2[m]);
        float E = P.alpha * xn * cmax + Ecm + P.xpk * x2[m];
        if (!BATCH && P.hop > 0.0f) {
          // |stored(-2c).stored(x) - (-2c).x| <= ||dc|| ||x|| + ||-2c|| ||dx|| + ||dc|| ||dx||
          const float dx = xd[m];
          E += P.hop * (cdmax * xn + 2.0f * cmax * dx + cdmax * dx);
        }
        const float twoE = 2.0f * E;
        const float b1p = fmaxf(b1 + E, 0.0f);
        // clear(b): every centroid ranked at b or later loses to the first
        auto clear = [&](float bb) {
          return (bb >= kFarTest) || (bb - b1 > twoE + P.gref * (fmaxf(bb + E, 0.0f) + b1p));
        };
        const bool uniq = clear(b2);
        const bool two = clear(b3);
        const bool three = clear(b4);
        // a point without any finite distance goes to the all-k pass, which
        // applies the first-min rule
        const bool sane = (b1 < kFarTest) && (i1 < Kv);
        P.labels[prow] = cbase + (sane ? i1 : 0);
        flag2 = sane && !uniq && (two || three);
        flagA = !sane || (!uniq && !two && !three);
        few3 = !two;
      }
      const uint64_t m2 = __ballot(flag2);
      const uint64_t mA = __ballot(flagA);
      if (m2) {
        const int leader = __ffsll((unsigned long long)m2) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&q2c_s, (unsigned)__popcll(m2));
        base = __shfl(base, leader, 64);
        if (flag2) {
          const unsigned pos = base + mbcnt64(m2);
          *reinterpret_cast<int4*>(q2 + 4 * (size_t)pos) =
              make_int4((int)prow, cbase + i1, cbase + i2, few3 ? cbase + i3 : -1);
        }
      }
      if (mA) {
        const int leader = __ffsll((unsigned long long)mA) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&qac_s, (unsigned)__popcll(mA));
        base = __shfl(base, leader, 64);
        if (flagA) qa[base + mbcnt64(mA)] = (int)prow;
      }
    }
    if constexpr (!BATCH) {
      if (kPrefetch) {
        cur[0] = nxt;
      } else if (ntile < t_end) {
#pragma unroll
        for (int m = 0; m < MS; ++m) {
          const long long r = row_of(ntile, m);
          load_xfrag<DP, PREC>(P, r, r < P.n, h, cur[m]);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    P.q2cnt[blockIdx.x] = q2c_s;
    P.qacnt[blockIdx.x] = qac_s;
  }
}

// ---------------------------------------------------------------------------
// Two-set distance kernel for the single-product forms (bf16, fp8)
// ---------------------------------------------------------------------------
// Insert v into a running list of the T smallest values b0 <= ... <= b(T-1)
// (T med3 operations; the k-th smallest of {b, v} is med3(b(k-1), b(k), v)).
// Insert packed value v into the ascending list b of the T smallest (T
// v_med3 operations; the value that falls off the end is dropped).
template <int T>
__device__ __forceinline__ void topT_push(float v, float (&b)[T], float ninf) {
#pragma unroll
  for (int k = T - 1; k >= 1; --k) b[k] = __builtin_amdgcn_fmed3f(b[k - 1], b[k], v);
  b[0] = __builtin_amdgcn_fmed3f(ninf, b[0], v);
}

// Rank one block's 16 accumulator values of a lane (each with its centroid
// index packed into the low mantissa bits) into L lists: value r goes to
// list r % L, each list keeps its own T smallest, and consecutive values
// update different lists, so their dependency chains interleave. A value
// costs 1 + T VALU operations (five for one list of four, four with T = 3).
template <int T, int L>
__device__ __forceinline__ void push_blockL(const kmc_f32x16& acc, uint32_t base, uint32_t keepv,
                                            float (&b)[L][T], float ninf) {
#pragma unroll
  for (int r = 0; r < 16; ++r)
    topT_push<T>(__uint_as_float((__float_as_uint(acc[r]) & keepv) |
                                 (base + (uint32_t)(8 * (r >> 2) + (r & 3)))),
                 b[r % L], ninf);
}

// The four smallest values kept by a lane's L lists (kFar where fewer) and
// t = the smallest of the lists' T-th values: every value a list dropped is
// at least t.
template <int T, int L>
__device__ __forceinline__ void merge_lists(const float (&l)[L][T], float ninf, float (&v)[4],
                                            float& t) {
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = k < T ? l[0][k < T ? k : 0] : kFar;
  t = l[0][T - 1];
#pragma unroll
  for (int j = 1; j < L; ++j) {
    t = __builtin_amdgcn_fmed3f(ninf, t, l[j][T - 1]);
#pragma unroll
    for (int k = 0; k < T; ++k) top4_push(l[j][k], v, ninf);
  }
}

// The A operand of one 32-centroid block as a lane reads it from LDS (its
// k-slices of -2c for centroid row cb + (lane & 31), and the block's
// ||c||^2 row as the accumulator start).
template <int DP, int PREC>
struct ABlock {
  static constexpr int NA = PREC == PREC_FP8 ? DP / 32 : DP / 16;
  uint4 a[NA];
  int sa;
  kmc_f32x16 c;
};

template <int DP, int PREC>
__device__ __forceinline__ void load_ablock(const unsigned char* rows, const float* cn_s, int cb,
                                            int p, int h, ABlock<DP, PREC>& ab) {
  using T = AssignTraits<DP, PREC>;
  const unsigned char* arow = rows + (size_t)(cb + p) * T::kRowBytes;
  ab.c = init_acc<false>(cn_s, cb, h, 0.0f);
  if constexpr (PREC == PREC_FP8) {
    ab.sa = *reinterpret_cast<const int*>(arow + DP);
#pragma unroll
    for (int s = 0; s < DP / 64; ++s) {
      ab.a[2 * s] = *reinterpret_cast<const uint4*>(arow + 64 * s + 32 * h);
      ab.a[2 * s + 1] = *reinterpret_cast<const uint4*>(arow + 64 * s + 32 * h + 16);
    }
  } else {
    ab.sa = 0;
    const uint16_t* ah = reinterpret_cast<const uint16_t*>(arow);
#pragma unroll
    for (int s = 0; s < DP / 16; ++s)
      ab.a[s] = *reinterpret_cast<const uint4*>(ah + 16 * s + 8 * h);
  }
}

// The MFMA chains of one block against two point sets: every A fragment
// feeds both chains; both start from the block's ||c||^2 row.
template <int DP, int PREC>
__device__ __forceinline__ void mfma_chain2(const ABlock<DP, PREC>& ab,
                                            const XFrag<DP, PREC>& x0,
                                            const XFrag<DP, PREC>& x1, kmc_f32x16& a0,
                                            kmc_f32x16& a1) {
  if constexpr (PREC == PREC_FP8) {
#pragma unroll
    for (int s = 0; s < DP / 64; ++s) {
      a0 = mfma_fp8_step(ab.a[2 * s], ab.a[2 * s + 1], ab.sa, x0.q[2 * s], x0.q[2 * s + 1],
                         x0.sc, s == 0 ? ab.c : a0);
      a1 = mfma_fp8_step(ab.a[2 * s], ab.a[2 * s + 1], ab.sa, x1.q[2 * s], x1.q[2 * s + 1],
                         x1.sc, s == 0 ? ab.c : a1);
    }
  } else {
    static_assert(PREC == PREC_BF16, "two-set chains: single-product forms only");
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) {
      const kmc_bf16x8 av = __builtin_bit_cast(kmc_bf16x8, ab.a[s]);
      a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(kmc_bf16x8, x0.hi[s]),
                                                   s == 0 ? ab.c : a0, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(kmc_bf16x8, x1.hi[s]),
                                                   s == 0 ? ab.c : a1, 0, 0, 0);
    }
  }
}

// Band test and output of both point sets of a two-set sweep in one pass:
// lane half h serves set h. Each lane merges its own k-half's lists of set h
// and receives the other half's merged lists of the same set (one exchange
// of up to four values and their bound t), so the merge and the band test
// run once per lane instead of once per set on every lane.
template <int T, int L>
__device__ __forceinline__ void band_out2(const AssignParams& P, const float (&la)[L][T],
                                          const float (&lb)[L][T], float x2, float xd,
                                          bool valid, long long prow, float cmax, float Ecm,
                                          float cdmax, float ninf, int h, int lane, int* q2,
                                          int* qa, unsigned* q2c_s, unsigned* qac_s) {
  // M values of each lane's merged lists cross to the other half
  constexpr int M = L * T < 4 ? L * T : 4;
  float v[4], s4[4], o[M], t, ts;
  merge_lists<T, L>(h ? lb : la, ninf, v, t);
  merge_lists<T, L>(h ? la : lb, ninf, s4, ts);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    v[k] = k < M ? __uint_as_float(__float_as_uint(v[k]) | (uint32_t)(4 * h)) : kFar;
#pragma unroll
  for (int k = 0; k < M; ++k)
    o[k] = __shfl_xor(__uint_as_float(__float_as_uint(s4[k]) | (uint32_t)(4 * h)), 32, 64);
  // dropped values of either half are at least t (bit 2 of a packed value
  // is free before the half is added, so adding it keeps the order)
  t = __builtin_amdgcn_fmed3f(
      ninf, __uint_as_float(__float_as_uint(t) | (uint32_t)(4 * h)),
      __shfl_xor(__uint_as_float(__float_as_uint(ts) | (uint32_t)(4 * h)), 32, 64));
#pragma unroll
  for (int k = 0; k < M; ++k) top4_push(o[k], v, ninf);
  float b1 = v[0];
  float b2 = T >= 2 ? v[1] : __builtin_amdgcn_fmed3f(ninf, v[1], t);
  float b3 = T >= 3 ? v[2] : __builtin_amdgcn_fmed3f(ninf, v[2], t);
  float b4 = T >= 4 ? v[3] : __builtin_amdgcn_fmed3f(ninf, v[3], t);
  const int i1 = (int)(__float_as_uint(v[0]) & P.idx_mask);
  const int i2 = (int)(__float_as_uint(v[1]) & P.idx_mask);
  const int i3 = (int)(__float_as_uint(v[2]) & P.idx_mask);
  // single-product forms: ||x||^2 is added after the ranking
  b1 += x2;
  b2 += x2;
  b3 += x2;
  b4 += x2;
  bool flag2 = false, flagA = false, few3 = false;
  if (valid) {
    const float xn = sqrtf(x2);
    float E = P.alpha * xn * cmax + Ecm + P.xpk * x2;
    if (P.hop > 0.0f) E += P.hop * (cdmax * xn + 2.0f * cmax * xd + cdmax * xd);
    const float twoE = 2.0f * E;
    const float b1p = fmaxf(b1 + E, 0.0f);
    auto clear = [&](float bb) {
      return (bb >= kFarTest) || (bb - b1 > twoE + P.gref * (fmaxf(bb + E, 0.0f) + b1p));
    };
    const bool uniq = clear(b2);
    const bool two = clear(b3);
    const bool three = clear(b4);
    const bool sane = (b1 < kFarTest) && (i1 < P.K);
    P.labels[prow] = sane ? i1 : 0;
    flag2 = sane && !uniq && (two || three);
    flagA = !sane || (!uniq && !two && !three);
    few3 = !two;
  }
  const uint64_t m2 = __ballot(flag2);
  const uint64_t mA = __ballot(flagA);
  if (m2) {
    const int leader = __ffsll((unsigned long long)m2) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(q2c_s, (unsigned)__popcll(m2));
    base = __shfl(base, leader, 64);
    if (flag2) {
      const unsigned pos = base + mbcnt64(m2);
      *reinterpret_cast<int4*>(q2 + 4 * (size_t)pos) =
          make_int4((int)prow, i1, i2, few3 ? i3 : -1);
    }
  }
  if (mA) {
    const int leader = __ffsll((unsigned long long)mA) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(qac_s, (unsigned)__popcll(mA));
    base = __shfl(base, leader, 64);
    if (flagA) qa[base + mbcnt64(mA)] = (int)prow;
  }
}

// Work split: the centroids are resident in LDS (one copy per workgroup);
// workgroup g takes the contiguous 32-point units [U g / G, U (g+1) / G) and
// splits them into contiguous runs, one per wave. A wave sweeps two units
// (64 points) at a time: each A-operand fragment and each block's ||c||^2
// row read from LDS feed both units' MFMA chains, which halves the LDS
// traffic per distance against one unit per wave, and the chains of block s
// are issued before the epilogue of block s-1 (two accumulator pairs
// alternate). The epilogue keeps the T smallest packed values of each of L
// lists per lane (1 + T VALU operations per value; band_out2 turns the lists
// into bounds: the two smallest distances stay exact for T >= 2, the third
// and fourth are exact unless one list held three of the four smallest).
// The A fragments are read from LDS step by step, so the point fragments of
// both units, two accumulator pairs and one A step fit the register budget
// of two or three waves per SIMD.
template <int DP, int PREC, int NW, int T, bool PF, bool PIPE, int WPE, int L>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_assign2(AssignParams P) {
  static_assert(PREC == PREC_BF16 || PREC == PREC_FP8, "single-product forms only");
  // the stop flag is read beside the first rows and the centroid norms
  // (valid memory either way; nothing global is written before the test)
  const int stop = *P.done;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ float red_s[NW];
  __shared__ unsigned q2c_s, qac_s;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p = lane & 31;
  const int h = lane >> 5;
  if (threadIdx.x == 0) {
    q2c_s = 0u;
    qac_s = 0u;
  }
  int* const q2 = P.q2 + 4 * (size_t)blockIdx.x * P.qstride;
  int* const qa = P.qa + (size_t)blockIdx.x * P.qstride;

  const long long U = (P.n + 31) / 32;
  const long long g0 = U * blockIdx.x / gridDim.x;
  const long long g1 = U * (blockIdx.x + 1) / gridDim.x;
  const long long w0 = g0 + (g1 - g0) * wave / NW;
  const long long w1 = g0 + (g1 - g0) * (wave + 1) / NW;

  // The fragments of a pair of units; lane half h also carries the row,
  // norm and measured operand error of its set h's point (the band test of
  // set h runs on half h). With PF the next pair is requested a whole pair
  // ahead into a second set of registers.
  XFrag<DP, PREC> x0, x1, n0, n1;
  long long prow = 0, nrow = 0;
  bool pvalid = false, nvalid = false;
  float px2 = 0.0f, pxd = 0.0f, nx2 = 0.0f, nxd = 0.0f;
  auto load_pair = [&](long long uu, XFrag<DP, PREC>& f0, XFrag<DP, PREC>& f1, long long& row,
                       bool& valid, float& x2, float& xd) {
    const long long ra = uu * 32 + p;
    const long long rb = ra + 32;
    const bool va = ra < P.n;
    const bool vb = uu + 1 < w1 && rb < P.n;
    load_xfrag_any<DP, PREC>(P, ra, va, h, f0);
    load_xfrag_any<DP, PREC>(P, rb, vb, h, f1);
    row = h ? rb : ra;
    valid = h ? vb : va;
    x2 = valid ? P.xn2[row] : 0.0f;
    xd = (P.hop > 0.0f && valid) ? P.xdn[row] : 0.0f;
  };
  // the first pairs' rows are requested before the centroids are staged
  if (w0 < w1) load_pair(w0, x0, x1, prow, pvalid, px2, pxd);
  if (PF && w0 + 2 < w1) load_pair(w0 + 2, n0, n1, nrow, nvalid, nx2, nxd);

  // max ||c|| over the live operand rows and max measured operand error of
  // the centroid rows (for the error bound)
  float cm = 0.0f, cd = 0.0f;
  for (int c = threadIdx.x; c < P.Kp; c += blockDim.x) {
    const float v = P.cn[c];
    if (c < P.K && v < kFarTest) cm = fmaxf(cm, v);  // excluded clusters carry kFar
    if (P.hop > 0.0f) cd = fmaxf(cd, P.cdn[c]);
  }
  cm = wave_max_f32(cm);
  cd = wave_max_f32(cd);
  if (lane == 0) red_s[wave] = cm;
  __syncthreads();
  float cmax = 0.0f;
#pragma unroll
  for (int w = 0; w < NW; ++w) cmax = fmaxf(cmax, red_s[w]);
  cmax = sqrtf(cmax);
  __syncthreads();  // red_s reused
  if (lane == 0) red_s[wave] = cd;
  stage_centroids<DP, PREC>(P, smem, 0, P.Kp);
  __syncthreads();
  if (stop) return;
  float cdmax = 0.0f;
#pragma unroll
  for (int w = 0; w < NW; ++w) cdmax = fmaxf(cdmax, red_s[w]);
  const float Ecm = P.beta * cmax * cmax;
  const uint32_t keepv = opaque_vgpr(~P.idx_mask);
  const float ninf = __uint_as_float(opaque_vgpr(0xff800000u));

  const float* cn_s = reinterpret_cast<const float*>(smem);
  const unsigned char* rows = smem + (size_t)P.Kp * 4;
  const int nblk = P.Kp / 32;  // even: Kp is a multiple of 64

  for (long long u = w0; u < w1; u += 2) {
    float ba[L][T], bb[L][T];
#pragma unroll
    for (int j = 0; j < L; ++j)
#pragma unroll
      for (int k = 0; k < T; ++k) {
        ba[j][k] = kFar;
        bb[j][k] = kFar;
      }
    // Block step: the block's LDS reads are issued first; the first half of
    // the previous block's epilogue (set 0's values) runs while they are in
    // flight; then the block's MFMA chains, interleaved with the second half
    // (set 1's values). Scheduling barriers keep the compiler from moving the
    // reads next to their use or the epilogue behind the wait for them.
    auto step = [&](int blk, kmc_f32x16& c0, kmc_f32x16& c1, const kmc_f32x16& p0,
                    const kmc_f32x16& p1) {
      ABlock<DP, PREC> ab;
      load_ablock<DP, PREC>(rows, cn_s, 32 * blk, p, h, ab);
      __builtin_amdgcn_sched_barrier(0);
      if (blk > 0) push_blockL<T, L>(p0, (uint32_t)(32 * (blk - 1)), keepv, ba, ninf);
      __builtin_amdgcn_sched_barrier(0);
      mfma_chain2<DP, PREC>(ab, x0, x1, c0, c1);
      if (blk > 0) push_blockL<T, L>(p1, (uint32_t)(32 * (blk - 1)), keepv, bb, ninf);
    };
    kmc_f32x16 pA0, pA1, pB0, pB1;
    int blk = 1;
    if constexpr (PIPE) {
      step(0, pA0, pA1, pB0, pB1);
      for (; blk + 1 < nblk; blk += 2) {
        step(blk, pB0, pB1, pA0, pA1);
        step(blk + 1, pA0, pA1, pB0, pB1);
      }
      step(blk, pB0, pB1, pA0, pA1);  // blk == nblk - 1: the last use of this pair's fragments
    } else {
      // register-lean form: each block's values are ranked right after its
      // chains; the other waves of the SIMD fill the gaps
      for (blk = 0; blk + 1 < nblk; ++blk) {
        ABlock<DP, PREC> ab;
        load_ablock<DP, PREC>(rows, cn_s, 32 * blk, p, h, ab);
        mfma_chain2<DP, PREC>(ab, x0, x1, pA0, pA1);
        push_blockL<T, L>(pA0, (uint32_t)(32 * blk), keepv, ba, ninf);
        push_blockL<T, L>(pA1, (uint32_t)(32 * blk), keepv, bb, ninf);
      }
      ABlock<DP, PREC> ab;  // blk == nblk - 1: the last use of this pair's fragments
      load_ablock<DP, PREC>(rows, cn_s, 32 * blk, p, h, ab);
      mfma_chain2<DP, PREC>(ab, x0, x1, pB0, pB1);
    }
    const long long crow = prow;
    const bool cvalid = pvalid;
    const float cx2 = px2, cxd = pxd;
    if constexpr (PF) {
      // the pair after next is requested now; the next one is already in
      // flight or here
      x0 = n0;
      x1 = n1;
      prow = nrow;
      pvalid = nvalid;
      px2 = nx2;
      pxd = nxd;
      if (u + 4 < w1) load_pair(u + 4, n0, n1, nrow, nvalid, nx2, nxd);
    } else {
      // the next pair's rows are requested now, so that their memory latency
      // passes while the last epilogues and the band test of this pair run
      if (u + 2 < w1) load_pair(u + 2, x0, x1, prow, pvalid, px2, pxd);
    }
    push_blockL<T, L>(pB0, (uint32_t)(32 * blk), keepv, ba, ninf);
    push_blockL<T, L>(pB1, (uint32_t)(32 * blk), keepv, bb, ninf);
    band_out2<T, L>(P, ba, bb, cx2, cxd, cvalid, crow, cmax, Ecm, cdmax, ninf, h, lane, q2, qa,
                 &q2c_s, &qac_s);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    P.q2cnt[blockIdx.x] = q2c_s;
    P.qacnt[blockIdx.x] = qac_s;
  }
}

// ---------------------------------------------------------------------------
// Refinement: reference float arithmetic, first-min rule.
//
// The assign workgroups leave their band points in per-workgroup queue
// regions. Each refinement block first forms the exclusive prefix of the
// region counts in LDS; entry t of the flattened queue then lives in the
// region r with pre[r] <= t < pre[r+1], so the work spreads evenly over the
// grid however the band points cluster in memory.
// ---------------------------------------------------------------------------
// Exclusive prefix of one value per thread over a 256-thread workgroup
// (shuffle scan inside each wave, the four wave totals through LDS: one
// barrier); *total receives the sum. Every thread of the workgroup calls it.
__device__ __forceinline__ unsigned block_excl_scan256(unsigned s, unsigned* total) {
  __shared__ unsigned wsum_s[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned incl = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum_s[wave] = incl;
  __syncthreads();
  unsigned base = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) base += w < wave ? wsum_s[w] : 0u;
  *total = wsum_s[0] + wsum_s[1] + wsum_s[2] + wsum_s[3];
  return base + incl - s;
}

__device__ unsigned region_prefix(const unsigned* cnt, int nregions, unsigned* pre_s) {
  const int per = (nregions + 255) / 256;
  const int lo = min(nregions, (int)threadIdx.x * per), hi = min(nregions, lo + per);
  unsigned s = 0;
  for (int r = lo; r < hi; ++r) s += cnt[r];
  unsigned total;
  unsigned run = block_excl_scan256(s, &total);
  for (int r = lo; r < hi; ++r) {
    pre_s[r] = run;
    run += cnt[r];
  }
  if (threadIdx.x == 0) pre_s[nregions] = total;
  __syncthreads();
  return total;
}

// region_prefix of cnt, and in the same pass the total of cnt2 (*total2)
__device__ unsigned region_prefix2(const unsigned* cnt, const unsigned* cnt2, int nregions,
                                   unsigned* pre_s, unsigned* total2) {
  __shared__ unsigned t2_s[4];
  const int per = (nregions + 255) / 256;
  const int lo = min(nregions, (int)threadIdx.x * per), hi = min(nregions, lo + per);
  unsigned s = 0, s2 = 0;
  for (int r = lo; r < hi; ++r) {
    s += cnt[r];
    s2 += cnt2[r];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o, 64);
  if ((threadIdx.x & 63) == 0) t2_s[threadIdx.x >> 6] = s2;
  unsigned total;
  unsigned run = block_excl_scan256(s, &total);
  for (int r = lo; r < hi; ++r) {
    pre_s[r] = run;
    run += cnt[r];
  }
  if (threadIdx.x == 0) pre_s[nregions] = total;
  __syncthreads();
  *total2 = t2_s[0] + t2_s[1] + t2_s[2] + t2_s[3];
  return total;
}

// the pruned all-k kernel (k_refine_prune) is launched, and k_refine_all
// leaves queues of prune_min or more entries to it, exactly when this holds
__host__ __device__ inline bool prune_active(const RefineParams& P) {
  return !P.bp && P.cn && P.xn2 && P.Phi && P.Plo && P.prune_min > 0;
}
// the fused refinement kernel (k_refine_fused) replaces k_refine2 and
// k_refine_all; the shapes are those whose all-k path fits 256 registers
// per lane without scratch
__host__ __device__ inline bool pq_active(const RefineParams& P) {
  return !P.bp && P.cn && P.xn2 && P.Phi && P.Plo && P.pq && pq_shape(P.K, P.DP);
}

// region holding flattened entry t (pre is nondecreasing, pre[0] = 0)
__device__ __forceinline__ int region_of(const unsigned* pre_s, int nregions, unsigned t) {
  int lo = 0, hi = nregions - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre_s[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Band points with two or three candidates (the rest provably lose): each
// candidate's distance in the reference's float arithmetic, first-min rule
// (the smallest distance, the lowest index among equal ones).
//
// A wave takes EPW queue entries (lane e < EPW owns entry e). Dimensions go
// in chunks of CH (32, or 16 when D is padded to 16): CH/4 lanes load one
// entry's slice of the point row and of each candidate row (16-byte loads,
// the next chunk's loads issued before the current chunk is summed), form
// the squared terms t*t with t = x - c (one rounding each, as in the
// reference), and leave them in LDS; lane e then adds its entry's terms
// strictly in dimension order. Padding dimensions are zero in X and C and add
// exact zeros. An entry with two candidates repeats the second.
constexpr int kR2Stride = 20;  // floats per LDS row of squared terms (b128 reads conflict-free)
constexpr int kRefinePts = 8;  // points per all-k group
constexpr size_t kRefinePreBytes = ((size_t)(kMaxRegions + 1) * 4 + 15) / 16 * 16;
constexpr size_t kRefineLds =
    kRefinePreBytes + (size_t)4 * 3 * 64 * kR2Stride * 4;  // two/three-candidate pass
static_assert(kRefinePts * kMaxDP * 4 + 48 * kRefinePts <= 4 * 3 * 64 * kR2Stride * 4,
              "all-k pass needs less LDS than the two/three-candidate pass");
template <int CH>
__device__ void refine2_body(const RefineParams& P, unsigned char* smem, unsigned bid,
                             unsigned nblk, unsigned known_total = ~0u) {
  constexpr int LPE = CH / 4;     // lanes per entry in the load phase (16 B each)
  constexpr int EPW = 256 / LPE;  // entries per wave (each lane loads 4 rows)
  constexpr int RS = CH + 4;      // LDS row stride in floats (b128 reads conflict-free)
  static_assert(3 * EPW * RS <= 3 * 64 * kR2Stride, "LDS slice of a wave");
  unsigned* const pre_s = reinterpret_cast<unsigned*>(smem);
  float* const sq_all = reinterpret_cast<float*>(smem + kRefinePreBytes);
  // known_total: the caller has left the queue's region prefix in pre_s
  const unsigned total =
      known_total != ~0u ? known_total : region_prefix(P.q2cnt, P.nregions, pre_s);
  if (total == 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* const sq = sq_all + wave * (3 * 64 * kR2Stride);
  const int sub = lane % LPE;  // which 16-byte piece of a CH-dimension slice
  const unsigned nbatch = (total + EPW - 1) / EPW;
  for (unsigned bt = bid * 4u + wave; bt < nbatch; bt += nblk * 4u) {
    const unsigned t = min(bt * EPW + min(lane, EPW - 1), total - 1);
    const int r = region_of(pre_s, P.nregions, t);
    const int4 q = *reinterpret_cast<const int4*>(P.q2 + 4 * ((size_t)r * P.qstride +
                                                              (t - pre_s[r])));
    const int cand0 = q.y, cand1 = q.z, cand2 = q.w >= 0 ? q.w : q.z;
    // row pointers of the four entries this lane loads for
    // (e = lane/LPE + (64/LPE)*rr)
    const float* xp[4];
    const float* cp[4][3];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int e = lane / LPE + (64 / LPE) * rr;
      xp[rr] = P.X + (size_t)__shfl(q.x, e, 64) * P.DP + 4 * sub;
      cp[rr][0] = P.C + (size_t)__shfl(cand0, e, 64) * P.DP + 4 * sub;
      cp[rr][1] = P.C + (size_t)__shfl(cand1, e, 64) * P.DP + 4 * sub;
      cp[rr][2] = P.C + (size_t)__shfl(cand2, e, 64) * P.DP + 4 * sub;
    }
    float4 xv[4], cv[4][3];
    auto load_chunk = [&](int j0) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        xv[rr] = *reinterpret_cast<const float4*>(xp[rr] + j0);
#pragma unroll
        for (int k = 0; k < 3; ++k) cv[rr][k] = *reinterpret_cast<const float4*>(cp[rr][k] + j0);
      }
    };
    float acc[3] = {0.0f, 0.0f, 0.0f};
    load_chunk(0);
    for (int j0 = 0; j0 < P.DP; j0 += CH) {
      // squared terms of this chunk into LDS
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int e = lane / LPE + (64 / LPE) * rr;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma clang fp contract(off)
          float4 sv;
          float tt;
          tt = xv[rr].x - cv[rr][k].x; sv.x = tt * tt;
          tt = xv[rr].y - cv[rr][k].y; sv.y = tt * tt;
          tt = xv[rr].z - cv[rr][k].z; sv.z = tt * tt;
          tt = xv[rr].w - cv[rr][k].w; sv.w = tt * tt;
          *reinterpret_cast<float4*>(sq + (k * EPW + e) * RS + 4 * sub) = sv;
        }
      }
      if (j0 + CH < P.DP) load_chunk(j0 + CH);
      __builtin_amdgcn_wave_barrier();
      // this lane's entry: its terms in dimension order
      if (lane < EPW) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float* row = sq + (k * EPW + lane) * RS;
#pragma unroll
          for (int v4 = 0; v4 < CH / 4; ++v4) {
#pragma clang fp contract(off)
            const float4 sv = *reinterpret_cast<const float4*>(row + 4 * v4);
            acc[k] = acc[k] + sv.x;
            acc[k] = acc[k] + sv.y;
            acc[k] = acc[k] + sv.z;
            acc[k] = acc[k] + sv.w;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    int best = cand0;
    float rbest = acc[0];
    if (acc[1] < rbest || (acc[1] == rbest && cand1 < best)) { rbest = acc[1]; best = cand1; }
    if (acc[2] < rbest || (acc[2] == rbest && cand2 < best)) { rbest = acc[2]; best = cand2; }
    if (lane < EPW && bt * EPW + lane < total) P.labels[q.x] = best;
  }
}

// refine2_body at a compile-time width with the latency hidden deeper, for
// the fused kernel (whose register budget is 256 per lane): the dimension
// chunks are requested two ahead, and the next batch's queue entry while the
// current batch's first chunks are in flight. The queue's region prefix is
// in the LDS; same arithmetic and order as refine2_body.
template <int CH, int DP>
__device__ void refine2_deep(const RefineParams& P, unsigned char* smem, unsigned bid,
                             unsigned nblk, unsigned total) {
  constexpr int LPE = CH / 4;
  constexpr int EPW = 256 / LPE;
  constexpr int RS = CH + 4;
  constexpr int NCH = DP / CH;
  static_assert(3 * EPW * RS <= 3 * 64 * kR2Stride, "LDS slice of a wave");
  static_assert(DP % CH == 0, "whole chunks");
  unsigned* const pre_s = reinterpret_cast<unsigned*>(smem);
  float* const sq_all = reinterpret_cast<float*>(smem + kRefinePreBytes);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* const sq = sq_all + wave * (3 * 64 * kR2Stride);
  const int sub = lane % LPE;
  const unsigned nbatch = (total + EPW - 1) / EPW;
  const unsigned stride = nblk * 4u;
  auto entry = [&](unsigned b) {
    const unsigned t = min(b * EPW + min(lane, EPW - 1), total - 1);
    const int r = region_of(pre_s, P.nregions, t);
    return *reinterpret_cast<const int4*>(P.q2 + 4 * ((size_t)r * P.qstride + (t - pre_s[r])));
  };
  unsigned bt = bid * 4u + wave;
  int4 qn = make_int4(0, 0, 0, -1);
  if (bt < nbatch) qn = entry(bt);
  for (; bt < nbatch; bt += stride) {
    const int4 q = qn;
    const int cand0 = q.y, cand1 = q.z, cand2 = q.w >= 0 ? q.w : q.z;
    const float* xp[4];
    const float* cp[4][3];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int e = lane / LPE + (64 / LPE) * rr;
      xp[rr] = P.X + (size_t)__shfl(q.x, e, 64) * DP + 4 * sub;
      cp[rr][0] = P.C + (size_t)__shfl(cand0, e, 64) * DP + 4 * sub;
      cp[rr][1] = P.C + (size_t)__shfl(cand1, e, 64) * DP + 4 * sub;
      cp[rr][2] = P.C + (size_t)__shfl(cand2, e, 64) * DP + 4 * sub;
    }
    float4 xv[2][4], cv[2][4][3];
    auto load_chunk = [&](int b, int j0) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        xv[b][rr] = *reinterpret_cast<const float4*>(xp[rr] + j0);
#pragma unroll
        for (int k = 0; k < 3; ++k)
          cv[b][rr][k] = *reinterpret_cast<const float4*>(cp[rr][k] + j0);
      }
    };
    load_chunk(0, 0);
    if (NCH > 1) load_chunk(1, CH);
    if (bt + stride < nbatch) qn = entry(bt + stride);
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
      const int b = ch & 1;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int e = lane / LPE + (64 / LPE) * rr;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma clang fp contract(off)
          float4 sv;
          float tt;
          tt = xv[b][rr].x - cv[b][rr][k].x; sv.x = tt * tt;
          tt = xv[b][rr].y - cv[b][rr][k].y; sv.y = tt * tt;
          tt = xv[b][rr].z - cv[b][rr][k].z; sv.z = tt * tt;
          tt = xv[b][rr].w - cv[b][rr][k].w; sv.w = tt * tt;
          *reinterpret_cast<float4*>(sq + (k * EPW + e) * RS + 4 * sub) = sv;
        }
      }
      if (ch + 2 < NCH) load_chunk(b, (ch + 2) * CH);
      __builtin_amdgcn_wave_barrier();
      if (lane < EPW) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float* row = sq + (k * EPW + lane) * RS;
#pragma unroll
          for (int v4 = 0; v4 < CH / 4; ++v4) {
#pragma clang fp contract(off)
            const float4 sv = *reinterpret_cast<const float4*>(row + 4 * v4);
            acc[k] = acc[k] + sv.x;
            acc[k] = acc[k] + sv.y;
            acc[k] = acc[k] + sv.z;
            acc[k] = acc[k] + sv.w;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    int best = cand0;
    float rbest = acc[0];
    if (acc[1] < rbest || (acc[1] == rbest && cand1 < best)) { rbest = acc[1]; best = cand1; }
    if (acc[2] < rbest || (acc[2] == rbest && cand2 < best)) { rbest = acc[2]; best = cand2; }
    if (lane < EPW && bt * EPW + lane < total) P.labels[q.x] = best;
  }
}

// Band points whose four nearest MFMA candidates are all ambiguous: every
// centroid in the reference's float arithmetic. A workgroup takes
// kRefinePts consecutive queue entries and stages their rows in LDS; wave w
// evaluates centroids w*64 + lane + 256*i from the transposed centroids
// (coalesced across lanes, the next RC dimensions prefetched while the
// current ones are used), reading the point values as LDS broadcasts. Each
// distance advances in the reference's float order (padding dimensions are
// zero in X and C and add exact zeros). First-min across lanes, then across
// the waves.
template <int RC>
__device__ void refine_all_body(const RefineParams& P, unsigned char* smem, unsigned bid,
                                unsigned nblk) {
  unsigned* const pre_s = reinterpret_cast<unsigned*>(smem);
  float(*const x_s)[kMaxDP] = reinterpret_cast<float(*)[kMaxDP]>(smem + kRefinePreBytes);
  float(*const bv_s)[kRefinePts] =
      reinterpret_cast<float(*)[kRefinePts]>(smem + kRefinePreBytes + kRefinePts * kMaxDP * 4);
  int(*const bi_s)[kRefinePts] = reinterpret_cast<int(*)[kRefinePts]>(
      smem + kRefinePreBytes + kRefinePts * kMaxDP * 4 + 16 * kRefinePts);
  int* const pid_s = reinterpret_cast<int*>(smem + kRefinePreBytes + kRefinePts * kMaxDP * 4 +
                                            32 * kRefinePts);
  const unsigned total = region_prefix(P.qacnt, P.nregions, pre_s);
  // k_refine_prune takes queues of prune_min or more entries
  if (total == 0 || (prune_active(P) && total >= (unsigned)P.prune_min)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned ngroups = (total + kRefinePts - 1) / kRefinePts;
  for (unsigned g = bid; g < ngroups; g += nblk) {
    __syncthreads();  // the previous group is done with the LDS
    if (threadIdx.x < kRefinePts) {
      const unsigned t = min(g * kRefinePts + threadIdx.x, total - 1);
      const int r = region_of(pre_s, P.nregions, t);
      pid_s[threadIdx.x] = P.qa[(size_t)r * P.qstride + (t - pre_s[r])];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kRefinePts * P.DP; e += 256) {
      const int q = e / P.DP, j = e - q * P.DP;
      x_s[q][j] = P.X[(size_t)pid_s[q] * P.DP + j];
    }
    __syncthreads();
    float best[kRefinePts];
    int bi[kRefinePts];
    // candidate clusters of each point: all K, or its own problem's in a batch
    int clo[kRefinePts], chi[kRefinePts];
    int cfrom = 0x7fffffff, cto = 0;
#pragma unroll
    for (int q = 0; q < kRefinePts; ++q) {
      best[q] = __builtin_inff();
      bi[q] = 0x7fffffff;
      clo[q] = 0;
      chi[q] = P.K;
      if (P.bp) {
        const BatchProblem bpr = P.bp[P.blk_prob[pid_s[q] / kBatchAlign]];
        clo[q] = bpr.coff;
        chi[q] = bpr.coff + bpr.k;
      }
      cfrom = min(cfrom, clo[q]);
      cto = max(cto, chi[q]);
    }
    for (int c = cfrom + wave * 64 + lane; c - lane < cto; c += 256) {
      const int cc = min(c, P.Kp - 1);
      const float* ct = P.CT + cc;
      float acc[kRefinePts];
#pragma unroll
      for (int q = 0; q < kRefinePts; ++q) acc[q] = 0.0f;
      float cv[RC], cnx[RC];
#pragma unroll
      for (int jj = 0; jj < RC; ++jj) cv[jj] = ct[(size_t)jj * P.Kp];
      for (int j0 = 0; j0 < P.DP; j0 += RC) {
        if (j0 + RC < P.DP) {
#pragma unroll
          for (int jj = 0; jj < RC; ++jj) cnx[jj] = ct[(size_t)(j0 + RC + jj) * P.Kp];
        }
#pragma unroll
        for (int q = 0; q < kRefinePts; ++q) {
          float xv[RC];
#pragma unroll
          for (int jj = 0; jj < RC; jj += 4) {
            const float4 v = *reinterpret_cast<const float4*>(&x_s[q][j0 + jj]);
            xv[jj] = v.x;
            xv[jj + 1] = v.y;
            xv[jj + 2] = v.z;
            xv[jj + 3] = v.w;
          }
#pragma unroll
          for (int jj = 0; jj < RC; ++jj) {
#pragma clang fp contract(off)
            const float t = xv[jj] - cv[jj];
            const float sq = t * t;
            acc[q] = acc[q] + sq;
          }
        }
#pragma unroll
        for (int jj = 0; jj < RC; ++jj) cv[jj] = cnx[jj];
      }
      const bool ok = c < P.K && P.alive[min(c, P.K - 1)];
#pragma unroll
      for (int q = 0; q < kRefinePts; ++q)
        if (ok && c >= clo[q] && c < chi[q] && acc[q] < best[q]) {
          best[q] = acc[q];
          bi[q] = c;
        }
    }
#pragma unroll
    for (int q = 0; q < kRefinePts; ++q) {
      float b = best[q];
      int ix = bi[q];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(b, o, 64);
        const int oi = __shfl_xor(ix, o, 64);
        if (ov < b || (ov == b && oi < ix)) {
          b = ov;
          ix = oi;
        }
      }
      if (lane == 0) {
        bv_s[wave][q] = b;
        bi_s[wave][q] = ix;
      }
    }
    __syncthreads();
    if (threadIdx.x < kRefinePts) {
      const int q = threadIdx.x;
      float b = bv_s[0][q];
      int ix = bi_s[0][q];
#pragma unroll
      for (int w = 1; w < 4; ++w) {
        const float ov = bv_s[w][q];
        const int oi = bi_s[w][q];
        if (ov < b || (ov == b && oi < ix)) {
          b = ov;
          ix = oi;
        }
      }
      // no finite distance (NaN data): the first cluster of the point's range
      if (g * kRefinePts + q < total) P.labels[pid_s[q]] = (ix == 0x7fffffff) ? clo[q] : ix;
    }
  }
}

// The two refinement passes (they touch disjoint points). LDS: one dynamic
// buffer each (the region prefix, then the pass's own arrays).
constexpr size_t kRefineAllLds = kRefinePreBytes + kRefinePts * kMaxDP * 4 + 48 * kRefinePts;
__global__ __launch_bounds__(256) void k_refine2(RefineParams P) {
  if (*P.done) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char rsm[];
  if (P.DP % 32 == 0)
    refine2_body<32>(P, rsm, blockIdx.x, gridDim.x);
  else
    refine2_body<16>(P, rsm, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(256) void k_refine_all(RefineParams P) {
  if (*P.done) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char rsm[];
  refine_all_body<16>(P, rsm, blockIdx.x, gridDim.x);  // 32-dimension steps measured slower
}

// ---------------------------------------------------------------------------
// All-k refinement with MFMA pruning (one problem; batches use k_refine_all).
// Stage 1 forms the squared distance of each queued point to every centroid
// in the bf16x3 distance form -- x split into bf16 hi + lo on the fly, the
// prepared hi and lo planes of -2c (prep_row), products hi.hi + hi.lo +
// lo.hi, accumulator started at ||c||^2 + ||x||^2 -- whose error the bf16x3
// margin constants bound: E = alpha3 ||x|| cmax + beta3 cmax^2, with gref3
// covering the reference's own float error. A first sweep finds each point's
// smallest value b1, a second collects every centroid the band test cannot
// clear against b1, so the reference's first minimum is among the
// candidates. Stage 2 evaluates the candidates in the reference's float
// arithmetic and keeps the smallest distance, the lower index among equal
// ones. A point with more than kPruneCap candidates, or without a finite
// stage-1 value (NaN data, no live cluster), gets every live centroid
// evaluated exactly by the whole workgroup instead.
// A workgroup takes kPrunePts points, 32 per wave, and the centroid planes
// pass through LDS in chunks of KC rows shared by the four waves, the next
// chunk fetched into registers while the current one feeds the MFMAs. A
// group's two sweeps run on one CU, so the pass only pays once the queue
// fills the chip: shorter queues stay with k_refine_all.
// ---------------------------------------------------------------------------
constexpr int kPruneWaves = 4;
constexpr int kPrunePts = 32 * kPruneWaves;
constexpr int kPruneCap = 16;
constexpr int kPruneBlocks = 256;
template <int DP>
constexpr int prune_chunk() {
  return DP <= 64 ? 256 : (DP == 128 ? 128 : 32);
}
template <int DP>
constexpr size_t prune_lds() {
  using T = AssignTraits<DP, PREC_BF16X3>;
  return kRefinePreBytes + (size_t)prune_chunk<DP>() * 4 +
         (size_t)prune_chunk<DP>() * T::kRowBytes * 2 +
         (size_t)kPrunePts * (4 + 4 + 8 + 4 + 4 * kPruneCap);
}

__device__ __forceinline__ unsigned long long dist_key(float d, int c) {
  return ((unsigned long long)__float_as_uint(d) << 32) | (unsigned)c;
}
__device__ __forceinline__ unsigned long long min_key_xor(unsigned long long k, int o) {
  const unsigned lo = __shfl_xor((unsigned)k, o, 64);
  const unsigned hi = __shfl_xor((unsigned)(k >> 32), o, 64);
  const unsigned long long ok = ((unsigned long long)hi << 32) | lo;
  return ok < k ? ok : k;
}

// Squared distance in the reference's float arithmetic over the DP padded
// dimensions (padding is zero in X and C and adds exact zeros): in index
// order, one rounding for each subtraction, product and addition; the next
// 16 dimensions (8 for DP = 256) of both rows are loaded before the current
// ones are summed.
template <int DP>
__device__ __forceinline__ float exact_dist_rows(const float* __restrict__ x,
                                                 const float* __restrict__ c) {
#pragma clang fp contract(off)
  constexpr int CH = DP >= 256 ? 8 : 16;
  constexpr int NV = CH / 4;
  float4 xa[NV], ca[NV], xb[NV], cb[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    xa[v] = *reinterpret_cast<const float4*>(x + 4 * v);
    ca[v] = *reinterpret_cast<const float4*>(c + 4 * v);
  }
  float acc = 0.0f;
#pragma unroll 1  // rolled: one chunk in use and one in flight, not every load hoisted
  for (int j0 = 0; j0 < DP; j0 += CH) {
    if (j0 + CH < DP) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        xb[v] = *reinterpret_cast<const float4*>(x + j0 + CH + 4 * v);
        cb[v] = *reinterpret_cast<const float4*>(c + j0 + CH + 4 * v);
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      float t, sq;
      t = xa[v].x - ca[v].x; sq = t * t; acc = acc + sq;
      t = xa[v].y - ca[v].y; sq = t * t; acc = acc + sq;
      t = xa[v].z - ca[v].z; sq = t * t; acc = acc + sq;
      t = xa[v].w - ca[v].w; sq = t * t; acc = acc + sq;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      xa[v] = xb[v];
      ca[v] = cb[v];
    }
  }
  return acc;
}

// x0, x1 = hi + lo, each half rounded to nearest even (the packed hardware
// conversion: for finite values the same as bf16_rne; a NaN row gets no
// finite MFMA value and its point every centroid evaluated exactly)
typedef __bf16 kmc_bf16x2v __attribute__((ext_vector_type(2)));
typedef float kmc_f32x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_pair(float x0, float x1, uint32_t& hw, uint32_t& lw) {
  const kmc_f32x2v xv = {x0, x1};
  hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(xv, kmc_bf16x2v));
  const kmc_f32x2v rv = {x0 - __uint_as_float(hw << 16), x1 - __uint_as_float(hw & 0xffff0000u)};
  lw = __builtin_bit_cast(uint32_t, __builtin_convertvector(rv, kmc_bf16x2v));
}

// B fragments of a point for the bf16x3 form: dims 16 s + 8 h + [0, 8) of
// the float row, x = hi + lo (zero for an invalid slot)
template <int DP>
__device__ __forceinline__ void split_x(const float* xr, bool valid, int h,
                                        XFrag<DP, PREC_BF16X3>& xf) {
#pragma unroll
  for (int s = 0; s < DP / 16; ++s) {
    float v[8];
    const float4 a = *reinterpret_cast<const float4*>(xr + 16 * s + 8 * h);
    const float4 b = *reinterpret_cast<const float4*>(xr + 16 * s + 8 * h + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    uint32_t hw[4], lw[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      split_pair(valid ? v[2 * e] : 0.0f, valid ? v[2 * e + 1] : 0.0f, hw[e], lw[e]);
    xf.hi[s] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    xf.lo[s] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
  }
}

// split_x on a row already in registers (xv[2 s], xv[2 s + 1]: dims
// 16 s + 8 h + [0, 8))
template <int DP>
__device__ __forceinline__ void split_x_regs(const float4* xv, bool valid,
                                             XFrag<DP, PREC_BF16X3>& xf) {
#pragma unroll
  for (int s = 0; s < DP / 16; ++s) {
    const float4 a = xv[2 * s], b = xv[2 * s + 1];
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t hw[4], lw[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      split_pair(valid ? v[2 * e] : 0.0f, valid ? v[2 * e + 1] : 0.0f, hw[e], lw[e]);
    xf.hi[s] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    xf.lo[s] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
  }
}

// bf16x3 chain of one 32-centroid block of the LDS chunk against a wave's
// 32 points (the order of mfma_chain_lds): A fragments are read four k-steps
// at a time, with a scheduling barrier between groups, so that no more than
// four steps' fragments are live beside the register-resident points.
template <int DP>
__device__ __forceinline__ kmc_f32x16 prune_chain(const unsigned char* rows, int cb, int p, int h,
                                                  const XFrag<DP, PREC_BF16X3>& x,
                                                  kmc_f32x16 acc) {
  using T = AssignTraits<DP, PREC_BF16X3>;
  constexpr int KC = prune_chunk<DP>();
  constexpr int NS = DP / 16;
  constexpr int G = NS < 4 ? NS : 4;
  const uint16_t* ah = reinterpret_cast<const uint16_t*>(rows + (size_t)(cb + p) * T::kRowBytes);
  const uint16_t* al = reinterpret_cast<const uint16_t*>(rows + (size_t)KC * T::kRowBytes +
                                                         (size_t)(cb + p) * T::kRowBytes);
#pragma unroll
  for (int s0 = 0; s0 < NS; s0 += G) {
    uint4 hv[G], lv[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
      hv[i] = *reinterpret_cast<const uint4*>(ah + 16 * (s0 + i) + 8 * h);
      lv[i] = *reinterpret_cast<const uint4*>(al + 16 * (s0 + i) + 8 * h);
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int st = s0 + i;
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, hv[i]),
                                                    __builtin_bit_cast(kmc_bf16x8, x.hi[st]),
                                                    acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, hv[i]),
                                                    __builtin_bit_cast(kmc_bf16x8, x.lo[st]),
                                                    acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, lv[i]),
                                                    __builtin_bit_cast(kmc_bf16x8, x.hi[st]),
                                                    acc, 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  return acc;
}

template <int DP>
__global__ __launch_bounds__(256) void k_refine_prune(RefineParams P) {
  if (*P.done) return;
  using T = AssignTraits<DP, PREC_BF16X3>;
  constexpr int KC = prune_chunk<DP>();
  extern __shared__ __attribute__((aligned(16))) unsigned char rsm[];
  unsigned* const pre_s = reinterpret_cast<unsigned*>(rsm);
  float* const cn_s = reinterpret_cast<float*>(rsm + kRefinePreBytes);
  unsigned char* const rows = rsm + kRefinePreBytes + (size_t)KC * 4;
  unsigned long long* const key_s =
      reinterpret_cast<unsigned long long*>(rows + (size_t)KC * T::kRowBytes * 2);
  int* const cnt_s = reinterpret_cast<int*>(key_s + kPrunePts);
  int* const pid_s = cnt_s + kPrunePts;
  int* const ovf_s = pid_s + kPrunePts;
  int* const cand_s = ovf_s + kPrunePts;
  __shared__ float red_s[kPruneWaves];

  const unsigned total = region_prefix(P.qacnt, P.nregions, pre_s);
  // shorter queues go to k_refine_all: a group's sweeps are long, so the
  // pruned pass pays only once its groups fill the chip
  if (total < (unsigned)max(P.prune_min, 1)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = lane & 31, h = lane >> 5;
  const float kInf = __builtin_inff();

  // max ||c|| over the live centroids (excluded ones carry kFar)
  float cm = 0.0f;
  for (int c = threadIdx.x; c < P.K; c += 256) {
    const float v = P.cn[c];
    if (v < kFarTest) cm = fmaxf(cm, v);
  }
  cm = wave_max_f32(cm);
  if (lane == 0) red_s[wave] = cm;
  __syncthreads();
  float cmax = 0.0f;
#pragma unroll
  for (int w = 0; w < kPruneWaves; ++w) cmax = fmaxf(cmax, red_s[w]);
  cmax = sqrtf(cmax);
  const float Ecm = P.beta3 * cmax * cmax;

  const unsigned ngroups = (total + kPrunePts - 1) / kPrunePts;
  const int nchunks = (P.Kp + KC - 1) / KC;
  // chunk rows [c0, c0 + KC) of the prepared hi and lo planes of -2c, in
  // 16-byte pieces: thread t holds piece t + 256 i of the chunk
  constexpr int kPieces = KC * (DP / 8);  // per plane
  constexpr int kFill = 2 * kPieces / 256;
  static_assert((2 * kPieces) % 256 == 0, "chunk fill splits evenly");
  uint4 pf[kFill];
  auto fetch_chunk = [&](int c0) {
#pragma unroll
    for (int i = 0; i < kFill; ++i) {
      const int v = threadIdx.x + 256 * i;
      const int plane = v / kPieces, u = v - plane * kPieces;
      const int r = u / (DP / 8), q = u - r * (DP / 8);
      const int c = c0 + r;
      const uint16_t* src = plane ? P.Plo : P.Phi;
      pf[i] = c < P.Kp ? *reinterpret_cast<const uint4*>(src + (size_t)c * DP + 8 * q)
                       : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto store_chunk = [&](int c0) {
#pragma unroll
    for (int i = 0; i < kFill; ++i) {
      const int v = threadIdx.x + 256 * i;
      const int plane = v / kPieces, u = v - plane * kPieces;
      const int r = u / (DP / 8), q = u - r * (DP / 8);
      *reinterpret_cast<uint4*>(rows + (size_t)plane * KC * T::kRowBytes +
                                (size_t)r * T::kRowBytes + 16 * q) = pf[i];
    }
    for (int r = threadIdx.x; r < KC; r += 256) cn_s[r] = c0 + r < P.Kp ? P.cn[c0 + r] : kFar;
  };
  for (unsigned g = blockIdx.x; g < ngroups; g += gridDim.x) {
    __syncthreads();  // the previous group is done with the LDS
    if (threadIdx.x < kPrunePts) {
      const unsigned t = g * kPrunePts + threadIdx.x;
      int pid = -1;
      if (t < total) {
        const int r = region_of(pre_s, P.nregions, t);
        pid = P.qa[(size_t)r * P.qstride + (t - pre_s[r])];
      }
      pid_s[threadIdx.x] = pid;
      cnt_s[threadIdx.x] = 0;
      key_s[threadIdx.x] = ~0ull;
    }
    __syncthreads();
    const int slot = wave * 32 + p;
    const int pid = pid_s[slot];
    const bool valid = pid >= 0;
    const float* const xr = P.X + (size_t)(valid ? pid : 0) * DP;
    XFrag<DP, PREC_BF16X3> xf;
    split_x<DP>(xr, valid, h, xf);
    const float x2 = valid ? P.xn2[pid] : 0.0f;

    float b1 = kInf, E = 0.0f, twoE = 0.0f, b1p = 0.0f;
    // the chunk sequence of the two sweeps (a single chunk is loaded once and
    // stays resident); the next chunk's rows are fetched into registers
    // while the current one feeds the MFMAs
    const int nfill = nchunks == 1 ? 1 : 2 * nchunks;
    fetch_chunk(0);
    int fill = 0;
    for (int pass = 0; pass < 2; ++pass) {
      for (int ch = 0; ch < nchunks; ++ch) {
        const int c0 = ch * KC;
        if (pass == 0 || nchunks > 1) {
          __syncthreads();  // the previous chunk's MFMAs are done with the LDS
          store_chunk(c0);
          __syncthreads();
          ++fill;
          if (fill < nfill) fetch_chunk((fill % nchunks) * KC);
        }
#pragma unroll 1
        for (int cb = 0; cb < KC && c0 + cb < P.Kp; cb += 32) {
          kmc_f32x16 acc = init_acc<true>(cn_s, cb, h, x2);
          acc = prune_chain<DP>(rows, cb, p, h, xf, acc);
          if (pass == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) b1 = fminf(b1, acc[i]);
          } else {
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) {
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float v = acc[4 * gq + j];
                const int c = c0 + cb + 8 * gq + 4 * h + j;
                const bool cand = valid && c < P.K && v < kFarTest &&
                                  !(v - b1 > twoE + P.gref3 * (fmaxf(v + E, 0.0f) + b1p));
                if (cand) {
                  const int pos = atomicAdd(&cnt_s[slot], 1);
                  if (pos < kPruneCap) cand_s[slot * kPruneCap + pos] = c;
                }
              }
            }
          }
        }
      }
      if (pass == 0) {
        b1 = fminf(b1, __shfl_xor(b1, 32, 64));
        E = P.alpha3 * sqrtf(x2) * cmax + Ecm;
        twoE = 2.0f * E;
        b1p = fmaxf(b1 + E, 0.0f);
      }
    }
    __syncthreads();

    // stage 2: lanes p and p + 32 take the even and odd candidates
    const int nc = cnt_s[slot];
    const bool ovf = valid && (nc > kPruneCap || !(b1 < kFarTest));
    unsigned long long best = ~0ull;
    if (valid && !ovf) {
#pragma unroll 1
      for (int i = h; i < nc; i += 2) {
        const int c = cand_s[slot * kPruneCap + i];
        const float d = exact_dist_rows<DP>(xr, P.C + (size_t)c * DP);
        if (d < kInf) {
          const unsigned long long k = dist_key(d, c);
          best = k < best ? k : best;
        }
      }
    }
    best = min_key_xor(best, 32);
    if (h == 0) {
      ovf_s[slot] = ovf ? 1 : 0;
      if (valid && !ovf) P.labels[pid] = best == ~0ull ? 0 : (int)(unsigned)best;
    }
    __syncthreads();

    // points the pruning could not narrow: every live centroid, exactly
    for (int q = 0; q < kPrunePts; ++q) {
      if (!ovf_s[q]) continue;
      const int qp = pid_s[q];
      const float* const xq = P.X + (size_t)qp * DP;
      unsigned long long bq = ~0ull;
#pragma unroll 1
      for (int c = threadIdx.x; c < P.K; c += 256) {
        if (!P.alive[c]) continue;
        const float d = exact_dist_rows<DP>(xq, P.C + (size_t)c * DP);
        if (d < kInf) {
          const unsigned long long k = dist_key(d, c);
          bq = k < bq ? k : bq;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) bq = min_key_xor(bq, o);
      if (lane == 0) atomicMin(&key_s[q], bq);
      __syncthreads();
      if (threadIdx.x == 0) P.labels[qp] = key_s[q] == ~0ull ? 0 : (int)(unsigned)key_s[q];
    }
  }
}

// ---------------------------------------------------------------------------
// All-k refinement of queues too short for k_refine_prune (one problem,
// pq_shape): the same bf16x3 candidate test and exact stage 2, shaped for
// latency. A workgroup takes 32 points, the four waves split the centroids
// (NB 32-centroid blocks each) and keep their MFMA values in registers, so
// one sweep serves both the minimum b1 (exchanged in LDS) and the band test.
// Stage 2 gives each point eight lanes for its candidates; a point with more
// than kPqCap candidates or without a finite value gets every live centroid
// evaluated exactly by the whole workgroup.
// ---------------------------------------------------------------------------
constexpr int kPqCap = 16;

// A fragments of G steps (from s0) of the 32-centroid block at c0, straight
// from the prepared planes in global memory
template <int DP>
struct PqA {
  static constexpr int NS = DP / 16;
  static constexpr int G = NS < 8 ? NS : 8;
  uint4 hv[G], lv[G];
};
template <int DP>
__device__ __forceinline__ void pq_load(const uint16_t* Phi, const uint16_t* Plo, int c0, int p,
                                        int h, int s0, PqA<DP>& a) {
  const uint16_t* ah = Phi + (size_t)(c0 + p) * DP;
  const uint16_t* al = Plo + (size_t)(c0 + p) * DP;
#pragma unroll
  for (int i = 0; i < PqA<DP>::G; ++i) {
    a.hv[i] = *reinterpret_cast<const uint4*>(ah + 16 * (s0 + i) + 8 * h);
    a.lv[i] = *reinterpret_cast<const uint4*>(al + 16 * (s0 + i) + 8 * h);
  }
}
template <int DP>
__device__ __forceinline__ kmc_f32x16 pq_mfma(const PqA<DP>& a, int s0,
                                              const XFrag<DP, PREC_BF16X3>& x, kmc_f32x16 acc) {
#pragma unroll
  for (int i = 0; i < PqA<DP>::G; ++i) {
    const int st = s0 + i;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a.hv[i]),
                                                  __builtin_bit_cast(kmc_bf16x8, x.hi[st]),
                                                  acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a.hv[i]),
                                                  __builtin_bit_cast(kmc_bf16x8, x.lo[st]),
                                                  acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a.lv[i]),
                                                  __builtin_bit_cast(kmc_bf16x8, x.hi[st]),
                                                  acc, 0, 0, 0);
  }
  return acc;
}
// bf16x3 chain of the block at c0 (the order of prune_chain), its first G
// steps already requested in a0; later groups are requested whole before
// their MFMAs (the scheduling barrier keeps the compiler from interleaving
// loads and waits, which would pay one memory latency per step)
template <int DP>
__device__ __forceinline__ kmc_f32x16 pq_chain(const uint16_t* Phi, const uint16_t* Plo, int c0,
                                               int p, int h, const XFrag<DP, PREC_BF16X3>& x,
                                               kmc_f32x16 acc, const PqA<DP>& a0) {
  acc = pq_mfma<DP>(a0, 0, x, acc);
#pragma unroll
  for (int s0 = PqA<DP>::G; s0 < PqA<DP>::NS; s0 += PqA<DP>::G) {
    PqA<DP> a;
    pq_load<DP>(Phi, Plo, c0, p, h, s0, a);
    __builtin_amdgcn_sched_barrier(0);
    acc = pq_mfma<DP>(a, s0, x, acc);
  }
  return acc;
}

// exact_dist_rows with the point row in LDS: the centroid row is requested
// 128 dimensions at a time (64 at DP = 256), all before the first is used
template <int DP>
__device__ __forceinline__ float exact_dist_lds(const float* xs, const float* __restrict__ c) {
#pragma clang fp contract(off)
  constexpr int CH = DP <= 128 ? DP : 64;
  float acc = 0.0f;
#pragma unroll
  for (int j0 = 0; j0 < DP; j0 += CH) {
    float4 cv[CH / 4];
#pragma unroll
    for (int v = 0; v < CH / 4; ++v) cv[v] = *reinterpret_cast<const float4*>(c + j0 + 4 * v);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int v = 0; v < CH / 4; ++v) {
      const float4 xv = *reinterpret_cast<const float4*>(xs + j0 + 4 * v);
      float t, sq;
      t = xv.x - cv[v].x; sq = t * t; acc = acc + sq;
      t = xv.y - cv[v].y; sq = t * t; acc = acc + sq;
      t = xv.z - cv[v].z; sq = t * t; acc = acc + sq;
      t = xv.w - cv[v].w; sq = t * t; acc = acc + sq;
    }
  }
  return acc;
}

// the all-k groups' LDS arrays, after the staged point rows
struct PqShared {
  float b1_s[4][32];
  int pid_s[32], cnt_s[32], cand_s[32 * kPqCap];
  unsigned long long key_s[32];
  float red_s[4];
  unsigned ovf_s;
};
template <int DP>
constexpr size_t pq_lds() {
  return kRefinePreBytes + (size_t)32 * (DP + 4) * sizeof(float) + sizeof(PqShared);
}

// The all-k groups g = bid, bid + nblk, ... of a queue of `total` points
// whose region prefix is in the LDS (kEarly: the first centroid block's
// planes are requested before the points' rows, when the registers allow).
template <int DP, int NB, bool kEarly>
__device__ void pq_body(const RefineParams& P, unsigned char* rsm, unsigned bid, unsigned nblk,
                        unsigned total) {
  constexpr int XS = DP + 4;  // LDS row stride of the staged point rows (floats)
  unsigned* const pre_s = reinterpret_cast<unsigned*>(rsm);
  float* const x_s = reinterpret_cast<float*>(rsm + kRefinePreBytes);
  PqShared& sh = *reinterpret_cast<PqShared*>(rsm + kRefinePreBytes +
                                              (size_t)32 * XS * sizeof(float));
  float(&b1_s)[4][32] = sh.b1_s;
  int(&pid_s)[32] = sh.pid_s;
  int(&cnt_s)[32] = sh.cnt_s;
  int(&cand_s)[32 * kPqCap] = sh.cand_s;
  unsigned long long(&key_s)[32] = sh.key_s;
  float(&red_s)[4] = sh.red_s;
  unsigned& ovf_s = sh.ovf_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = lane & 31, h = lane >> 5;
  const float kInf = __builtin_inff();

  constexpr int NCN = (4 * 32 * NB + 255) / 256;
  float cnv[NCN];
#pragma unroll
  for (int i = 0; i < NCN; ++i) {
    const int c = threadIdx.x + 256 * i;
    cnv[i] = c < P.K ? P.cn[c] : 0.0f;
  }
  // max ||c|| over the live centroids (excluded ones carry kFar)
  float cm = 0.0f;
#pragma unroll
  for (int i = 0; i < NCN; ++i)
    if (cnv[i] < kFarTest) cm = fmaxf(cm, cnv[i]);
  cm = wave_max_f32(cm);
  if (lane == 0) red_s[wave] = cm;
  __syncthreads();
  const float cmax = sqrtf(fmaxf(fmaxf(red_s[0], red_s[1]), fmaxf(red_s[2], red_s[3])));
  const float Ecm = P.beta3 * cmax * cmax;
  const int cw = 32 * NB * wave;  // this wave's first centroid

  const unsigned ngroups = (total + 31) / 32;
  for (unsigned g = bid; g < ngroups; g += nblk) {
    __syncthreads();  // the previous group is done with the LDS
    if (threadIdx.x < 32) {
      const unsigned t = g * 32 + threadIdx.x;
      int pid = -1;
      if (t < total) {
        const int r = region_of(pre_s, P.nregions, t);
        pid = P.qa[(size_t)r * P.qstride + (t - pre_s[r])];
      }
      pid_s[threadIdx.x] = pid;
      cnt_s[threadIdx.x] = 0;
      key_s[threadIdx.x] = ~0ull;
    }
    if (threadIdx.x == 0) ovf_s = 0u;
    // the first block's planes do not depend on the points
    PqA<DP> a[2];
    if (kEarly && cw < P.Kp) pq_load<DP>(P.Phi, P.Plo, cw, p, h, 0, a[0]);
    __syncthreads();
    const int pid = pid_s[p];
    const bool valid = pid >= 0;
    // the point's row (this lane's half of each 16-dimension step) is
    // requested in full before any of it is split into bf16 hi + lo
    const float* const xr = P.X + (size_t)(valid ? pid : 0) * DP;
    float4 xv[DP / 8];
#pragma unroll
    for (int st = 0; st < DP / 16; ++st) {
      xv[2 * st] = *reinterpret_cast<const float4*>(xr + 16 * st + 8 * h);
      xv[2 * st + 1] = *reinterpret_cast<const float4*>(xr + 16 * st + 8 * h + 4);
    }
    const float x2 = valid ? P.xn2[pid] : 0.0f;
    __builtin_amdgcn_sched_barrier(0);
    XFrag<DP, PREC_BF16X3> xf;
    split_x_regs<DP>(xv, valid, xf);
    if (wave == 0) {  // the rows for stage 2
#pragma unroll
      for (int st = 0; st < DP / 16; ++st) {
        *reinterpret_cast<float4*>(x_s + p * XS + 16 * st + 8 * h) = xv[2 * st];
        *reinterpret_cast<float4*>(x_s + p * XS + 16 * st + 8 * h + 4) = xv[2 * st + 1];
      }
    }

    // stage 1: this wave's centroid blocks, values kept for the band test;
    // the next block's first planes are requested before this block's MFMAs
    kmc_f32x16 acc[NB];
    float bl = kInf;
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      const int c0 = cw + 32 * cb;
      if (c0 < P.Kp) {
        if (!kEarly) {
          pq_load<DP>(P.Phi, P.Plo, c0, p, h, 0, a[cb & 1]);
          __builtin_amdgcn_sched_barrier(0);
        } else if (cb + 1 < NB && c0 + 32 < P.Kp) {
          pq_load<DP>(P.Phi, P.Plo, c0 + 32, p, h, 0, a[(cb + 1) & 1]);
          __builtin_amdgcn_sched_barrier(0);
        }
        acc[cb] = pq_chain<DP>(P.Phi, P.Plo, c0, p, h, xf, init_acc<true>(P.cn, c0, h, x2),
                               a[cb & 1]);
#pragma unroll
        for (int i = 0; i < 16; ++i) bl = fminf(bl, acc[cb][i]);
      }
    }
    bl = fminf(bl, __shfl_xor(bl, 32, 64));
    if (h == 0) b1_s[wave][p] = bl;
    __syncthreads();
    const float b1 = fminf(fminf(b1_s[0][p], b1_s[1][p]), fminf(b1_s[2][p], b1_s[3][p]));
    const float E = P.alpha3 * sqrtf(x2) * cmax + Ecm;
    const float twoE = 2.0f * E;
    const float b1p = fmaxf(b1 + E, 0.0f);
    // candidates as a bit set (bit 16 cb + 4 gq + j), one LDS atomic per lane
    uint64_t bits = 0;
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      const int c0 = cw + 32 * cb;
      if (c0 < P.Kp) {
#pragma unroll
        for (int gq = 0; gq < 4; ++gq) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float v = acc[cb][4 * gq + j];
            const int c = c0 + 8 * gq + 4 * h + j;
            const bool cand = valid && c < P.K && v < kFarTest &&
                              !(v - b1 > twoE + P.gref3 * (fmaxf(v + E, 0.0f) + b1p));
            if (cand) bits |= 1ull << (16 * cb + 4 * gq + j);
          }
        }
      }
    }
    if (bits) {
      int pos = atomicAdd(&cnt_s[p], __popcll(bits));
      while (bits) {
        const int b = __ffsll((long long)bits) - 1;
        bits &= bits - 1;
        const int r = b & 15;
        if (pos < kPqCap) cand_s[p * kPqCap + pos] = cw + 32 * (b >> 4) + 8 * (r >> 2) + 4 * h + (r & 3);
        ++pos;
      }
    }
    __syncthreads();

    // stage 2: lanes 8 q + [0, 8) take point q's candidates
    {
      const int q = threadIdx.x >> 3, sub = threadIdx.x & 7;
      const int qp = pid_s[q];
      const int nc = cnt_s[q];
      const float bq1 = fminf(fminf(b1_s[0][q], b1_s[1][q]), fminf(b1_s[2][q], b1_s[3][q]));
      const bool ovf = qp >= 0 && (nc > kPqCap || !(bq1 < kFarTest));
      if (ovf && sub == 0) atomicOr(&ovf_s, 1u << q);
      unsigned long long best = ~0ull;
      if (qp >= 0 && !ovf) {
#pragma unroll 1
        for (int i = sub; i < nc; i += 8) {
          const int c = cand_s[q * kPqCap + i];
          const float d = exact_dist_lds<DP>(x_s + q * XS, P.C + (size_t)c * DP);
          if (d < kInf) {
            const unsigned long long k = dist_key(d, c);
            best = k < best ? k : best;
          }
        }
      }
      best = min_key_xor(best, 4);
      best = min_key_xor(best, 2);
      best = min_key_xor(best, 1);
      if (sub == 0 && qp >= 0 && !ovf) P.labels[qp] = best == ~0ull ? 0 : (int)(unsigned)best;
    }
    __syncthreads();

    // points the pruning could not narrow: every live centroid, exactly
    unsigned m = ovf_s;
    while (m) {
      const int q = __ffs(m) - 1;
      m &= m - 1;
      const int qp = pid_s[q];
      unsigned long long bq = ~0ull;
#pragma unroll 1
      for (int c = threadIdx.x; c < P.K; c += 256) {
        if (!P.alive[c]) continue;
        const float d = exact_dist_lds<DP>(x_s + q * XS, P.C + (size_t)c * DP);
        if (d < kInf) {
          const unsigned long long k = dist_key(d, c);
          bq = k < bq ? k : bq;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) bq = min_key_xor(bq, o);
      if (lane == 0) atomicMin(&key_s[q], bq);
      __syncthreads();
      if (threadIdx.x == 0) P.labels[qp] = key_s[q] == ~0ull ? 0 : (int)(unsigned)key_s[q];
    }
  }
}

// Both refinement passes in one launch (one problem, pq_shape): the first G
// workgroups take the all-k groups (G: the group count up to a quarter of
// the grid, for a longer queue the share of the grid its estimated work
// asks for; none when the all-k queue goes to k_refine_prune), the others
// the two/three-candidate batches, so that the
// all-k chain runs beside the batches instead of after them. The grid is the
// workgroups resident at once, two per CU for the two/three-candidate LDS;
// the all-k path fits that budget without its early plane requests.
template <int DP, int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_refine_fused(RefineParams P) {
  // the stop flag is read beside the queue counts of the region prefix (valid
  // memory either way; the prefix only writes LDS), so its latency overlaps
  const int stop = *P.done;
  extern __shared__ __attribute__((aligned(16))) unsigned char rsm[];
  unsigned* const pre_s = reinterpret_cast<unsigned*>(rsm);
  unsigned ta = 0;
  const unsigned t2 = region_prefix2(P.q2cnt, P.qacnt, P.nregions, pre_s, &ta);
  if (stop) return;
  const unsigned ngA = (ta + 31) / 32;
  unsigned G = 0;
  if (ta > 0 && !(prune_active(P) && ta >= (unsigned)P.prune_min)) {
    constexpr unsigned EPW2 = 256 / ((DP % 32 == 0 ? 32 : 16) / 4);
    const unsigned nb2 = (t2 + EPW2 - 1) / EPW2;  // two/three-candidate wave batches
    G = min(ngA, gridDim.x / 4);
    if (ngA > G) {
      // a long all-k queue: the grid is split by estimated work, an all-k
      // group counting as eight wave batches (measured best of 3, 5, 8, 12
      // and 20 on fp8's 100 k-point queue at N=1M, D=128, K=256)
      const unsigned long long wA = 8ull * ngA;
      G = max(G, (unsigned)((unsigned long long)gridDim.x * wA / (wA + nb2)));
      G = min(G, nb2 > 0 ? gridDim.x - 1 : gridDim.x);
    }
  }
  if (blockIdx.x < G) {
    __syncthreads();  // every thread has read the two/three-candidate total
    unsigned dummy;
    const unsigned tq = region_prefix2(P.qacnt, P.qacnt, P.nregions, pre_s, &dummy);
    pq_body<DP, NB, false>(P, rsm, blockIdx.x, G, tq);
  } else if (t2 > 0) {
    refine2_deep<(DP % 32 == 0) ? 32 : 16, DP>(P, rsm, blockIdx.x - G, gridDim.x - G, t2);
  }
}
static_assert(pq_lds<kMaxDP>() <= kRefineLds, "the all-k groups fit the fused LDS");

template <int DP, int NB>
hipError_t launch_fused1(const RefineParams& P, hipStream_t s) {
  static int grid = 0;
  if (!grid) {
    const void* fn = reinterpret_cast<const void*>(k_refine_fused<DP, NB>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)kRefineLds);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, kRefineLds) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    grid = std::min(kRefineBlocks, per_cu * cus);
  }
  hipLaunchKernelGGL((k_refine_fused<DP, NB>), dim3(grid), dim3(256), kRefineLds, s, P);
  return hipGetLastError();
}
template <int DP>
hipError_t launch_fused(const RefineParams& P, hipStream_t s) {
  if (!pq_shape(P.K, DP)) return hipErrorInvalidValue;
  if (P.K <= 128) return launch_fused1<DP, 1>(P, s);
  if (P.K <= 256) return launch_fused1<DP, 2>(P, s);
  if constexpr (DP <= 64) return launch_fused1<DP, 4>(P, s);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stable counting sort of point ids by label. One wave per block of CH points
// (CH a multiple of 1024); every lane loads its 16 labels of a 1024-point
// round before any is used, so one memory latency covers the round.
// bcount is cluster-major: bcount[c * nblk + b].
// ---------------------------------------------------------------------------
constexpr int kSortRound = 1024;
__global__ __launch_bounds__(64) void k_hist(SortParams P) {
  if (*P.done) return;
  extern __shared__ int cnt_s[];
  const int b = blockIdx.x;
  for (int c = threadIdx.x; c < P.K; c += 64) cnt_s[c] = 0;
  __syncthreads();
  const long long lo = (long long)b * P.CH;
  const long long hi = min(lo + P.CH, (long long)P.n);
  for (long long i0 = lo; i0 < hi; i0 += kSortRound) {
    int lab[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const long long i = i0 + u * 64 + threadIdx.x;
      lab[u] = i < hi ? P.labels[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (lab[u] >= 0) atomicAdd(&cnt_s[lab[u]], 1);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < P.K; c += 64) P.bcount[(size_t)c * P.nblk + b] = cnt_s[c];
}

// Exclusive scan over blocks for each cluster (one wave per cluster); each
// lane takes 8 consecutive block counts per round, all loaded first.
__global__ __launch_bounds__(256) void k_scan_blocks(SortParams P) {
  if (*P.done) return;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= P.K) return;
  int* a = P.bcount + (size_t)c * P.nblk;
  int carry = 0;
  for (int base = 0; base < P.nblk; base += 512) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = base + lane * 8 + u;
      v[u] = idx < P.nblk ? a[idx] : 0;
    }
    int tot = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = v[u];
      v[u] = tot;
      tot += t;
    }
    int incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    const int off = carry + incl - tot;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = base + lane * 8 + u;
      if (idx < P.nblk) a[idx] = off + v[u];
    }
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) P.count[c] = carry;
}

// Scatter point ids to their sorted slot; within a block the points are
// visited in index order, 64 at a time. The lanes holding the same label are
// found with one ballot per label bit (a lane keeps the lanes that agree with
// it on every bit), ranked by lane order, and the lowest lane of each group
// advances the group's cursor, so the sort is stable.
__global__ __launch_bounds__(64) void k_scatter(SortParams P) {
  if (*P.done) return;
  extern __shared__ int cur_s[];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  // start[] = exclusive scan of the cluster sizes: every block forms it
  // (lane l scans a contiguous run of clusters), block 0 publishes it
  {
    const int S = (P.K + 63) / 64;
    const int c0 = min(P.K, lane * S), c1 = min(P.K, c0 + S);
    int tot = 0;
    for (int c = c0; c < c1; c += 4) {
      int v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = c + u < c1 ? P.count[c + u] : 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) tot += v[u];
    }
    int incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    int run = incl - tot;
    for (int c = c0; c < c1; c += 4) {
      int n_c[4], bc[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        n_c[u] = c + u < c1 ? P.count[c + u] : 0;
        bc[u] = c + u < c1 ? P.bcount[(size_t)(c + u) * P.nblk + b] : 0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (c + u < c1) {
          if (b == 0) P.start[c + u] = run;
          cur_s[c + u] = run + bc[u];
        }
        run += n_c[u];
      }
    }
  }
  __syncthreads();
  const long long lo = (long long)b * P.CH;
  const long long hi = min(lo + P.CH, (long long)P.n);
  const uint64_t lt = lanemask_lt();
  for (long long r0 = lo; r0 < hi; r0 += kSortRound) {
    int labs[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const long long i = r0 + u * 64 + lane;
      labs[u] = i < hi ? P.labels[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const long long i = r0 + u * 64 + lane;
      const bool valid = labs[u] >= 0;
      const int lab = valid ? labs[u] : 0;
      uint64_t match = __ballot(valid);
      for (int bit = 0; bit < P.lbits; ++bit) {
        const bool on = (lab >> bit) & 1;
        const uint64_t B = __ballot(on);
        match &= on ? B : ~B;
      }
      const int rank = __popcll(match & lt);
      const int pos = valid ? cur_s[lab] + rank : 0;
      __builtin_amdgcn_wave_barrier();
      if (valid && rank == 0) cur_s[lab] = pos + __popcll(match);
      if (valid) P.members[pos] = (int)i;
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// ---------------------------------------------------------------------------
// k_update: one wave per (cluster, 64-dim slab). Float sums over the cluster's
// members in index order (bit-identical to the reference CPU update), scaled
// by (float)(1.0 / (double)max(count, 1)); FP64 partial inertia against the
// centroids the labels were assigned to.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_update(UpdateParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int S = (P.D + 63) / 64;
  if (gw >= P.K * S) return;
  const int c = gw / S;
  const int j = (gw - c * S) * 64 + lane;
  const bool act = j < P.D;
  // wave-uniform cluster extent (scalar registers)
  const int cnt = __builtin_amdgcn_readfirstlane(P.count[c]);
  const int* mem = P.members + __builtin_amdgcn_readfirstlane(P.start[c]);
  const float cold = act ? P.C[(size_t)c * P.DP + j] : 0.0f;
  const float* X = P.X + (act ? j : 0);
  const size_t DP = (size_t)P.DP;
  float sum = 0.0f;
  double in = 0.0;
  auto take = [&](float x) {
#pragma clang fp contract(off)
    sum = sum + x;
    const float df = x - cold;
    in = fma((double)df, (double)df, in);
  };
  // Rolling pipeline over full 64-row batches: the indices of a batch arrive
  // in one vector load (lane u holds the u-th) two batches ahead, each row
  // address is a readlane, and the load of row t+64 is issued right after
  // row t is consumed, so ~64 row loads stay in flight while the float sum
  // advances strictly in member order.
  float xv[64];
  int t = 0;
  if (cnt >= 128) {
    const int icur = mem[lane];
    int inext = mem[64 + lane];
#pragma unroll
    for (int u = 0; u < 64; ++u) xv[u] = X[(size_t)__builtin_amdgcn_readlane(icur, u) * DP];
    for (; t + 192 <= cnt; t += 64) {
      const int inn = mem[t + 128 + lane];
#pragma unroll
      for (int u = 0; u < 64; ++u) {
        take(xv[u]);
        xv[u] = X[(size_t)__builtin_amdgcn_readlane(inext, u) * DP];
      }
      inext = inn;
    }
#pragma unroll
    for (int u = 0; u < 64; ++u) {
      take(xv[u]);
      xv[u] = X[(size_t)__builtin_amdgcn_readlane(inext, u) * DP];
    }
#pragma unroll
    for (int u = 0; u < 64; ++u) take(xv[u]);
    t += 128;
  }
  // remaining rows (< 192 when the pipeline ran, < 128 otherwise): batches
  // of 64 loads issued together, then consumed in order
  while (t < cnt) {
    const int m = min(64, cnt - t);
    const int iv = (lane < m) ? mem[t + lane] : 0;
#pragma unroll
    for (int u = 0; u < 64; ++u)
      if (u < m) xv[u] = X[(size_t)__builtin_amdgcn_readlane(iv, u) * DP];
#pragma unroll
    for (int u = 0; u < 64; ++u)
      if (u < m) take(xv[u]);
    t += m;
  }
  if (act) {
    const float scale = (float)(1.0 / (double)(cnt > 1 ? cnt : 1));
    P.Cnext[(size_t)c * P.DP + j] = sum * scale;
  } else {
    in = 0.0;
  }
  in = wave_sum_f64(in);
  if (lane == 0) P.ipart[gw] = in;
}

// ---------------------------------------------------------------------------
// Tree update: the sorted member list is cut into chunks of kTreeRows rows,
// one wave per chunk. The wave first stages the chunk's member ids and their
// labels in LDS (coalesced loads, two memory latencies per chunk), then walks
// the chunk segment by segment (a segment = the part of one cluster inside
// the chunk, found by a ballot over the staged labels) and sums each
// segment's rows in FP64, with every row load of a batch issued before any
// is used. One partial per segment goes to index cluster + chunk (injective,
// because clusters occupy increasing chunk ranges); k_finish adds a
// cluster's partials in chunk order. Deterministic; the FP64 sums are
// rounded once, so centroids are at least as accurate as the reference's
// float sums (not bit-identical).
// ---------------------------------------------------------------------------
constexpr int kTreeWaves = 4;

// the moved points of an incremental iteration (defined with k_update_fx)
__device__ void update_moved_points(const UpdateParams& P);

// stage member ids and labels of local rows [0, L) of a chunk; the exact
// form also keeps each point's label in prev and returns how many of the
// lane's staged points changed cluster since the last update
template <bool EX>
__device__ __forceinline__ unsigned tree_stage(const UpdateParams& P, long long p0, int L,
                                               int* mem_s, int* lab_s, int lane) {
  int id[kTreeRows / 64], lab[kTreeRows / 64];
#pragma unroll
  for (int m = 0; m < kTreeRows / 64; ++m) {
    const int i = m * 64 + lane;
    id[m] = i < L ? P.members[p0 + i] : 0;
  }
#pragma unroll
  for (int m = 0; m < kTreeRows / 64; ++m) {
    const int i = m * 64 + lane;
    lab[m] = i < L ? P.labels[id[m]] : -1;
  }
  unsigned moved = 0;
  if constexpr (EX) {
    int old[kTreeRows / 64];
#pragma unroll
    for (int m = 0; m < kTreeRows / 64; ++m) old[m] = m * 64 + lane < L ? P.prev[id[m]] : 0;
#pragma unroll
    for (int m = 0; m < kTreeRows / 64; ++m) {
      if (m * 64 + lane < L) {
        moved += old[m] != lab[m] ? 1u : 0u;
        P.prev[id[m]] = lab[m];
      }
    }
  }
#pragma unroll
  for (int m = 0; m < kTreeRows / 64; ++m) {
    mem_s[m * 64 + lane] = id[m];
    lab_s[m * 64 + lane] = lab[m];
  }
  return moved;
}

// first local index in (r, L) whose label differs from c, or L
__device__ __forceinline__ int tree_seg_end(const int* lab_s, int r, int L, int c, int lane) {
  for (int w = r + 1; w < L; w += 64) {
    const int i = w + lane;
    const uint64_t m = __ballot(i >= L || lab_s[i < L ? i : 0] != c);
    if (m) return w + __ffsll((unsigned long long)m) - 1;
  }
  return L;
}

// Wide rows (DP = 128 with VW = 2, DP = 256 with VW = 4): lane owns dimensions
// [VW*lane, VW*lane + VW) of every row, one contiguous 256*VW-byte read per
// row, batches of 16 rows. EX: the exact form (UpdateParams::cl_scale): a
// segment's sum is the 64-bit integer sum of round(x * scale), its cluster's
// fixed-point scale; the iterations that are not a full pass update the
// totals from the moved points only.
template <int VW, bool EX>
__global__ __launch_bounds__(64 * kTreeWaves) void k_update_tree(UpdateParams P) {
  if (*P.done) return;
  if (EX && !fx_full_pass(P.state, P.n)) {
    update_moved_points(P);
    return;
  }
  __shared__ int mem_s[kTreeWaves][kTreeRows], lab_s[kTreeWaves][kTreeRows];
  __shared__ unsigned moved_s;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (EX && threadIdx.x == 0) moved_s = 0u;
  __syncthreads();
  const long long q = (long long)blockIdx.x * kTreeWaves + wv;
  const long long nchunks = (P.n + kTreeRows - 1) / kTreeRows;
  const long long p0 = q * kTreeRows;
  const int L = q < nchunks ? (int)min((long long)kTreeRows, P.n - p0) : 0;
  const unsigned mv = tree_stage<EX>(P, p0, L, mem_s[wv], lab_s[wv], lane);
  if (EX && mv) atomicAdd(&moved_s, mv);
  __syncthreads();
  if (EX && threadIdx.x == 0 && moved_s)
    atomicAdd(&P.state->moved_acc, (unsigned long long)moved_s);
  if (L == 0) return;
  const int* ms = mem_s[wv];
  const int j0 = VW * lane;
  const bool act = j0 < P.DP;
  const float* X = P.X + (act ? j0 : 0);
  double in = 0.0;
  int r = 0;
  while (r < L) {
    const int c = lab_s[wv][r];
    const int e = tree_seg_end(lab_s[wv], r, L, c, lane);
    double cold[VW], sum[VW];
    long long isum[VW];
    const float fs = EX ? P.cl_scale[c] : 0.0f;
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      cold[v] = act ? (double)P.C[(size_t)c * P.DP + j0 + v] : 0.0;
      sum[v] = 0.0;
      isum[v] = 0;
    }
    // batches of U rows; the last, partial batch of a segment loads a valid
    // row in place of the rows past the end and masks them out
    constexpr int U = 16;
    auto batch = [&](int t, bool partial) {
      float xv[U][VW];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int tt = partial ? min(t + u, e - 1) : t + u;
        const float* row = X + (size_t)ms[tt] * P.DP;
        if (VW == 4) {
          const float4 f = *reinterpret_cast<const float4*>(row);
          xv[u][0] = f.x; xv[u][1 % VW] = f.y; xv[u][2 % VW] = f.z; xv[u][3 % VW] = f.w;
        } else {
          const float2 f = *reinterpret_cast<const float2*>(row);
          xv[u][0] = f.x; xv[u][1 % VW] = f.y;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = !partial || t + u < e;  // wave-uniform
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          if constexpr (EX) {
            isum[v] += ok ? __float2ll_rn(xv[u][v] * fs) : 0ll;
          } else {
            const double xd = (double)xv[u][v];
            const double df = xd - cold[v];
            sum[v] = ok ? sum[v] + xd : sum[v];
            in = ok ? fma(df, df, in) : in;
          }
        }
      }
    };
    int t = r;
    for (; t + U <= e; t += U) batch(t, false);
    if (t < e) batch(t, true);
    if (act) {
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        const size_t o = (size_t)(c + q) * P.DP + j0 + v;
        if constexpr (EX) reinterpret_cast<long long*>(P.tpart)[o] = isum[v];
        else P.tpart[o] = sum[v];
      }
    }
    r = e;
  }
  if (!act) in = 0.0;
  in = wave_sum_f64(in);
  if (lane == 0) P.ipart[q] = in;
}

// Narrow rows (DP = 16, 32, 64): a lane takes four dimensions (one 16-byte
// load), so one wave load covers G = 256 / DP rows and a batch of U loads 64
// rows. Lane group g sums the rows t + u*G + g of a segment; at the segment's
// end the G group sums are added by a fixed butterfly, so the result is
// deterministic. EX as for k_update_tree.
template <int DP, bool EX>
__global__ __launch_bounds__(64 * kTreeWaves) void k_update_tree_rows(UpdateParams P) {
  if (*P.done) return;
  if (EX && !fx_full_pass(P.state, P.n)) {
    update_moved_points(P);
    return;
  }
  constexpr int LPR = DP / 4;    // lanes per row
  constexpr int G = 64 / LPR;    // rows per wave load
  constexpr int U = G >= 4 ? 64 / G : 16;  // loads per batch
  __shared__ int mem_s[kTreeWaves][kTreeRows], lab_s[kTreeWaves][kTreeRows];
  __shared__ unsigned moved_s;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (EX && threadIdx.x == 0) moved_s = 0u;
  __syncthreads();
  const long long q = (long long)blockIdx.x * kTreeWaves + wv;
  const long long nchunks = (P.n + kTreeRows - 1) / kTreeRows;
  const long long p0 = q * kTreeRows;
  const int L = q < nchunks ? (int)min((long long)kTreeRows, P.n - p0) : 0;
  const unsigned mv = tree_stage<EX>(P, p0, L, mem_s[wv], lab_s[wv], lane);
  if (EX && mv) atomicAdd(&moved_s, mv);
  __syncthreads();
  if (EX && threadIdx.x == 0 && moved_s)
    atomicAdd(&P.state->moved_acc, (unsigned long long)moved_s);
  if (L == 0) return;
  const int* ms = mem_s[wv];
  const int g = lane / LPR, j0 = 4 * (lane % LPR);
  const float* X = P.X + j0;
  double in = 0.0;
  int r = 0;
  while (r < L) {
    const int c = lab_s[wv][r];
    const int e = tree_seg_end(lab_s[wv], r, L, c, lane);
    const float4 cf = *reinterpret_cast<const float4*>(P.C + (size_t)c * DP + j0);
    const double cold[4] = {(double)cf.x, (double)cf.y, (double)cf.z, (double)cf.w};
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    long long isum[4] = {0, 0, 0, 0};
    const float fs = EX ? P.cl_scale[c] : 0.0f;
    auto batch = [&](int t, bool partial) {
      float4 xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int rr = t + u * G + g;
        const int tt = partial ? min(rr, e - 1) : rr;
        xv[u] = *reinterpret_cast<const float4*>(X + (size_t)ms[tt] * DP);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = !partial || t + u * G + g < e;
        const float xf[4] = {xv[u].x, xv[u].y, xv[u].z, xv[u].w};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if constexpr (EX) {
            isum[v] += ok ? __float2ll_rn(xf[v] * fs) : 0ll;
          } else {
            const double xd = (double)xf[v];
            const double df = xd - cold[v];
            sum[v] = ok ? sum[v] + xd : sum[v];
            in = ok ? fma(df, df, in) : in;
          }
        }
      }
    };
    int t = r;
    for (; t + U * G <= e; t += U * G) batch(t, false);
    if (t < e) batch(t, true);
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if constexpr (EX) isum[v] += __shfl_xor(isum[v], o, 64);
        else sum[v] += __shfl_xor(sum[v], o, 64);
      }
    }
    if (g == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const size_t o = (size_t)(c + q) * DP + j0 + v;
        if constexpr (EX) reinterpret_cast<long long*>(P.tpart)[o] = isum[v];
        else P.tpart[o] = sum[v];
      }
    }
    r = e;
  }
  in = wave_sum_f64(in);
  if (lane == 0) P.ipart[q] = in;
}

// ---------------------------------------------------------------------------
// k_update_fx (UPDATE_FIXED): one streaming pass over X in point order, no
// sort. Workgroup w of dimension slice s owns the points [w*per, (w+1)*per)
// and the DS dimensions [s*DS, (s+1)*DS); each value becomes the integer
// round(x * fscale) (fscale a power of two chosen from max|x| and n, so every
// |x| >= 2^-14 converts exactly) and is added to the cluster's LDS
// accumulator with a 64-bit integer atomic. Integer sums do not depend on
// the order of the additions, so the result is deterministic however the
// waves interleave. Slice-0 workgroups also count the members and add the
// fixed-point ||x||^2 (for the FP64 inertia). k_finish adds the workgroup
// partials exactly and rounds the means once.
// ---------------------------------------------------------------------------
constexpr int kFxThreads = 512;
// round(v) as a 64-bit integer (ties to even, as __float2ll_rn). With M
// (|v| < 2^50): v + 1.5 * 2^52 in double rounds v to an integer and leaves it
// in the low mantissa bits, so the bits of the sum less those of 1.5 * 2^52
// are round(v) -- three operations instead of the general conversion's
// eleven.
template <bool M>
__device__ __forceinline__ unsigned long long fx_int(float v) {
  if constexpr (M) {
    return (unsigned long long)__double_as_longlong((double)v + 0x1.8p52) -
           0x4338000000000000ull;
  } else {
    return (unsigned long long)__float2ll_rn(v);
  }
}
// LDS row stride of a cluster's accumulators: one extra word so that lanes
// of different points (different clusters) fall in different banks
__host__ __device__ constexpr int fx_stride(int ds) { return ds + 1; }
// moved-point form (FX_MOVED): labels a lane scans per step, and the LDS
// lists of the waves' moved points (point offset, new and old label)
constexpr int kFxScan = 4;
constexpr int kFxRound = kFxThreads * kFxScan;
size_t update_fx_lds(int Kp, int ds) {
  return (size_t)Kp * fx_stride(ds) * 8 + (size_t)Kp * 8 + (size_t)Kp * 4 +
         (size_t)kFxRound * sizeof(int2);
}

// Incremental form of the fixed-point update (the iterations that are not a
// full pass; run by k_update_fx's grid): every point whose label differs
// from the one the cluster totals stand for has its converted row (the same
// round(x * fscale) integers as the full pass) subtracted from the old
// cluster's totals and added to the new one's, with its fixed-point ||x||^2
// and count, by 64-bit integer atomics. Integer sums do not depend on the
// order, so the totals equal those a full pass would form. A wave scans 64
// points at a time and takes its moved points one by one, a lane per
// dimension.
__device__ void update_moved_points(const UpdateParams& P) {
  __shared__ unsigned moved_s;  // this workgroup's moved points (one global add)
  if (threadIdx.x == 0) moved_s = 0u;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  const long long w0 = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  // the totals this wave adds to: its delta copy, or the totals themselves
  const int copy = (int)(w0 % kDeltaCopies);
  long long* const St = P.Sdel ? P.Sdel + (size_t)copy * P.Kp * P.DP : P.Stot;
  long long* const Qt = P.Sdel ? (P.Qtot ? P.Qdel + (size_t)copy * P.Kp : nullptr) : P.Qtot;
  int* const Nt = P.Sdel ? P.Ndel + (size_t)copy * P.Kp : P.Ntot;
  // one moved point: its converted row out of the old cluster's totals and
  // into the new one's, a lane per dimension (both clusters are of the
  // point's problem: one scale)
  auto move_point = [&](long long pi, int nl, int ol) {
    const float fs = P.cl_scale ? P.cl_scale[nl] : P.fscale;
    const float* xr = P.X + (size_t)pi * P.DP;
    unsigned long long* sn = reinterpret_cast<unsigned long long*>(St + (size_t)nl * P.DP);
    unsigned long long* so = reinterpret_cast<unsigned long long*>(St + (size_t)ol * P.DP);
    // the row (a lane per dimension) and ||x||^2 are requested before any
    // atomic is issued: one memory latency per moved point
    constexpr int JT = kMaxDP / 64;
    float xv[JT];
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int j = lane + 64 * t;
      xv[t] = j < P.D ? xr[j] : 0.0f;
    }
    const double xq = (lane == 0 && Qt) ? P.xn64[pi] : 0.0;
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int j = lane + 64 * t;
      if (j < P.D) {
        const long long q = __float2ll_rn(xv[t] * fs);
        atomicAdd(sn + j, (unsigned long long)q);
        atomicAdd(so + j, (unsigned long long)(-q));
      }
    }
    if (lane == 0) {
      if (Qt) {
        const long long qq = __double2ll_rn(xq * P.qscale);
        atomicAdd(reinterpret_cast<unsigned long long*>(Qt + nl), (unsigned long long)qq);
        atomicAdd(reinterpret_cast<unsigned long long*>(Qt + ol), (unsigned long long)(-qq));
      }
      atomicAdd(Nt + nl, 1);
      atomicAdd(Nt + ol, -1);
    }
  };
  // a wave takes points base + u * nw * 64 + lane, u < UD: the labels and
  // previous labels of all UD groups are loaded before any is compared
  constexpr int UD = 8;
  for (long long base = w0 * 64; base < P.n; base += nw * 64 * UD) {
    int labv[UD], oldv[UD];
#pragma unroll
    for (int u = 0; u < UD; ++u) {
      const long long i = base + u * nw * 64 + lane;
      labv[u] = i < P.n ? P.labels[i] : 0;
      oldv[u] = i < P.n ? P.prev[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < UD; ++u) {
      const long long i = base + u * nw * 64 + lane;
      const bool mv = i < P.n && labv[u] != oldv[u];
      uint64_t m = __ballot(mv);
      if (!m) continue;
      if (lane == 0) atomicAdd(&moved_s, (unsigned)__popcll(m));
      if (mv) P.prev[i] = labv[u];
      while (m) {
        const int src = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        move_point(__shfl(i, src, 64), __shfl(labv[u], src, 64), __shfl(oldv[u], src, 64));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && moved_s) atomicAdd(&P.state->moved_acc, (unsigned long long)moved_s);
}

// The moved points only (FX_MOVED), run by a k_update_fx workgroup over its
// points [lo, hi) and its dimension slice: a point whose label differs from
// the one the totals stand for adds its converted slice to the new cluster's
// LDS accumulators and subtracts it from the old one's (on slice 0 its
// fixed-point ||x||^2 and count likewise), so the workgroup partials are
// changes of the totals. Each wave works on its own: a step scans 64 *
// kFxScan labels and previous labels (the next step's load while this one's
// rows are read), lists the moved points in the wave's LDS list, and reads
// their rows with every load of a batch issued before any is used; no
// workgroup barrier. prev is left as it is: the workgroups of every slice
// read it, and k_finish copies the labels into it.
template <int DS, bool M>
__device__ __forceinline__ void fx_moved_rows(const UpdateParams& P, unsigned long long* acc,
                                              unsigned long long* qacc, unsigned* cacc,
                                              int2* lists, int slice, long long lo, long long hi,
                                              unsigned& mv_lane) {
  constexpr int ST = fx_stride(DS);
  constexpr int LPP = DS / 4;    // lanes per point (a float4 each)
  constexpr int PPW = 64 / LPP;  // points per wave load
  constexpr int UM = 16;         // loads of a batch
  constexpr int NW = kFxThreads / 64;
  constexpr int WSTEP = 64 * kFxScan;  // points a wave scans per step
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int pp = lane / LPP, dg = lane % LPP;
  const int col = slice * DS + 4 * dg;
  const bool pt_lane = slice == 0 && dg == 0;
  const float fs = P.fscale;
  int2* const list = lists + wave * WSTEP;
  const unsigned long long below = (1ull << lane) - 1ull;
  int la[kFxScan], pa[kFxScan];
  auto load_labels = [&](long long s0) {
#pragma unroll
    for (int u = 0; u < kFxScan; ++u) {
      const long long i = s0 + u * 64 + lane;
      la[u] = i < hi ? P.labels[i] : 0;
      pa[u] = i < hi ? P.prev[i] : 0;  // equal past the range: not moved
    }
  };
  const long long wstride = (long long)NW * WSTEP;
  long long s0 = lo + (long long)wave * WSTEP;
  load_labels(s0);
  for (; s0 < hi; s0 += wstride) {
    // the step's moved points into the wave's list, ranked by ballot
    // (labels < 2^16: k is limited by the LDS)
    unsigned cnt = 0;
#pragma unroll
    for (int u = 0; u < kFxScan; ++u) {
      const bool mv = la[u] != pa[u];
      const unsigned long long m = __ballot(mv);
      if (mv)
        list[cnt + (unsigned)__popcll(m & below)] =
            make_int2((int)(s0 + u * 64 + lane - lo), (la[u] << 16) | pa[u]);
      cnt += (unsigned)__popcll(m);
    }
    if (slice == 0 && lane == 0) mv_lane += cnt;
    load_labels(s0 + wstride);
    __builtin_amdgcn_wave_barrier();
    for (unsigned e0 = 0; e0 < cnt; e0 += UM * PPW) {
      float4 v[UM];
      double q[UM];
      int ln[UM], lp[UM];
#pragma unroll
      for (int u = 0; u < UM; ++u) {
        const unsigned e = e0 + u * PPW + pp;
        v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        q[u] = 0.0;
        ln[u] = -1;
        lp[u] = 0;
        if (e < cnt) {
          const int2 t = list[e];
          const long long i = lo + t.x;
          ln[u] = t.y >> 16;
          lp[u] = t.y & 0xffff;
          v[u] = *reinterpret_cast<const float4*>(P.X + (size_t)i * P.DP + col);
          if (pt_lane) q[u] = P.xn64[i];
        }
      }
#pragma unroll
      for (int u = 0; u < UM; ++u) {
        if (ln[u] < 0) continue;
        unsigned long long* an = acc + (size_t)ln[u] * ST + dg;
        unsigned long long* ao = acc + (size_t)lp[u] * ST + dg;
        const unsigned long long x0 = fx_int<M>(v[u].x * fs), x1 = fx_int<M>(v[u].y * fs);
        const unsigned long long x2 = fx_int<M>(v[u].z * fs), x3 = fx_int<M>(v[u].w * fs);
        atomicAdd(an + 0 * LPP, x0);
        atomicAdd(an + 1 * LPP, x1);
        atomicAdd(an + 2 * LPP, x2);
        atomicAdd(an + 3 * LPP, x3);
        atomicAdd(ao + 0 * LPP, 0ull - x0);  // subtraction modulo 2^64
        atomicAdd(ao + 1 * LPP, 0ull - x1);
        atomicAdd(ao + 2 * LPP, 0ull - x2);
        atomicAdd(ao + 3 * LPP, 0ull - x3);
        if (pt_lane) {
          const long long qq = __double2ll_rn(q[u] * P.qscale);
          atomicAdd(&qacc[ln[u]], (unsigned long long)qq);
          atomicAdd(&qacc[lp[u]], (unsigned long long)(-qq));
          atomicAdd(&cacc[ln[u]], 1u);
          atomicAdd(&cacc[lp[u]], ~0u);  // -1 modulo 2^32
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // the list is rewritten by the next step
  }
}

// The moved points only by global atomics (FX_ATOMIC), run by k_update_fx's
// grid (its waves interleaved over the points 64 at a time, as
// update_moved_points): a wave lists a round's moved points (256 scanned) in
// its LDS list and takes them eight at a time, a lane per dimension, with
// the eight rows and norms requested before any atomic, so a wave with many
// moved points pays one memory latency per eight instead of one each.
template <bool M>
__device__ void fx_atomic_rows(const UpdateParams& P, int2* lists) {
  __shared__ unsigned moved_s;  // this workgroup's moved points (one global add)
  if (threadIdx.x == 0) moved_s = 0u;
  __syncthreads();
  constexpr int UD = kFxScan;  // groups of 64 points a wave scans per round
  constexpr int JT = kMaxDP / 64, UB = 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
  const long long w0 = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  // the totals this wave adds to: its delta copy, or the totals themselves
  const int copy = (int)(w0 % kDeltaCopies);
  long long* const St = P.Sdel ? P.Sdel + (size_t)copy * P.Kp * P.DP : P.Stot;
  long long* const Qt = P.Sdel ? (P.Qtot ? P.Qdel + (size_t)copy * P.Kp : nullptr) : P.Qtot;
  int* const Nt = P.Sdel ? P.Ndel + (size_t)copy * P.Kp : P.Ntot;
  int2* const list = lists + wave * (64 * UD);
  const unsigned long long below = (1ull << lane) - 1ull;
  const float fs = P.fscale;
  unsigned mv_lane = 0;
  for (long long base = w0 * 64; base < P.n; base += nw * 64 * UD) {
    int labv[UD], oldv[UD];
#pragma unroll
    for (int u = 0; u < UD; ++u) {
      const long long i = base + u * nw * 64 + lane;
      labv[u] = i < P.n ? P.labels[i] : 0;
      oldv[u] = i < P.n ? P.prev[i] : 0;
    }
    // the round's moved points into the wave's list, ranked by ballot
    // (labels < 2^16: k is limited by the LDS)
    unsigned cnt = 0;
#pragma unroll
    for (int u = 0; u < UD; ++u) {
      const long long i = base + u * nw * 64 + lane;
      const bool mv = i < P.n && labv[u] != oldv[u];
      const unsigned long long m = __ballot(mv);
      if (mv) {
        list[cnt + (unsigned)__popcll(m & below)] =
            make_int2(u * 64 + lane, (labv[u] << 16) | oldv[u]);
        P.prev[i] = labv[u];
      }
      cnt += (unsigned)__popcll(m);
    }
    if (lane == 0) mv_lane += cnt;
    __builtin_amdgcn_wave_barrier();
    for (unsigned e0 = 0; e0 < cnt; e0 += UB) {
      float xv[UB][JT];
      double xq[UB];
      int nl[UB], ol[UB];
#pragma unroll
      for (int k = 0; k < UB; ++k) {
        nl[k] = -1;
        ol[k] = 0;
        xq[k] = 0.0;
#pragma unroll
        for (int t = 0; t < JT; ++t) xv[k][t] = 0.0f;
        if (e0 + k < cnt) {
          const int2 en = list[e0 + k];
          const long long i = base + (en.x >> 6) * nw * 64 + (en.x & 63);
          nl[k] = en.y >> 16;
          ol[k] = en.y & 0xffff;
          const float* xr = P.X + (size_t)i * P.DP;
#pragma unroll
          for (int t = 0; t < JT; ++t) {
            const int j = lane + 64 * t;
            if (j < P.D) xv[k][t] = xr[j];
          }
          if (lane == 0 && Qt) xq[k] = P.xn64[i];
        }
      }
#pragma unroll
      for (int k = 0; k < UB; ++k) {
        if (nl[k] < 0) continue;
        unsigned long long* sn = reinterpret_cast<unsigned long long*>(St + (size_t)nl[k] * P.DP);
        unsigned long long* so = reinterpret_cast<unsigned long long*>(St + (size_t)ol[k] * P.DP);
#pragma unroll
        for (int t = 0; t < JT; ++t) {
          const int j = lane + 64 * t;
          if (j < P.D) {
            const unsigned long long q = fx_int<M>(xv[k][t] * fs);
            atomicAdd(sn + j, q);
            atomicAdd(so + j, 0ull - q);  // subtraction modulo 2^64
          }
        }
        if (lane == 0) {
          if (Qt) {
            const long long qq = __double2ll_rn(xq[k] * P.qscale);
            atomicAdd(reinterpret_cast<unsigned long long*>(Qt + nl[k]), (unsigned long long)qq);
            atomicAdd(reinterpret_cast<unsigned long long*>(Qt + ol[k]),
                      (unsigned long long)(-qq));
          }
          atomicAdd(Nt + nl[k], 1);
          atomicAdd(Nt + ol[k], -1);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // the list is rewritten by the next round
  }
  if (mv_lane) atomicAdd(&moved_s, mv_lane);
  __syncthreads();
  if (threadIdx.x == 0 && moved_s) atomicAdd(&P.state->moved_acc, (unsigned long long)moved_s);
}

template <int DS, bool M>
__global__ __launch_bounds__(kFxThreads) void k_update_fx(UpdateParams P) {
  // the stop flag and the state are read together (one memory latency); the
  // early return tests both, so neither load waits behind the other
  const int stop = *P.done;
  const int mode0 = P.Stot ? fx_mode(P.state, P.n) : FX_FULL;
  const int mode = stop ? -1 : mode0;
  if (mode < 0) return;
  constexpr int ST = fx_stride(DS);
  extern __shared__ __attribute__((aligned(16))) unsigned char fsm[];
  if (mode == FX_ATOMIC) {  // few moved points: global atomics
    fx_atomic_rows<M>(P, reinterpret_cast<int2*>(fsm + (size_t)P.Kp * ST * 8 +
                                                 (size_t)P.Kp * 8 + (size_t)P.Kp * 4));
    return;
  }
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(fsm);
  unsigned long long* qacc = acc + (size_t)P.Kp * ST;
  unsigned* cacc = reinterpret_cast<unsigned*>(qacc + P.Kp);
  const int slice = blockIdx.x % P.nslice, wg = blockIdx.x / P.nslice;
  __shared__ unsigned moved_s;  // points whose label changed (one global add)
  if (threadIdx.x == 0) moved_s = 0u;
  for (int t = threadIdx.x; t < P.Kp * ST + P.Kp; t += kFxThreads) acc[t] = 0ull;
  for (int t = threadIdx.x; t < P.Kp; t += kFxThreads) cacc[t] = 0u;
  __syncthreads();
  const long long per = (P.n + P.wgps - 1) / P.wgps;
  const long long lo = (long long)wg * per, hi = min(P.n, lo + per);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // lane = (point of the step, 4-dimension group): a float4 load per lane
  constexpr int LPP = DS / 4;    // lanes per point
  constexpr int PPW = 64 / LPP;  // points per wave step
  constexpr int U = 8;           // steps whose loads are issued together
  constexpr int NW = kFxThreads / 64;
  const int pp = lane / LPP, dg = lane % LPP;
  const int col = slice * DS + 4 * dg;
  const float fs = P.fscale;
  // software pipeline: the loads of step i+1 are in flight while the
  // atomics of step i run
  const long long stride = (long long)NW * PPW * U;
  float4 v[U], vn[U];
  int lab[U], labn[U];
  // the lanes that add a point's ||x||^2 and count (slice 0, first lane of
  // the point) also load its ||x||^2 and previous label with the step
  const bool pt_lane = slice == 0 && dg == 0;
  double q[U], qn[U];
  int pv[U], pvn[U];
  auto load_step = [&](long long i0, float4(&vv)[U], int(&ll)[U], double(&qq)[U],
                       int(&pq)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = i0 + u * PPW + pp;
      const bool ok = i < hi;
      const long long ic = ok ? i : lo;  // a valid row; its values are masked
      ll[u] = ok ? P.labels[ic] : 0;
      vv[u] = *reinterpret_cast<const float4*>(P.X + (size_t)ic * P.DP + col);
      if (!ok) vv[u] = make_float4(0.f, 0.f, 0.f, 0.f);  // adds exact zeros
      qq[u] = (pt_lane && ok) ? P.xn64[ic] : 0.0;
      pq[u] = (pt_lane && ok && P.Stot) ? P.prev[ic] : 0;
    }
  };
  unsigned mv_lane = 0;  // this lane's moved points, added to moved_s once
  long long i0 = lo + (long long)wave * PPW * U;
  if (mode == FX_MOVED) {
    fx_moved_rows<DS, M>(P, acc, qacc, cacc, reinterpret_cast<int2*>(cacc + P.Kp), slice, lo,
                         hi, mv_lane);
    i0 = hi;  // the full-pass loop below is skipped
  } else if (i0 < hi) {
    load_step(i0, v, lab, q, pv);
  }
  for (; i0 < hi; i0 += stride) {
    const bool more = i0 + stride < hi;
    if (more) load_step(i0 + stride, vn, labn, qn, pvn);
    // a cluster's accumulator row is component-major (dimension 4 g + j
    // at j * LPP + g), so the LPP lanes of one point add to consecutive
    // 8-byte words in every atomic instruction: no LDS bank conflict among
    // them
#pragma unroll
    for (int u = 0; u < U; ++u) {
      unsigned long long* a = acc + (size_t)lab[u] * ST + dg;
      atomicAdd(a + 0 * LPP, fx_int<M>(v[u].x * fs));
      atomicAdd(a + 1 * LPP, fx_int<M>(v[u].y * fs));
      atomicAdd(a + 2 * LPP, fx_int<M>(v[u].z * fs));
      atomicAdd(a + 3 * LPP, fx_int<M>(v[u].w * fs));
    }
    if (pt_lane) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long i = i0 + u * PPW + pp;
        const bool ok = i < hi;
        atomicAdd(&qacc[lab[u]], (unsigned long long)__double2ll_rn(q[u] * P.qscale));
        atomicAdd(&cacc[lab[u]], ok ? 1u : 0u);
        if (ok && P.Stot) {  // the labels the totals will stand for
          mv_lane += pv[u] != lab[u] ? 1u : 0u;
          P.prev[i] = lab[u];
        }
      }
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        v[u] = vn[u];
        lab[u] = labn[u];
        q[u] = qn[u];
        pv[u] = pvn[u];
      }
    }
  }
  if (mv_lane) atomicAdd(&moved_s, mv_lane);
  __syncthreads();
  if (threadIdx.x == 0 && moved_s) atomicAdd(&P.state->moved_acc, (unsigned long long)moved_s);
  long long* out = P.fpart + ((size_t)slice * P.wgps + wg) * (size_t)P.Kp * DS;
  for (int t = threadIdx.x; t < P.Kp * DS; t += kFxThreads) {
    const int c = t / DS, d = t - c * DS;
    out[t] = (long long)acc[(size_t)c * ST + (d & 3) * LPP + (d >> 2)];
  }
  if (slice == 0) {
    for (int t = threadIdx.x; t < P.Kp; t += kFxThreads) {
      P.qpart[(size_t)wg * P.Kp + t] = (long long)qacc[t];
      P.cpart[(size_t)wg * P.Kp + t] = (int)cacc[t];
    }
  }
}

// max |x| (as bits of a non-negative float) and max ||x||^2 (bits of a
// non-negative double) for the fixed-point scales; bit patterns of
// non-negative numbers order like the numbers
__global__ __launch_bounds__(256) void k_absmax(const float* X, long long count,
                                                const double* xn64, long long n, unsigned* amax,
                                                unsigned long long* qmax) {
  unsigned a = 0;
  unsigned long long q = 0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (long long)gridDim.x * 256)
    a = max(a, __float_as_uint(fabsf(X[i])));
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    q = max(q, (unsigned long long)__double_as_longlong(xn64[i]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a = max(a, (unsigned)__shfl_xor((int)a, o, 64));
    q = max(q, (unsigned long long)__shfl_xor((long long)q, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(amax, a);
    atomicMax(qmax, q);
  }
}

// ---------------------------------------------------------------------------
// k-means++ seeding (D^2 sampling). k_pp_dist folds the newest centroid into
// every point's FP64 squared distance to the chosen set (sequential over the
// dimensions) and forms per-block sums in a fixed order: a block of 256
// threads owns kPPBlock consecutive points, thread t adds points t, t+256,
// t+512, t+768 in that order, then a pairwise tree over the threads.
// k_pp_pick (one workgroup) draws the next centroid: with u the round's
// uniform in [0,1) and T the sum of the block sums in block order, the first
// point whose running sum (block prefix, then the block's points in order)
// exceeds u*T. Round 0 takes point floor(u*n).
// ---------------------------------------------------------------------------
constexpr int kPPBlock = 1024;
__global__ __launch_bounds__(256) void k_pp_dist(const float* X, long long n, int DP,
                                                 const float* crow, int first, double* D2,
                                                 double* bsum) {
  __shared__ float c_s[kMaxDP];
  __shared__ double r_s[256];
  for (int j = threadIdx.x; j < DP; j += 256) c_s[j] = crow[j];
  __syncthreads();
  const long long base = (long long)blockIdx.x * kPPBlock;
  double part = 0.0;
#pragma unroll
  for (int u = 0; u < kPPBlock / 256; ++u) {
    const long long i = base + threadIdx.x + 256 * u;
    if (i < n) {
      const float* x = X + (size_t)i * DP;
      double acc = 0.0;
      for (int j = 0; j < DP; j += 4) {
        const float4 v = *reinterpret_cast<const float4*>(x + j);
        double t;
        t = (double)v.x - (double)c_s[j];
        acc = fma(t, t, acc);
        t = (double)v.y - (double)c_s[j + 1];
        acc = fma(t, t, acc);
        t = (double)v.z - (double)c_s[j + 2];
        acc = fma(t, t, acc);
        t = (double)v.w - (double)c_s[j + 3];
        acc = fma(t, t, acc);
      }
      const double dv = first ? acc : fmin(D2[i], acc);
      D2[i] = dv;
      part += dv;
    }
  }
  r_s[threadIdx.x] = part;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) r_s[threadIdx.x] += r_s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = r_s[0];
}

__global__ __launch_bounds__(256) void k_pp_pick(const double* bsum, int nb, const double* D2,
                                                 long long n, const float* X, int DP, float* C,
                                                 int r, const double* uni, long long* chosen) {
  __shared__ long long idx_s;
  if (threadIdx.x == 0) {
    const double u = uni[r];
    long long idx = min((long long)(u * (double)n), n - 1);
    if (r > 0) {
      double total = 0.0;
      for (int b = 0; b < nb; ++b) total += bsum[b];
      if (total > 0.0) {
        const double target = u * total;
        double run = 0.0;
        int b = 0;
        for (; b < nb - 1; ++b) {
          if (run + bsum[b] > target) break;
          run += bsum[b];
        }
        const long long lo = (long long)b * kPPBlock, hi = min(n, lo + kPPBlock);
        long long pick = -1, last_pos = -1;
        for (long long i = lo; i < hi; ++i) {
          const double dv = D2[i];
          if (dv > 0.0) last_pos = i;
          run += dv;
          if (run > target && dv > 0.0) {
            pick = i;
            break;
          }
        }
        if (pick < 0) pick = last_pos;  // rounding left the target unreached
        if (pick >= 0) idx = pick;
      }
    }
    chosen[r] = idx;
    idx_s = idx;
  }
  __syncthreads();
  const long long idx = idx_s;
  for (int j = threadIdx.x; j < DP; j += 256) C[(size_t)r * DP + j] = X[(size_t)idx * DP + j];
}

// Greedy k-means++ (several D^2-sampled trials per round, the one leaving
// the smallest potential kept: the s = 2 + ceil(ln k) scheme of
// src/kmeans.cu:463-769). k_pp_trials: workgroup t draws trial t of the round
// as k_pp_pick does, from uniform uni[1 + (r-1) s + t]. k_pp_trial_cost:
// per block of kPPBlock points, for every trial, sum_i min(D2_i, ||x_i -
// trial||^2) in the k_pp_dist order. k_pp_choose: each trial's potential is
// its block sums added in block order; the first smallest wins and becomes
// centroid r.
__global__ __launch_bounds__(256) void k_pp_trials(const double* bsum, int nb, const double* D2,
                                                   long long n, int r, int s_trials,
                                                   const double* uni, long long* cand) {
  if (threadIdx.x != 0) return;
  const int t = blockIdx.x;
  const double u = uni[1 + (size_t)(r - 1) * s_trials + t];
  long long idx = min((long long)(u * (double)n), n - 1);
  double total = 0.0;
  for (int b = 0; b < nb; ++b) total += bsum[b];
  if (total > 0.0) {
    const double target = u * total;
    double run = 0.0;
    int b = 0;
    for (; b < nb - 1; ++b) {
      if (run + bsum[b] > target) break;
      run += bsum[b];
    }
    const long long lo = (long long)b * kPPBlock, hi = min(n, lo + kPPBlock);
    long long pick = -1, last_pos = -1;
    for (long long i = lo; i < hi; ++i) {
      const double dv = D2[i];
      if (dv > 0.0) last_pos = i;
      run += dv;
      if (run > target && dv > 0.0) {
        pick = i;
        break;
      }
    }
    if (pick < 0) pick = last_pos;
    if (pick >= 0) idx = pick;
  }
  cand[t] = idx;
}

constexpr int kPPMaxTrials = 16;
__global__ __launch_bounds__(256) void k_pp_trial_cost(const float* X, long long n, int DP,
                                                       const double* D2, const long long* cand,
                                                       int s_trials, double* tsum) {
  __shared__ float c_s[kPPMaxTrials][kMaxDP];
  __shared__ double r_s[256];
  for (int e = threadIdx.x; e < s_trials * DP; e += 256) {
    const int t = e / DP, j = e - t * DP;
    c_s[t][j] = X[(size_t)cand[t] * DP + j];
  }
  __syncthreads();
  const long long base = (long long)blockIdx.x * kPPBlock;
  for (int t = 0; t < s_trials; ++t) {
    double part = 0.0;
#pragma unroll
    for (int u = 0; u < kPPBlock / 256; ++u) {
      const long long i = base + threadIdx.x + 256 * u;
      if (i < n) {
        const float* x = X + (size_t)i * DP;
        double acc = 0.0;
        for (int j = 0; j < DP; j += 4) {
          const float4 v = *reinterpret_cast<const float4*>(x + j);
          double d;
          d = (double)v.x - (double)c_s[t][j];
          acc = fma(d, d, acc);
          d = (double)v.y - (double)c_s[t][j + 1];
          acc = fma(d, d, acc);
          d = (double)v.z - (double)c_s[t][j + 2];
          acc = fma(d, d, acc);
          d = (double)v.w - (double)c_s[t][j + 3];
          acc = fma(d, d, acc);
        }
        part += fmin(D2[i], acc);
      }
    }
    __syncthreads();  // r_s free
    r_s[threadIdx.x] = part;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) r_s[threadIdx.x] += r_s[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) tsum[(size_t)t * gridDim.x + blockIdx.x] = r_s[0];
  }
}

__global__ __launch_bounds__(256) void k_pp_choose(const double* tsum, int nb, int s_trials,
                                                   const long long* cand, const float* X,
                                                   int DP, float* C, int r, long long* chosen) {
  __shared__ long long idx_s;
  if (threadIdx.x == 0) {
    double best = 0.0;
    int bt = -1;
    for (int t = 0; t < s_trials; ++t) {
      double pot = 0.0;
      for (int b = 0; b < nb; ++b) pot += tsum[(size_t)t * nb + b];
      if (bt < 0 || pot < best) {
        best = pot;
        bt = t;
      }
    }
    chosen[r] = cand[bt];
    idx_s = cand[bt];
  }
  __syncthreads();
  const long long idx = idx_s;
  for (int j = threadIdx.x; j < DP; j += 256) C[(size_t)r * DP + j] = X[(size_t)idx * DP + j];
}

hipError_t launch_plusplus(const float* X, long long n, int DP, float* C, int K,
                           const double* uni, double* D2, double* bsum, long long* chosen,
                           hipStream_t s, int trials, long long* cand, double* tsum) {
  if (trials > 1 && (trials > kPPMaxTrials || !cand || !tsum)) return hipErrorInvalidValue;
  const int nb = (int)((n + kPPBlock - 1) / kPPBlock);
  hipLaunchKernelGGL(k_pp_pick, dim3(1), dim3(256), 0, s, bsum, nb, D2, n, X, DP, C, 0, uni,
                     chosen);
  for (int r = 1; r < K; ++r) {
    hipLaunchKernelGGL(k_pp_dist, dim3(nb), dim3(256), 0, s, X, n, DP,
                       C + (size_t)(r - 1) * DP, r == 1 ? 1 : 0, D2, bsum);
    if (trials <= 1) {
      hipLaunchKernelGGL(k_pp_pick, dim3(1), dim3(256), 0, s, bsum, nb, D2, n, X, DP, C, r, uni,
                         chosen);
    } else {
      hipLaunchKernelGGL(k_pp_trials, dim3(trials), dim3(64), 0, s, bsum, nb, D2, n, r, trials,
                         uni, cand);
      hipLaunchKernelGGL(k_pp_trial_cost, dim3(nb), dim3(256), 0, s, X, n, DP, D2, cand, trials,
                         tsum);
      hipLaunchKernelGGL(k_pp_choose, dim3(1), dim3(256), 0, s, tsum, nb, trials, cand, X, DP, C,
                         r, chosen);
    }
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Deterministic FP64 inertia / shift reductions, refinement counts, the
// iteration counter and the stop flag; one workgroup of kStatsThreads.
// ---------------------------------------------------------------------------
constexpr int kStatsThreads = 512;
__device__ void stats_reduce(const StatsParams& P) {
  __shared__ double sd[kStatsThreads];
  __shared__ double sm[kStatsThreads];
  __shared__ double ss[kStatsThreads];
  __shared__ int sc[kStatsThreads];
  __shared__ unsigned long long sq2[kStatsThreads], sqa[kStatsThreads];
  __shared__ DevState st_s;
  const int t = threadIdx.x;
  // every load is issued before the stop flag is known (they read memory
  // that is valid either way) and parked in LDS, so the kernel waits on
  // memory about once
  const int stop = *P.done;
  if (t == 0) st_s = *P.state;
  double in = 0.0, mx = 0.0, sh = 0.0;
  int all = 1;
  unsigned long long r2 = 0, ra = 0;
  {
    // the first eight entries per thread of every array (all of them up to
    // 8 * kStatsThreads); fixed order: deterministic
    unsigned a[8], b[8];
    double v[8], w[8];
    int cv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = u * kStatsThreads + t;
      a[u] = r < P.nregions ? P.q2cnt[r] : 0u;
      b[u] = r < P.nregions ? P.qacnt[r] : 0u;
      v[u] = r < P.nparts ? P.ipart[r] : 0.0;
      w[u] = r < P.K ? P.shift2[r] : 0.0;
      cv[u] = r < P.K ? P.conv[r] : 1;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      r2 += a[u];
      ra += b[u];
      in += v[u];
      mx = fmax(mx, w[u]);
      sh += w[u];
      all &= cv[u];
    }
  }
  // the rest, eight loads in flight per thread and round
  for (int r0 = 8 * kStatsThreads; r0 < P.nregions; r0 += 8 * kStatsThreads) {
    unsigned a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int r = r0 + u * kStatsThreads + t;
      a[u] = r < P.nregions ? P.q2cnt[r] : 0u;
      b[u] = r < P.nregions ? P.qacnt[r] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      r2 += a[u];
      ra += b[u];
    }
  }
  for (int i0 = 8 * kStatsThreads; i0 < P.nparts; i0 += 8 * kStatsThreads) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * kStatsThreads + t;
      v[u] = i < P.nparts ? P.ipart[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) in += v[u];
  }
  for (int c0 = 8 * kStatsThreads; c0 < P.K; c0 += 8 * kStatsThreads) {
    double v[8];
    int cv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int c = c0 + u * kStatsThreads + t;
      v[u] = c < P.K ? P.shift2[c] : 0.0;
      cv[u] = c < P.K ? P.conv[c] : 1;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      mx = fmax(mx, v[u]);
      sh += v[u];
      all &= cv[u];
    }
  }
  sd[t] = in;
  sm[t] = mx;
  ss[t] = sh;
  sc[t] = all;
  sq2[t] = r2;
  sqa[t] = ra;
  if (stop) return;
  __syncthreads();
  if (t >= 64) return;
  // The FP64 sums keep the order of the halving tree over the 512 values
  // (value t + o added to value t for o = 256, 128, ..., 1): wave 0 forms
  // its lanes' first three levels from LDS, shuffles the other six. The
  // integer sums, the maximum and the conjunction are exact in any order.
  static_assert(kStatsThreads == 512, "three levels through LDS");
  auto tree8 = [&](const double* v) {
    return ((v[t] + v[t + 256]) + (v[t + 128] + v[t + 384])) +
           ((v[t + 64] + v[t + 320]) + (v[t + 192] + v[t + 448]));
  };
  double vd = tree8(sd), vs = tree8(ss), vm = sm[t];
  int vc = sc[t];
  unsigned long long v2 = sq2[t], va = sqa[t];
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    vm = fmax(vm, sm[t + 64 * k]);
    vc &= sc[t + 64 * k];
    v2 += sq2[t + 64 * k];
    va += sqa[t + 64 * k];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    vd += __shfl_down(vd, o, 64);
    vs += __shfl_down(vs, o, 64);
    vm = fmax(vm, __shfl_down(vm, o, 64));
    vc &= __shfl_down(vc, o, 64);
    v2 += __shfl_down(v2, o, 64);
    va += __shfl_down(va, o, 64);
  }
  if (t == 0) {
    DevState sv = st_s;  // the state as the kernel found it
    DevState* s = &sv;
    if (P.incremental) {
      // the iteration's update is complete: its moved count decides the
      // next one, and the cluster totals now stand for its labels
      if (P.fx_modes ? fx_mode(s, P.n) == FX_FULL : fx_full_pass(s, P.n))
        s->full_passes += 1;
      else
        s->moved_total += s->moved_acc;
      s->moved = s->moved_acc;
      s->moved_acc = 0ull;
      s->sums_valid = 1;
    }
    const unsigned long long it = s->iter + 1;
    s->iter = it;
    s->inertia = vd;
    s->max_shift2 = vm;
    s->sum_shift2 = vs;
    s->refined_two = v2;
    s->refined_all = va;
    const double score = vd - s->sum_x2;
    if (P.mode == CONV_SHIFT && it >= 2 && vc) {
      // comparisons/CPU_Kmeans/src/kmeans.cpp:205-208: tested from iteration 2
      s->converged_at = it;
      s->done = 1;
    }
    if (P.mode == CONV_SCORE && it >= 2 && it < s->maxiter &&
        fabsf((float)score - (float)s->last_score) < P.tol) {
      // src/kmeans.cu:1064-1073: float scores, tested from iteration 2,
      // never at the last iteration
      s->converged_at = it;
      s->done = 1;
    }
    s->last_score = score;
    if (P.mode == CONV_SHIFT_SUM && vs < (double)P.tol && !s->done) {
      // experiments/raft-bench/src/kmeans.cu:168-212: sum over all centroid
      // entries of the squared change, tested every iteration
      s->converged_at = it;
      s->done = 1;
    }
    if (P.cost_ratio && it >= 2 && !s->done && s->prior_cost > 0.0 &&
        vd / s->prior_cost > 1.0 - (double)P.tol) {
      // experiments/raft-bench/src/kmeans.cu:186-204: the clustering cost
      // stopped falling by more than a tol fraction
      s->converged_at = it;
      s->done = 1;
    }
    s->prior_cost = vd;
    if (it >= s->maxiter) s->done = 1;
    *P.state = sv;
  }
}

// ---------------------------------------------------------------------------
// Finish kernels: the new centroid from the update's sums (k_finish_fx for
// the fixed-point update, k_finish for the tree and ordered updates,
// k_finish_waves for batches), then, on one wave, the float shift in the
// reference's order on lane 0 (comparisons/CPU_Kmeans/src/kmeans.cpp:225-234),
// the FP64 shift, C <- C_new lane-parallel, and the next iteration's distance
// operands for the centroid (prep_row).
// ---------------------------------------------------------------------------
// The end of a centroid's update, run by one wave: the empty-cluster rule,
// the float shift in the reference's order (squared terms lane-parallel,
// summed on one lane in dimension order), the FP64 shift, C <- C_new and the
// next iteration's distance operands. ns: the new centroid (LDS), sq_s: LDS
// scratch of the wave, os: the old centroid (LDS or the row of C), alive:
// alive[c] as the iteration found it.
__device__ __forceinline__ void finish_tail(const FinishParams& P, int c, int cnt, float* ns,
                                            float* sq_s, int lane, const float* os, bool alive) {
  float* cr = P.C + (size_t)c * P.DP;
  float* nrow = P.Cnext + (size_t)c * P.DP;
  if (P.keep_empty && cnt == 0) {
    // an empty cluster keeps its centroid (the explicit-Lloyd rule of
    // experiments/raft-bench/src/kmeans.cu:158-166)
    for (int j = lane; j < P.D; j += 64) {
      ns[j] = os[j];
      nrow[j] = os[j];
    }
  }
  __builtin_amdgcn_wave_barrier();
  double s64 = 0.0;
  for (int j = lane; j < P.D; j += 64) {
    const float o = os[j], nv = ns[j];
    const double td = (double)nv - (double)o;
    s64 = fma(td, td, s64);
    {
#pragma clang fp contract(off)
      const float t = nv - o;
      sq_s[j] = t * t;
    }
  }
  for (int j = P.D + lane; j < ((P.D + 3) & ~3); j += 64) sq_s[j] = 0.0f;
  s64 = wave_sum_f64(s64);
  __builtin_amdgcn_wave_barrier();
  if (P.exclude_empty && cnt == 0) alive = false;
  if (lane == 0) {
    // the squared terms in dimension order, 32 of them loaded before they
    // are added (the padding up to a multiple of 4 holds zeros: s + 0 = s
    // for s >= 0)
    float s = 0.0f;
    for (int j0 = 0; j0 < P.D; j0 += 32) {
      float4 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        v[u] = j0 + 4 * u < P.D ? *reinterpret_cast<const float4*>(&sq_s[j0 + 4 * u])
                                : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
#pragma clang fp contract(off)
        s = s + v[u].x;
        s = s + v[u].y;
        s = s + v[u].z;
        s = s + v[u].w;
      }
    }
    P.conv[c] = (__fsqrt_rn(s) > P.tol) ? 0 : 1;
    P.shift2[c] = s64;
    if (!alive) P.alive[c] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  for (int j = lane; j < P.D; j += 64) cr[j] = ns[j];
  // the next iteration's distance operands for this centroid
  if (P.do_prep) {
    __builtin_amdgcn_wave_barrier();
    prep_row(P.prep, c, lane, alive, [&](int j) { return j < P.D ? ns[j] : 0.0f; });
  }
}

constexpr int kFinishWaves = 8;

// k_finish_fx: k_finish for the fixed-point update (one workgroup of
// kFinishWaves waves per centroid). The exact integer sums -- a full pass's
// workgroup partials, or the totals plus the changes of an incremental
// iteration (workgroup partials of the moved points, or the delta copies of
// the global atomics, wave w taking copy w) -- are rounded once into the new
// centroid, and the FP64 inertia of the assignment is formed from them:
//   sum ||x - c||^2 = Q - 2 c.S + n ||c||^2
// Every load that depends neither on the stop flag nor on the update's form
// (the delta copy, the totals, the old centroid, alive[c]) is issued at the
// start, beside the loads of the stop flag and the state, and parked in
// LDS, so an iteration of the incremental form waits on memory about once.
__global__ __launch_bounds__(64 * kFinishWaves) void k_finish_fx(FinishParams P) {
  __shared__ long long fs_s[kFinishWaves][kMaxDP];
  __shared__ long long tot_s[kMaxDP];
  __shared__ float os_s[kMaxDP];
  __shared__ float ns[kMaxDP];
  __shared__ __attribute__((aligned(16))) float sq_s[kMaxDP];
  __shared__ long long q_s[kFinishWaves];
  __shared__ int n_s[kFinishWaves];
  __shared__ long long qtot_s;
  __shared__ int ntot_s, alive_s;
  constexpr int JT = kMaxDP / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x;
  float* cr = P.C + (size_t)c * P.DP;
  float* nrow = P.Cnext + (size_t)c * P.DP;
  const size_t dco = (size_t)wave * P.Kp + c;  // this wave's delta copy row
  const int stop = *P.done;
  const int fxm = P.Stot ? fx_mode(P.stats.state, P.n) : FX_FULL;
  {
    long long dv[JT], tv[JT];
    float ov[JT];
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int j = lane + 64 * t;
      const bool in = j < P.D;
      dv[t] = (P.Sdel && in) ? P.Sdel[dco * P.DP + j] : 0;
      tv[t] = (wave == 0 && P.Stot && in) ? P.Stot[(size_t)c * P.DP + j] : 0;
      ov[t] = (wave == 0 && in) ? cr[j] : 0.0f;
    }
    long long qd = 0, qt0 = 0;
    int nd = 0, nt0 = 0, al = 1;
    if (lane == 0) {
      if (P.Sdel) {
        qd = P.Qdel ? P.Qdel[dco] : 0;
        nd = P.Ndel[dco];
      }
      if (wave == 0) {
        if (P.Stot) {
          qt0 = P.Qtot[c];
          nt0 = P.Ntot[c];
        }
        al = P.alive[c];
      }
    }
#pragma unroll
    for (int t = 0; t < JT; ++t) {
      const int j = lane + 64 * t;
      if (j < P.D) {
        fs_s[wave][j] = dv[t];
        if (wave == 0) {
          tot_s[j] = tv[t];
          os_s[j] = ov[t];
        }
      }
    }
    if (lane == 0) {
      q_s[wave] = qd;
      n_s[wave] = nd;
      if (wave == 0) {
        qtot_s = qt0;
        ntot_s = nt0;
        alive_s = al;
      }
    }
  }
  if (stop) return;
  const bool from_totals = fxm != FX_FULL;
  // the workgroup partials: the sums of a full pass, or (FX_MOVED) the
  // changes of the totals
  const bool parts = fxm != FX_ATOMIC;
  // the delta copies hold this iteration's moves (FX_ATOMIC); cleared for
  // the next iteration
  static_assert(kDeltaCopies == kFinishWaves, "one delta copy per wave");
  const bool deltas = fxm == FX_ATOMIC && P.Sdel;
  if (deltas) {
    for (int j = lane; j < P.D; j += 64) P.Sdel[dco * P.DP + j] = 0;
    if (lane == 0) {
      if (P.Qdel) P.Qdel[dco] = 0;
      P.Ndel[dco] = 0;
    }
  }
  if (fxm == FX_MOVED && P.prev) {
    // the totals now stand for this iteration's labels: prev <- labels
    // (the update left prev alone), 16 bytes a lane, two loads in flight
    const long long n4 = P.n >> 2, nt = (long long)gridDim.x * blockDim.x;
    const int4* src = reinterpret_cast<const int4*>(P.labels);
    int4* dst = reinterpret_cast<int4*>(P.prev);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += 2 * nt) {
      const int4 a = src[i];
      const bool two = i + nt < n4;
      int4 b = a;
      if (two) b = src[i + nt];
      dst[i] = a;
      if (two) dst[i + nt] = b;
    }
    if (blockIdx.x == 0 && threadIdx.x < (P.n & 3)) {
      const long long i = (n4 << 2) + threadIdx.x;
      P.prev[i] = P.labels[i];
    }
  }
  if (parts) {
    // wave w adds the partials of workgroups w, w+8, ...: every dimension's
    // loads of a round (up to 8 partials) are issued before any is added
    // (integer sums: the order does not matter)
    constexpr int JT = kMaxDP / 64, UF = 8;
    long long accj[JT];
#pragma unroll
    for (int jt = 0; jt < JT; ++jt) accj[jt] = 0;
    const size_t wst = (size_t)P.Kp * P.ds;  // one workgroup further
    for (int w0 = wave; w0 < P.wgps; w0 += UF * kFinishWaves) {
      long long t[JT][UF];
#pragma unroll
      for (int jt = 0; jt < JT; ++jt) {
        const int j = lane + 64 * jt;
        const int sl = j / P.ds, dj = j - sl * P.ds;
        const long long* fp = P.fpart + ((size_t)sl * P.wgps * P.Kp + c) * P.ds + dj;
#pragma unroll
        for (int u = 0; u < UF; ++u) {
          const int w = w0 + u * kFinishWaves;
          t[jt][u] = (j < P.D && w < P.wgps) ? fp[(size_t)w * wst] : 0;
        }
      }
#pragma unroll
      for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int u = 0; u < UF; ++u) accj[jt] += t[jt][u];
    }
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
      if (lane + 64 * jt < P.D) fs_s[wave][lane + 64 * jt] = accj[jt];
    {
      // the fixed-point ||x||^2 and counts: a lane per partial, then the
      // wave's sum
      long long qsum = 0;
      int nsum = 0;
      for (int w = wave + kFinishWaves * lane; w < P.wgps; w += 64 * kFinishWaves) {
        qsum += P.qpart[(size_t)w * P.Kp + c];
        nsum += P.cpart[(size_t)w * P.Kp + c];
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        qsum += __shfl_xor(qsum, o, 64);
        nsum += __shfl_xor(nsum, o, 64);
      }
      if (lane == 0) {
        q_s[wave] = qsum;
        n_s[wave] = nsum;
      }
    }
  }
  __syncthreads();
  if (wave != 0) return;
  // totals (when they are kept and valid) plus the waves' delta copies or
  // partials, stored back as the new totals
  const bool adds = deltas || parts;
  long long qt = from_totals ? qtot_s : 0;
  int nt = from_totals ? ntot_s : 0;
  if (adds) {
#pragma unroll
    for (int w = 0; w < kFinishWaves; ++w) {
      qt += q_s[w];
      nt += n_s[w];
    }
    if (P.Stot && lane == 0) {
      P.Qtot[c] = qt;
      P.Ntot[c] = nt;
    }
  }
  double cs = 0.0, cc = 0.0;
  for (int j = lane; j < P.D; j += 64) {
    long long sj = from_totals ? tot_s[j] : 0;
    if (adds) {
#pragma unroll
      for (int w = 0; w < kFinishWaves; ++w) sj += fs_s[w][j];
      if (P.Stot) P.Stot[(size_t)c * P.DP + j] = sj;
    }
    const double S = (double)sj * P.inv_fscale;
    const float v = (float)(S / (double)(nt > 1 ? nt : 1));
    ns[j] = v;
    nrow[j] = v;
    const double co = (double)os_s[j];
    cs = fma(co, S, cs);
    cc = fma(co, co, cc);
  }
  cs = wave_sum_f64(cs);
  cc = wave_sum_f64(cc);
  if (lane == 0) {
    P.count_out[c] = nt;
    P.ipart_out[c] = (double)qt * P.inv_qscale - 2.0 * cs + (double)nt * cc;
  }
  finish_tail(P, c, nt, ns, sq_s, lane, os_s, alive_s != 0);
}

// k_finish for the tree and ordered updates: one workgroup of kFinishWaves
// waves per centroid. With the tree update the waves first add the
// cluster's chunk partials: wave w takes partials q0+w, q0+w+kFinishWaves,
// ... in order, and the wave sums are added in wave order, so the result
// depends only on the partials (a cluster of any size costs the same few
// load rounds per wave). c_new = sum / count, rounded once (an empty cluster
// becomes the zero vector); then finish_tail on wave 0.
__global__ __launch_bounds__(64 * kFinishWaves) void k_finish(FinishParams P) {
  if (*P.done) return;
  __shared__ double part_s[kFinishWaves][kMaxDP];
  __shared__ float ns[kMaxDP];
  __shared__ __attribute__((aligned(16))) float sq_s[kMaxDP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x;
  float* cr = P.C + (size_t)c * P.DP;
  float* nrow = P.Cnext + (size_t)c * P.DP;
  int cnt = P.count[c];
  if (P.tree == 2) {
    // exact tree form: the 64-bit integer segment sums of a full pass (the
    // cluster's chunks of the sorted members, any order), stored as the
    // cluster totals, or the totals the moved points kept up to date; the
    // means rounded once, as the fixed-point update's
    long long(*is_s)[kMaxDP] = reinterpret_cast<long long(*)[kMaxDP]>(part_s);
    const bool from_totals = !fx_full_pass(P.fx_state, P.n);
    if (from_totals) cnt = P.Ntot[c];
    long long q0 = 0, q1 = -1;
    if (!from_totals && cnt > 0) {
      q0 = (long long)P.start[c] / kTreeRows;
      q1 = ((long long)P.start[c] + cnt - 1) / kTreeRows;
    }
    constexpr int S = kFinishWaves;
    const long long* tp64 = reinterpret_cast<const long long*>(P.tpart);
    for (int j = lane; j < P.D && !from_totals; j += 64) {
      const long long* tp = tp64 + (size_t)c * P.DP + j;
      long long sum = 0;
      long long q = q0 + wave;
      for (; q + 3 * S <= q1; q += 4 * S) {
        long long v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = tp[(size_t)(q + u * S) * P.DP];
#pragma unroll
        for (int u = 0; u < 4; ++u) sum += v[u];
      }
      for (; q <= q1; q += S) sum += tp[(size_t)q * P.DP];
      is_s[wave][j] = sum;
    }
    __syncthreads();
    if (wave == 0) {
      const double inv = P.cl_inv_scale[c];
      for (int j = lane; j < P.D; j += 64) {
        long long sj = 0;
        if (from_totals) {
          sj = P.Stot[(size_t)c * P.DP + j];
        } else {
#pragma unroll
          for (int w = 0; w < S; ++w) sj += is_s[w][j];
          P.Stot[(size_t)c * P.DP + j] = sj;
        }
        const double Sd = (double)sj * inv;
        const float v = (float)(Sd / (double)(cnt > 1 ? cnt : 1));
        ns[j] = v;
        nrow[j] = v;
      }
      if (!from_totals && lane == 0) P.Ntot[c] = cnt;
    }
  } else if (P.tree) {
    long long q0 = 0, q1 = -1;
    if (cnt > 0) {
      q0 = (long long)P.start[c] / kTreeRows;
      q1 = ((long long)P.start[c] + cnt - 1) / kTreeRows;
    }
    constexpr int S = kFinishWaves;
    for (int j = lane; j < P.D; j += 64) {
      const double* tp = P.tpart + (size_t)c * P.DP + j;
      double sum = 0.0;
      long long q = q0 + wave;
      for (; q + 3 * S <= q1; q += 4 * S) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = tp[(size_t)(q + u * S) * P.DP];
#pragma unroll
        for (int u = 0; u < 4; ++u) sum += v[u];
      }
      for (; q <= q1; q += S) sum += tp[(size_t)q * P.DP];
      part_s[wave][j] = sum;
    }
    __syncthreads();
    if (wave == 0) {
      for (int j = lane; j < P.D; j += 64) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < S; ++w) sum += part_s[w][j];
        const float v = (float)(sum / (double)(cnt > 1 ? cnt : 1));
        ns[j] = v;
        nrow[j] = v;
      }
    }
  } else if (wave == 0) {
    for (int j = lane; j < P.D; j += 64) ns[j] = nrow[j];
  }
  if (wave == 0) finish_tail(P, c, cnt, ns, sq_s, lane, cr, P.alive[c] != 0);
}

// k_finish for the exact tree form with many small clusters (batches): one
// wave per cluster, four clusters per workgroup. The wave adds the integer
// partials of the cluster's chunks itself (a batch cluster spans few chunks)
// or reads the totals, rounds the means as k_finish does, and runs the same
// tail. No distance operands are written (the batch prepares its own).
constexpr int kFinishClustersPerBlock = 4;
__global__ __launch_bounds__(64 * kFinishClustersPerBlock) void k_finish_waves(FinishParams P) {
  if (*P.done) return;
  __shared__ float ns_s[kFinishClustersPerBlock][kMaxDP];
  __shared__ __attribute__((aligned(16))) float sq_s[kFinishClustersPerBlock][kMaxDP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x * kFinishClustersPerBlock + wave;
  if (c >= P.K) return;
  float* ns = ns_s[wave];
  float* nrow = P.Cnext + (size_t)c * P.DP;
  const bool from_totals = !fx_full_pass(P.fx_state, P.n);
  const int cnt = from_totals ? P.Ntot[c] : P.count[c];
  long long q0 = 0, q1 = -1;
  if (!from_totals && cnt > 0) {
    q0 = (long long)P.start[c] / kTreeRows;
    q1 = ((long long)P.start[c] + cnt - 1) / kTreeRows;
  }
  const long long* tp64 = reinterpret_cast<const long long*>(P.tpart);
  const double inv = P.cl_inv_scale[c];
  for (int j = lane; j < P.D; j += 64) {
    long long sj = 0;
    if (from_totals) {
      sj = P.Stot[(size_t)c * P.DP + j];
    } else {
      const long long* tp = tp64 + (size_t)c * P.DP + j;
      for (long long q = q0; q <= q1; ++q) sj += tp[(size_t)q * P.DP];
      P.Stot[(size_t)c * P.DP + j] = sj;
    }
    const double Sd = (double)sj * inv;
    const float v = (float)(Sd / (double)(cnt > 1 ? cnt : 1));
    ns[j] = v;
    nrow[j] = v;
  }
  if (!from_totals && lane == 0) P.Ntot[c] = cnt;
  finish_tail(P, c, cnt, ns, sq_s[wave], lane, P.C + (size_t)c * P.DP, P.alive[c] != 0);
}

// k_stats: one workgroup.
__global__ __launch_bounds__(kStatsThreads) void k_stats(StatsParams P) {
  stats_reduce(P);  // returns at once (after its loads) when the stop flag is set
}


// ---------------------------------------------------------------------------
// Set-up kernels
// ---------------------------------------------------------------------------

// Padded fp32 copy of X, ||x||^2 (fp32 and FP64), bf16 hi/lo planes.
__global__ __launch_bounds__(256) void k_load_points(LoadParams P) {
  const int lane = threadIdx.x & 63;
  const long long i = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= P.n) return;
  const float* src = P.src + (size_t)i * P.ld;
  if (P.Xq) {
    // fp8 plane: the row quantised with its own power-of-two scale
    float amax = 0.0f;
    for (int j = lane; j < P.D; j += 64) amax = fmaxf(amax, fabsf(src[j]));
    amax = wave_max_f32(amax);
    const int e = fp8_scale_exp(amax);
    double d2 = 0.0;
    for (int j = lane; j < P.DP; j += 64) {
      const float v = j < P.D ? src[j] : 0.0f;
      const uint32_t code = fp8_e4m3_rne(ldexpf(v, -e));
      P.Xq[(size_t)i * P.DP + j] = (uint8_t)code;
      const double dv = (double)ldexpf(fp8_e4m3_value(code), e) - (double)v;
      d2 = fma(dv, dv, d2);
    }
    d2 = wave_sum_f64(d2);
    if (lane == 0) {
      P.xsc[i] = e + 127;
      if (P.xdn) P.xdn[i] = norm_up(d2);
    }
  }
  float s = 0.0f;
  double s64 = 0.0, d2 = 0.0;
  for (int j = lane; j < P.DP; j += 64) {
    const float v = j < P.D ? src[j] : 0.0f;
    if (P.X) P.X[(size_t)i * P.DP + j] = v;  // null: norms only (rows already padded)
    s = fmaf(v, v, s);
    s64 += (double)v * (double)v;
    if (P.Xhi) {
      const uint16_t hi = bf16_rne(v);
      P.Xhi[(size_t)i * P.DP + j] = hi;
      if (P.Xlo) P.Xlo[(size_t)i * P.DP + j] = bf16_rne(v - bf16_to_f32(hi));
      const double dv = (double)bf16_to_f32(hi) - (double)v;
      d2 = fma(dv, dv, d2);
    }
  }
  if (P.xdn && !P.Xq) {
    d2 = wave_sum_f64(d2);
    if (lane == 0) P.xdn[i] = norm_up(d2);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    s64 += __shfl_xor(s64, o, 64);
  }
  if (lane == 0) {
    P.xn2[i] = s;
    if (P.xn2_64) P.xn2_64[i] = s64;
  }
}

// Sum of FP64 values in a fixed order: stage 1 (grid of 256 blocks) writes
// one partial per block into part[], stage 2 (one block) adds them.
__global__ __launch_bounds__(256) void k_sum_f64_part(const double* v, long long n, double* part) {
  __shared__ double sd[256];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += 256LL * gridDim.x)
    s += v[i];
  sd[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sd[threadIdx.x] += sd[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sd[0];
}

__global__ __launch_bounds__(256) void k_sum_f64(const double* v, long long n, double* out) {
  __shared__ double sd[256];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += 256) s += v[i];
  sd[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sd[threadIdx.x] += sd[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sd[0];
}

// Fill helpers (graph-safe replacement for memset).
__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Host-side launch helpers
// ---------------------------------------------------------------------------
size_t assign_lds(int DP, int prec, int kchunk) {
#define KMC_LDS_CASE(dp)                                                             \
  case dp:                                                                           \
    return prec == PREC_FP32     ? assign_lds_bytes<dp, PREC_FP32>(kchunk)           \
           : prec == PREC_BF16X3 ? assign_lds_bytes<dp, PREC_BF16X3>(kchunk)         \
           : prec == PREC_FP8    ? assign_lds_bytes<dp, PREC_FP8>(kchunk)            \
                                 : assign_lds_bytes<dp, PREC_BF16>(kchunk);
  switch (DP) {
    KMC_LDS_CASE(16)
    KMC_LDS_CASE(32)
    KMC_LDS_CASE(64)
    KMC_LDS_CASE(128)
    KMC_LDS_CASE(256)
    default: return 0;
  }
#undef KMC_LDS_CASE
}

// Launch variants of the distance kernel: (waves per workgroup, point sets
// per wave, prefetch, waves per SIMD the register allocation targets (0 =
// none), kernel (0 = k_assign, 1 = the two-set k_assign2, 2 = its form
// without the block software pipeline), values kept per list, lists per
// lane (k_assign2 only)). Two point sets per wave need both fragment sets in
// registers; where they do not fit (kFragRegs > 64) the variant is not
// built. Targets above
// four waves per SIMD select the register-lean loop and let two workgroups
// share a CU when their centroid copies fit the LDS together. k_assign2 is
// built for the single-product forms only (bf16 up to DP = 128, fp8) and
// needs every centroid resident in LDS.
struct VariantShape {
  int nw;
  int ms;
  bool pf;
  int wpe;
  int two;
  int topk;
  int lists = 1;
};
constexpr VariantShape kVariants[] = {
    {8, 1, true, 0, 0, 4},     {8, 1, false, 0, 0, 4},    {12, 1, false, 0, 0, 4},
    {16, 1, false, 0, 0, 4},   {12, 1, false, 6, 0, 4},   {4, 2, false, 0, 0, 4},
    {8, 2, false, 0, 0, 4},    {8, 1, false, 8, 0, 4},    {8, 2, false, 0, 1, 4},
    {8, 2, false, 0, 1, 3, 2}, {4, 2, false, 0, 1, 4},    {4, 2, false, 0, 2, 4},
    {8, 2, false, 0, 1, 3, 4}, {4, 2, false, 0, 2, 3, 2}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
static_assert(kNumVariants <= KMC_MAX_VARIANTS, "variant table exceeds kmc_tuning");

template <int DP, int PREC, int V>
static const void* kernel_of() {
  constexpr VariantShape S = kVariants[V];
  if constexpr (S.two) {
    constexpr bool kBuilt = (PREC == PREC_BF16 && DP <= 128) || PREC == PREC_FP8;
    if constexpr (kBuilt)
      return reinterpret_cast<const void*>(
          k_assign2<DP, PREC, S.nw, S.topk, S.pf, S.two == 1, S.wpe, S.lists>);
    else
      return nullptr;
  } else {
    // a register target leaves room for the point fragments plus ~40 others
    constexpr bool kFits = S.wpe == 0 || frag_regs<DP, PREC>() * S.ms + 40 <= 512 / S.wpe;
    if constexpr ((S.ms > 1 && frag_regs<DP, PREC>() > 64) || !kFits) {
      return nullptr;
    } else {
      return reinterpret_cast<const void*>(k_assign<DP, PREC, S.nw, S.ms, S.pf, S.wpe>);
    }
  }
}

template <int DP, int PREC>
static const void* kernel_of_v(int v) {
  switch (v) {
    case 0: return kernel_of<DP, PREC, 0>();
    case 1: return kernel_of<DP, PREC, 1>();
    case 2: return kernel_of<DP, PREC, 2>();
    case 3: return kernel_of<DP, PREC, 3>();
    case 4: return kernel_of<DP, PREC, 4>();
    case 5: return kernel_of<DP, PREC, 5>();
    case 6: return kernel_of<DP, PREC, 6>();
    case 7: return kernel_of<DP, PREC, 7>();
    case 8: return kernel_of<DP, PREC, 8>();
    case 9: return kernel_of<DP, PREC, 9>();
    case 10: return kernel_of<DP, PREC, 10>();
    case 11: return kernel_of<DP, PREC, 11>();
    case 12: return kernel_of<DP, PREC, 12>();
    case 13: return kernel_of<DP, PREC, 13>();
    default: return nullptr;
  }
}

template <int PREC>
static const void* kernel_of_dv(int DP, int v) {
  if constexpr (PREC == PREC_FP8) {  // one fp8 step covers 64 dimensions
    switch (DP) {
      case 64: return kernel_of_v<64, PREC>(v);
      case 128: return kernel_of_v<128, PREC>(v);
      case 256: return kernel_of_v<256, PREC>(v);
      default: return nullptr;
    }
  } else {
    switch (DP) {
      case 16: return kernel_of_v<16, PREC>(v);
      case 32: return kernel_of_v<32, PREC>(v);
      case 64: return kernel_of_v<64, PREC>(v);
      case 128: return kernel_of_v<128, PREC>(v);
      case 256: return kernel_of_v<256, PREC>(v);
      default: return nullptr;
    }
  }
}

const void* assign_kernel_ptr(int DP, int prec, int v) {
  switch (prec) {
    case PREC_FP32: return kernel_of_dv<PREC_FP32>(DP, v);
    case PREC_BF16X3: return kernel_of_dv<PREC_BF16X3>(DP, v);
    case PREC_BF16: return kernel_of_dv<PREC_BF16>(DP, v);
    case PREC_FP8: return kernel_of_dv<PREC_FP8>(DP, v);
    default: return nullptr;
  }
}

// The dynamic-LDS limit of a kernel is process-wide state: handles of other
// shapes launch the same kernel, so it is only ever raised.
hipError_t raise_lds_limit(const void* fn, size_t bytes) {
  static std::mutex mu;
  static std::map<const void*, size_t> set;
  std::lock_guard<std::mutex> lock(mu);
  auto it = set.find(fn);
  if (it != set.end() && it->second >= bytes) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)bytes);
  if (e == hipSuccess) set[fn] = bytes;
  return e;
}

int assign_variants() { return kNumVariants; }
bool assign_variant_resident_only(int v) { return kVariants[v].two != 0; }
int assign_threads(int v) { return 64 * kVariants[v].nw; }
int assign_tile(int v) { return 32 * kVariants[v].ms * kVariants[v].nw; }

hipError_t launch_assign(int DP, int prec, int v, const AssignParams& P, int grid, size_t lds,
                         hipStream_t s) {
  if (v < 0 || v >= kNumVariants) return hipErrorInvalidValue;
  const void* fn = assign_kernel_ptr(DP, prec, v);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<AssignParams*>(&P)};
  return hipLaunchKernel(fn, dim3(grid), dim3(assign_threads(v)), args, lds, s);
}

// Batch mode of the distance kernel: one shape (8 waves, one point set, 256
// points per tile) for every (DP, precision).
template <int PREC>
static const void* batch_kernel_of(int DP) {
  switch (DP) {
    case 16: return reinterpret_cast<const void*>(k_assign<16, PREC, 8, 1, false, 0, true>);
    case 32: return reinterpret_cast<const void*>(k_assign<32, PREC, 8, 1, false, 0, true>);
    case 64: return reinterpret_cast<const void*>(k_assign<64, PREC, 8, 1, false, 0, true>);
    case 128: return reinterpret_cast<const void*>(k_assign<128, PREC, 8, 1, false, 0, true>);
    case 256: return reinterpret_cast<const void*>(k_assign<256, PREC, 8, 1, false, 0, true>);
    default: return nullptr;
  }
}
const void* assign_batch_kernel_ptr(int DP, int prec) {
  switch (prec) {
    case PREC_FP32: return batch_kernel_of<PREC_FP32>(DP);
    case PREC_BF16X3: return batch_kernel_of<PREC_BF16X3>(DP);
    case PREC_BF16: return batch_kernel_of<PREC_BF16>(DP);
    default: return nullptr;
  }
}
hipError_t launch_assign_batch(int DP, int prec, const AssignParams& P, int grid, size_t lds,
                               hipStream_t s) {
  const void* fn = assign_batch_kernel_ptr(DP, prec);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<AssignParams*>(&P)};
  return hipLaunchKernel(fn, dim3(grid), dim3(kBatchAssignThreads), args, lds, s);
}

hipError_t launch_prep(const PrepParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_prep, dim3((P.Kp + 3) / 4), dim3(256), 0, s, P);
  return hipGetLastError();
}
template <int DP>
static hipError_t launch_pruned(const RefineParams& P, hipStream_t s) {
  static bool lds_set = false;
  if (!lds_set) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_refine_prune<DP>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)prune_lds<DP>());
    if (e != hipSuccess) return e;
    lds_set = true;
  }
  hipLaunchKernelGGL(k_refine_prune<DP>, dim3(kPruneBlocks), dim3(256), prune_lds<DP>(), s, P);
  return hipGetLastError();
}

hipError_t launch_refine(const RefineParams& P, hipStream_t s) {
  // k_refine2's grid: the workgroups resident at once for its LDS (a second
  // round of workgroups would only add a tail: 1024 measured slower than
  // the 512 that fit on 256 CUs); set up once per process
  static int grid2 = 0;
  if (!grid2) {
    const void* fn = reinterpret_cast<const void*>(k_refine2);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)kRefineLds);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, kRefineLds) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    grid2 = std::min(kRefineBlocks, per_cu * cus);
  }
  if (pq_active(P)) {
    hipError_t e = hipSuccess;
    switch (P.DP) {
      case 16: e = launch_fused<16>(P, s); break;
      case 32: e = launch_fused<32>(P, s); break;
      case 64: e = launch_fused<64>(P, s); break;
      case 128: e = launch_fused<128>(P, s); break;
      default: e = hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(k_refine2, dim3(grid2), dim3(256), kRefineLds, s, P);
    hipLaunchKernelGGL(k_refine_all, dim3(kRefineBlocks), dim3(256), kRefineAllLds, s, P);
  }
  if (!prune_active(P)) return hipGetLastError();
  switch (P.DP) {
    case 16: return launch_pruned<16>(P, s);
    case 32: return launch_pruned<32>(P, s);
    case 64: return launch_pruned<64>(P, s);
    case 128: return launch_pruned<128>(P, s);
    case 256: return launch_pruned<256>(P, s);
    default: return hipErrorInvalidValue;
  }
}
hipError_t launch_sort(const SortParams& P, hipStream_t s) {
  const size_t lds = (size_t)P.K * sizeof(int);
  hipLaunchKernelGGL(k_hist, dim3(P.nblk), dim3(64), lds, s, P);
  hipLaunchKernelGGL(k_scan_blocks, dim3((P.K + 3) / 4), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_scatter, dim3(P.nblk), dim3(64), lds, s, P);
  return hipGetLastError();
}
hipError_t launch_update(const UpdateParams& P, int mode, hipStream_t s) {
  const int S = (P.D + 63) / 64;
  if (mode == UPDATE_FIXED) {
    const size_t lds = update_fx_lds(P.Kp, P.ds);
    const void* fn =
        P.magic ? (P.ds == 64   ? reinterpret_cast<const void*>(k_update_fx<64, true>)
                   : P.ds == 32 ? reinterpret_cast<const void*>(k_update_fx<32, true>)
                                : reinterpret_cast<const void*>(k_update_fx<16, true>))
                : (P.ds == 64   ? reinterpret_cast<const void*>(k_update_fx<64, false>)
                   : P.ds == 32 ? reinterpret_cast<const void*>(k_update_fx<32, false>)
                                : reinterpret_cast<const void*>(k_update_fx<16, false>));
    if (P.ds != 64 && P.ds != 32 && P.ds != 16) return hipErrorInvalidValue;
    void* args[] = {const_cast<UpdateParams*>(&P)};
    return hipLaunchKernel(fn, dim3(P.nslice * P.wgps), dim3(kFxThreads), args, lds, s);
  }
  if (mode == UPDATE_TREE) {
    const long long waves = (P.n + kTreeRows - 1) / kTreeRows;
    const dim3 g((unsigned)((waves + kTreeWaves - 1) / kTreeWaves));
    const dim3 b(64 * kTreeWaves);
    const bool ex = P.cl_scale != nullptr;
    if (P.DP >= 256) {
      if (ex) hipLaunchKernelGGL((k_update_tree<4, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((k_update_tree<4, false>), g, b, 0, s, P);
    } else if (P.DP >= 128) {
      if (ex) hipLaunchKernelGGL((k_update_tree_rows<128, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((k_update_tree_rows<128, false>), g, b, 0, s, P);
    } else if (P.DP == 64) {
      if (ex) hipLaunchKernelGGL((k_update_tree_rows<64, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((k_update_tree_rows<64, false>), g, b, 0, s, P);
    } else if (P.DP == 32) {
      if (ex) hipLaunchKernelGGL((k_update_tree_rows<32, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((k_update_tree_rows<32, false>), g, b, 0, s, P);
    } else if (P.DP == 16) {
      if (ex) hipLaunchKernelGGL((k_update_tree_rows<16, true>), g, b, 0, s, P);
      else hipLaunchKernelGGL((k_update_tree_rows<16, false>), g, b, 0, s, P);
    } else {
      return hipErrorInvalidValue;
    }
  } else {
    hipLaunchKernelGGL(k_update, dim3((P.K * S + 3) / 4), dim3(256), 0, s, P);
  }
  return hipGetLastError();
}
hipError_t launch_finish(const FinishParams& F, hipStream_t s) {
  if (F.tree == 2 && !F.do_prep)
    hipLaunchKernelGGL(k_finish_waves,
                       dim3((F.K + kFinishClustersPerBlock - 1) / kFinishClustersPerBlock),
                       dim3(64 * kFinishClustersPerBlock), 0, s, F);
  else if (F.fixed)
    hipLaunchKernelGGL(k_finish_fx, dim3(F.K), dim3(64 * kFinishWaves), 0, s, F);
  else
    hipLaunchKernelGGL(k_finish, dim3(F.K), dim3(64 * kFinishWaves), 0, s, F);
  if (F.stats.state) hipLaunchKernelGGL(k_stats, dim3(1), dim3(kStatsThreads), 0, s, F.stats);
  return hipGetLastError();
}
hipError_t launch_absmax(const float* X, long long count, const double* xn64, long long n,
                         unsigned* amax, unsigned long long* qmax, hipStream_t s) {
  hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, X, count, xn64, n, amax, qmax);
  return hipGetLastError();
}
hipError_t launch_load_points(const LoadParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_load_points, dim3((unsigned)((P.n + 3) / 4)), dim3(256), 0, s, P);
  return hipGetLastError();
}
hipError_t launch_sum_f64(const double* v, long long n, double* out, double* part256,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_sum_f64_part, dim3(256), dim3(256), 0, s, v, n, part256);
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s, part256, 256LL, out);
  return hipGetLastError();
}
hipError_t launch_fill_i32(int* p, long long n, int v, hipStream_t s) {
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096)),
                     dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace kmc
