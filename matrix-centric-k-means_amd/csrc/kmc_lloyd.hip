// kmc_lloyd.hip -- gfx950 kernels of one Lloyd iteration (linear kernel).
//
// Per iteration, in stream order:
//   k_prep       centroid operands (-2c in fp32 / bf16 hi+lo), ||c||^2;
//                clears the refinement queue counters
//   k_assign     fused distance GEMM on MFMA (||c||^2 - 2 x.c accumulated
//                from the norm), running top-3 per point, margin band test;
//                writes labels and queues the points inside the band
//   k_refine2    band points with two candidates: both distances in the
//                reference's float arithmetic, first-min rule
//   k_refine_all band points with three or more candidates: all k distances
//   k_hist / k_scan_blocks / k_scan_clusters / k_scatter
//                stable counting sort of point ids by label (the one-hot
//                V matrix in CSR form: row offsets + column ids)
//   k_update     C = V X with per-cluster float sums in point order, plus the
//                FP64 inertia of the assignment
//   k_finish     per-centroid shift test, C <- C_next
//   k_stats      FP64 inertia / shift reductions, iteration counter, stop flag
//
// Every kernel returns at once when the device stop flag is set, so the host
// may queue iterations ahead of the convergence decision.
#include "kmc_device.h"
#include "kmc_kernels.h"

namespace kmc {

// ---------------------------------------------------------------------------
// k_prep: one wave per (padded) centroid row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prep(PrepParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= P.Kp) return;
  const size_t row = (size_t)c * P.DP;
  float s = 0.0f;
  for (int j = lane; j < P.DP; j += 64) {
    const float v = (c < P.K) ? P.C[row + j] : 0.0f;
    const float w = -2.0f * v;  // exact scaling
    s = fmaf(v, v, s);
    P.CT[(size_t)j * P.Kp + c] = v;
    if (P.prec == PREC_FP32) {
      P.Cw[row + j] = w;
    } else {
      const uint16_t hi = bf16_rne(w);
      P.Chi[row + j] = hi;
      if (P.prec == PREC_BF16X3) P.Clo[row + j] = bf16_rne(w - bf16_to_f32(hi));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) P.cn[c] = (c < P.K && P.alive[c]) ? s : kFar;
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
  constexpr int kVecPerRow = (PREC == PREC_FP32) ? DP / 4 : DP / 8;  // 16-B units
  const int total = cnt * kVecPerRow;
  for (int plane = 0; plane < T::kPlanes; ++plane) {
    const uint4* src;
    if (PREC == PREC_FP32)
      src = reinterpret_cast<const uint4*>(P.Cw + (size_t)c0 * DP);
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
template <int DP, int PREC>
struct XFrag {
  static constexpr int NF = (PREC == PREC_FP32) ? DP / 2 : 1;
  static constexpr int NH = (PREC == PREC_FP32) ? 1 : DP / 16;
  static constexpr int NL = (PREC == PREC_BF16X3) ? DP / 16 : 1;
  float f[NF];
  uint4 hi[NH];
  uint4 lo[NL];
};

template <int DP, int PREC>
__device__ __forceinline__ void load_xfrag(const AssignParams& P, long long prow, bool valid,
                                           int h, XFrag<DP, PREC>& x) {
  const size_t row = (size_t)(valid ? prow : 0) * DP;
  if (PREC == PREC_FP32) {
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

// One 32-centroid block on MFMA: all A fragments of the block are read from
// LDS first, then the MFMA chain runs from ||c||^2.
template <int DP, int PREC>
__device__ __forceinline__ kmc_f32x16 mfma_block(const XFrag<DP, PREC>& x, const float* cn_s,
                                                 const unsigned char* rows, int cnt, int cb,
                                                 int p, int h, float x2) {
  using T = AssignTraits<DP, PREC>;
  kmc_f32x16 acc;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 n4 = *reinterpret_cast<const float4*>(cn_s + cb + 8 * g + 4 * h);
    acc[4 * g + 0] = n4.x + x2;
    acc[4 * g + 1] = n4.y + x2;
    acc[4 * g + 2] = n4.z + x2;
    acc[4 * g + 3] = n4.w + x2;
  }
  const unsigned char* arow = rows + (size_t)(cb + p) * T::kRowBytes;
  if (PREC == PREC_FP32) {
    const float* a = reinterpret_cast<const float*>(arow);
    float4 a4[DP / 8];
#pragma unroll
    for (int q = 0; q < DP / 8; ++q) a4[q] = *reinterpret_cast<const float4*>(a + 8 * q + 4 * h);
#pragma unroll
    for (int q = 0; q < DP / 8; ++q) {
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[q].x, x.f[4 * q + 0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[q].y, x.f[4 * q + 1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[q].z, x.f[4 * q + 2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[q].w, x.f[4 * q + 3], acc, 0, 0, 0);
    }
  } else {
    const uint16_t* ah = reinterpret_cast<const uint16_t*>(arow);
    const uint16_t* al = reinterpret_cast<const uint16_t*>(
        rows + (size_t)cnt * T::kRowBytes + (size_t)(cb + p) * T::kRowBytes);
    uint4 a_hi[DP / 16], a_lo[PREC == PREC_BF16X3 ? DP / 16 : 1];
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) {
      a_hi[s] = *reinterpret_cast<const uint4*>(ah + 16 * s + 8 * h);
      if (PREC == PREC_BF16X3) a_lo[s] = *reinterpret_cast<const uint4*>(al + 16 * s + 8 * h);
    }
#pragma unroll
    for (int s = 0; s < DP / 16; ++s) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a_hi[s]),
                                                    __builtin_bit_cast(kmc_bf16x8, x.hi[s]),
                                                    acc, 0, 0, 0);
      if (PREC == PREC_BF16X3) {
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a_hi[s]),
                                                      __builtin_bit_cast(kmc_bf16x8, x.lo[s]),
                                                      acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a_lo[s]),
                                                      __builtin_bit_cast(kmc_bf16x8, x.hi[s]),
                                                      acc, 0, 0, 0);
      }
    }
  }
  return acc;
}

// Top-4 update over one block's 16 accumulator values. Each value carries
// its centroid index in the low mantissa bits (idx[r] | (v & ~mask)), so the
// running top-4 is four v_med3_f32 per value and the indices travel with
// the values. The values are full squared distances, so the perturbation is
// relative to the (small) distances of the leading candidates; it is part of
// the margin-band bound.
// Insert v into the running top-4 b1 <= b2 <= b3 <= b4 (the k-th smallest of
// {b1..b4, v} is med3(b_{k-1}, b_k, v)).
__device__ __forceinline__ void top4_push(float v, float& b1, float& b2, float& b3, float& b4) {
  b4 = __builtin_amdgcn_fmed3f(b3, b4, v);
  b3 = __builtin_amdgcn_fmed3f(b2, b3, v);
  b2 = __builtin_amdgcn_fmed3f(b1, b2, v);
  b1 = __builtin_amdgcn_fmed3f(-__builtin_inff(), b1, v);
}

__device__ __forceinline__ void push_block(const kmc_f32x16& acc, const uint32_t* idx,
                                           uint32_t keep, float& b1, float& b2, float& b3,
                                           float& b4) {
#pragma unroll
  for (int r = 0; r < 16; ++r)
    top4_push(__uint_as_float((__float_as_uint(acc[r]) & keep) | idx[r]), b1, b2, b3, b4);
}

// Work split: a workgroup of NW waves shares one LDS copy of the centroid
// operands (all of them when they fit, else chunk by chunk); each wave owns
// 32 points per tile, keeps their operand fragments in registers for the
// whole sweep over the centroids and, with PF, prefetches the fragments of
// its next tile while sweeping. The grid is persistent. (NW, PF) are chosen
// per problem by the autotuner (kmc_autotune); the LDS copy limits a CU to
// one workgroup when all centroids are resident, so NW sets the waves per
// SIMD.
template <int DP, int PREC, int NW, bool PF>
__global__ __launch_bounds__(64 * NW) void k_assign(AssignParams P) {
  if (*P.done) return;
  constexpr int kAssignWaves = NW;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ float red_s[kAssignWaves];
  __shared__ unsigned q2c_s, qac_s;  // this workgroup's refinement queue fill

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p = lane & 31;
  const int h = lane >> 5;
  constexpr int kTile = 32 * kAssignWaves;  // points per workgroup tile
  if (threadIdx.x == 0) {
    q2c_s = 0u;
    qac_s = 0u;
  }
  // each workgroup appends to its own queue region (no global atomics)
  int* const q2 = P.q2 + 4 * (size_t)blockIdx.x * P.qstride;
  int* const qa = P.qa + (size_t)blockIdx.x * P.qstride;

  // max ||c|| over the live centroids (for the error bound)
  float cm = 0.0f;
  for (int c = threadIdx.x; c < P.K; c += blockDim.x) {
    const float v = P.cn[c];
    if (v < kFarTest) cm = fmaxf(cm, v);  // excluded clusters carry kFar
  }
  cm = wave_max_f32(cm);
  if (lane == 0) red_s[wave] = cm;
  __syncthreads();
  float cm_all = 0.0f;
#pragma unroll
  for (int w = 0; w < kAssignWaves; ++w) cm_all = fmaxf(cm_all, red_s[w]);
  const float cmax = sqrtf(cm_all);
  const float Ecm = P.beta * cmax * cmax;
  const uint32_t keep = ~P.idx_mask;

  const bool resident = P.kchunk >= P.Kp;
  if (resident) {
    stage_centroids<DP, PREC>(P, smem, 0, P.Kp);
    __syncthreads();
  }

  const long long ntiles = (P.n + kTile - 1) / kTile;
  // prefetch the next tile only while two fragment sets fit the registers
  constexpr int kFragRegs = PREC == PREC_FP32 ? DP / 2 : (PREC == PREC_BF16X3 ? DP / 2 : DP / 4);
  constexpr bool kPrefetch = PF && kFragRegs <= 64;
  XFrag<DP, PREC> cur, nxt;
  long long tile = blockIdx.x;
  if (tile < ntiles) {
    const long long prow = tile * kTile + wave * 32 + p;
    load_xfrag<DP, PREC>(P, prow, prow < P.n, h, cur);
  }
  for (; tile < ntiles; tile += gridDim.x) {
    const long long prow = tile * kTile + wave * 32 + p;
    const bool valid = prow < P.n;
    const long long ntile = tile + gridDim.x;
    if (kPrefetch && ntile < ntiles) {
      const long long nrow = ntile * kTile + wave * 32 + p;
      load_xfrag<DP, PREC>(P, nrow, nrow < P.n, h, nxt);
    }

    float b1 = kFar, b2 = kFar, b3 = kFar, b4 = kFar;
    const float x2 = valid ? P.xn2[prow] : 0.0f;
    for (int c0 = 0; c0 < P.Kp; c0 += P.kchunk) {
      const int cnt = min(P.kchunk, P.Kp - c0);
      if (!resident) {
        __syncthreads();
        stage_centroids<DP, PREC>(P, smem, c0, cnt);
        __syncthreads();
      }
      const float* cn_s = reinterpret_cast<const float*>(smem);
      const unsigned char* rows = smem + (size_t)cnt * 4;
      // centroid index of accumulator r of the first block of this chunk
      uint32_t idx0[16], idx1[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        idx0[r] = (uint32_t)(c0 + 4 * h + 8 * (r >> 2) + (r & 3));
        idx1[r] = idx0[r] + 32u;
      }
      // blocks in pairs: the second block's MFMAs run while the first
      // block's values go through the top-3
      for (int cb = 0; cb < cnt; cb += 64) {
        const kmc_f32x16 a0 = mfma_block<DP, PREC>(cur, cn_s, rows, cnt, cb, p, h, x2);
        const kmc_f32x16 a1 = mfma_block<DP, PREC>(cur, cn_s, rows, cnt, cb + 32, p, h, x2);
        push_block(a0, idx0, keep, b1, b2, b3, b4);
        push_block(a1, idx1, keep, b1, b2, b3, b4);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          idx0[r] += 64u;
          idx1[r] += 64u;
        }
      }
    }

    // ---- merge the two k-slices' halves of the centroid rows ----
    {
      const float o1 = __shfl_xor(b1, 32, 64);
      const float o2 = __shfl_xor(b2, 32, 64);
      const float o3 = __shfl_xor(b3, 32, 64);
      const float o4 = __shfl_xor(b4, 32, 64);
      top4_push(o1, b1, b2, b3, b4);
      top4_push(o2, b1, b2, b3, b4);
      top4_push(o3, b1, b2, b3, b4);
      top4_push(o4, b1, b2, b3, b4);
    }
    const int i1 = (int)(__float_as_uint(b1) & P.idx_mask);
    const int i2 = (int)(__float_as_uint(b2) & P.idx_mask);
    const int i3 = (int)(__float_as_uint(b3) & P.idx_mask);

    // ---- margin band test and output ----
    // b1 <= b2 <= b3 are full squared distances (plus the packed index bits);
    // E bounds the MFMA error of each, gref the reference's own float error
    // plus the packing perturbation, both relative to the distances.
    bool flag2 = false, flagA = false, few3 = false;
    if (h == 0 && valid) {
      const float E = P.alpha * sqrtf(x2) * cmax + Ecm;
      const float twoE = 2.0f * E;
      const float b1p = fmaxf(b1 + E, 0.0f);
      // clear(b): every centroid ranked at b or later loses to the first
      auto clear = [&](float b) {
        return (b >= kFarTest) || (b - b1 > twoE + P.gref * (fmaxf(b + E, 0.0f) + b1p));
      };
      const bool uniq = clear(b2);
      const bool two = clear(b3);
      const bool three = clear(b4);
      // a point without any finite distance goes to the all-k pass, which
      // applies the first-min rule
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
      if (lane == leader) base = atomicAdd(&q2c_s, (unsigned)__popcll(m2));
      base = __shfl(base, leader, 64);
      if (flag2) {
        const unsigned pos = base + mbcnt64(m2);
        q2[4 * (size_t)pos + 0] = (int)prow;
        q2[4 * (size_t)pos + 1] = i1;
        q2[4 * (size_t)pos + 2] = i2;
        q2[4 * (size_t)pos + 3] = few3 ? i3 : -1;
      }
    }
    if (mA) {
      const int leader = __ffsll((unsigned long long)mA) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(&qac_s, (unsigned)__popcll(mA));
      base = __shfl(base, leader, 64);
      if (flagA) qa[base + mbcnt64(mA)] = (int)prow;
    }
    if (kPrefetch) {
      cur = nxt;
    } else if (ntile < ntiles) {
      const long long nrow = ntile * kTile + wave * 32 + p;
      load_xfrag<DP, PREC>(P, nrow, nrow < P.n, h, cur);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    P.q2cnt[blockIdx.x] = q2c_s;
    P.qacnt[blockIdx.x] = qac_s;
  }
}

// ---------------------------------------------------------------------------
// Refinement: reference float arithmetic, first-min rule.
// ---------------------------------------------------------------------------
// Band points with two or three candidates (the rest provably lose): each
// candidate's distance in the reference's float arithmetic, first-min rule
// (the smallest distance, the lowest index among equal ones). Block r
// handles queue region r (written by assign workgroup r).
__global__ __launch_bounds__(256) void k_refine2(RefineParams P) {
  if (*P.done) return;
  for (int r = blockIdx.x; r < P.nregions; r += gridDim.x) {
    const unsigned cnt = P.q2cnt[r];
    const int* q2 = P.q2 + 4 * (size_t)r * P.qstride;
    for (unsigned e = threadIdx.x; e < cnt; e += blockDim.x) {
      const int4 q = *reinterpret_cast<const int4*>(q2 + 4 * (size_t)e);
      const int i = q.x;
      const float* x = P.X + (size_t)i * P.DP;
      int best = q.y;
      float rbest = direct_dist_f32(x, P.C + (size_t)q.y * P.DP, P.D);
      const float rb = direct_dist_f32(x, P.C + (size_t)q.z * P.DP, P.D);
      if (rb < rbest || (rb == rbest && q.z < best)) { rbest = rb; best = q.z; }
      if (q.w >= 0) {
        const float rc = direct_dist_f32(x, P.C + (size_t)q.w * P.DP, P.D);
        if (rc < rbest || (rc == rbest && q.w < best)) { rbest = rc; best = q.w; }
      }
      P.labels[i] = best;
    }
  }
}

constexpr int kRefineSplit = 4;

// One wave per queued point; lane c evaluates centroids c, c+64, c+128,
// c+192 from the transposed centroids (coalesced across lanes). Dimensions go
// in chunks of 16: the 64 centroid values and 16 point values of a chunk are
// loaded first, then each of the four distances advances in the reference's
// float order (padding dimensions are zero in X and C and add exact zeros).
// First-min across lanes.
__global__ __launch_bounds__(256) void k_refine_all(RefineParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  // kRefineSplit blocks share one region; wave w of block part q takes
  // entries w*kRefineSplit + q, stepping by 4*kRefineSplit
  const int part = blockIdx.x % kRefineSplit;
  const int wave = (threadIdx.x >> 6) * kRefineSplit + part;
  for (int r = blockIdx.x / kRefineSplit; r < P.nregions; r += gridDim.x / kRefineSplit) {
    const unsigned cnt = P.qacnt[r];
    const int* qa = P.qa + (size_t)r * P.qstride;
    for (unsigned e = wave; e < cnt; e += 4 * kRefineSplit) {
      const int i = qa[e];
      const float* x = P.X + (size_t)i * P.DP;
      float best = __builtin_inff();
      int bi = 0x7fffffff;
      for (int cg = 0; cg < P.K; cg += 256) {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        int cc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) cc[u] = min(cg + 64 * u + lane, P.Kp - 1);
        for (int j0 = 0; j0 < P.DP; j0 += 16) {
          float xv[16], cv[4][16];
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) xv[jj] = x[j0 + jj];
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)
              cv[u][jj] = P.CT[(size_t)(j0 + jj) * P.Kp + cc[u]];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
#pragma clang fp contract(off)
              const float t = xv[jj] - cv[u][jj];
              const float q = t * t;
              acc[u] = acc[u] + q;
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c = cg + 64 * u + lane;
          if (c < P.K && P.alive[c] && acc[u] < best) { best = acc[u]; bi = c; }
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const float ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov < best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (lane == 0) P.labels[i] = (bi == 0x7fffffff) ? 0 : bi;
    }
  }
}

// ---------------------------------------------------------------------------
// Stable counting sort of point ids by label. One wave per block of CH points.
// bcount is cluster-major: bcount[c * nblk + b].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_hist(SortParams P) {
  if (*P.done) return;
  extern __shared__ int cnt_s[];
  const int b = blockIdx.x;
  for (int c = threadIdx.x; c < P.K; c += 64) cnt_s[c] = 0;
  __syncthreads();
  const long long lo = (long long)b * P.CH;
  const long long hi = min(lo + P.CH, (long long)P.n);
  for (long long i = lo + threadIdx.x; i < hi; i += 64) atomicAdd(&cnt_s[P.labels[i]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < P.K; c += 64) P.bcount[(size_t)c * P.nblk + b] = cnt_s[c];
}

// Exclusive scan over blocks for each cluster (one wave per cluster).
__global__ __launch_bounds__(256) void k_scan_blocks(SortParams P) {
  if (*P.done) return;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= P.K) return;
  int* a = P.bcount + (size_t)c * P.nblk;
  int carry = 0;
  for (int base = 0; base < P.nblk; base += 64) {
    const int idx = base + lane;
    const int v = idx < P.nblk ? a[idx] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (idx < P.nblk) a[idx] = carry + incl - v;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) P.count[c] = carry;
}

// Exclusive scan of the cluster sizes (one workgroup).
__global__ __launch_bounds__(1024) void k_scan_clusters(SortParams P) {
  if (*P.done) return;
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < P.K; base += 1024) {
    const int idx = base + threadIdx.x;
    const int v = idx < P.K ? P.count[idx] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; ++w) woff += wsum[w];
    const int carry = carry_s;
    if (idx < P.K) P.start[idx] = carry + woff + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry_s = carry + woff + incl;
    __syncthreads();
  }
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
  for (int c = lane; c < P.K; c += 64)
    cur_s[c] = P.start[c] + P.bcount[(size_t)c * P.nblk + b];
  __syncthreads();
  const long long lo = (long long)b * P.CH;
  const long long hi = min(lo + P.CH, (long long)P.n);
  const uint64_t lt = lanemask_lt();
  for (long long i0 = lo; i0 < hi; i0 += 64) {
    const long long i = i0 + lane;
    const bool valid = i < hi;
    const int lab = valid ? P.labels[i] : 0;
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
// Tree update: the sorted member list is cut into chunks of kTreeRows rows;
// one wave per (chunk, 64-dim slab) sums its rows segment by segment (a
// segment = the part of one cluster inside the chunk) in FP64 and writes one
// partial per segment at index cluster + chunk (injective, because clusters
// occupy increasing chunk ranges). k_update_combine adds a cluster's partials
// in chunk order. Deterministic; the FP64 sums are rounded once, so centroids
// are at least as accurate as the reference's float sums (not bit-identical).
// ---------------------------------------------------------------------------
template <int VW>
__global__ __launch_bounds__(256) void k_update_tree(UpdateParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  const long long q = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long nchunks = (P.n + kTreeRows - 1) / kTreeRows;
  if (q >= nchunks) return;
  // lane owns dimensions [VW*lane, VW*lane + VW) of every row (whole rows per
  // wave: one contiguous 4*VW*64-byte read per row)
  const int j0 = VW * lane;
  const bool act = j0 < P.DP;
  const long long p0 = q * kTreeRows;
  const long long p1 = min(p0 + (long long)kTreeRows, P.n);
  const float* X = P.X + (act ? j0 : 0);
  double in = 0.0;
  long long pos = p0;
  while (pos < p1) {
    const int c = P.labels[P.members[pos]];
    const long long seg_end = min(p1, (long long)P.start[c] + P.count[c]);
    double cold[VW], sum[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) {
      cold[v] = act ? (double)P.C[(size_t)c * P.DP + j0 + v] : 0.0;
      sum[v] = 0.0;
    }
    long long t = pos;
    constexpr int U = 16;
    for (; t + U <= seg_end; t += U) {
      float xv[U][VW];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float* r = X + (size_t)P.members[t + u] * P.DP;
        if (VW == 4) {
          const float4 f = *reinterpret_cast<const float4*>(r);
          xv[u][0] = f.x; xv[u][1 % VW] = f.y; xv[u][2 % VW] = f.z; xv[u][3 % VW] = f.w;
        } else if (VW == 2) {
          const float2 f = *reinterpret_cast<const float2*>(r);
          xv[u][0] = f.x; xv[u][1 % VW] = f.y;
        } else {
          xv[u][0] = r[0];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int v = 0; v < VW; ++v) {
          const double xd = (double)xv[u][v];
          sum[v] += xd;
          const double df = xd - cold[v];
          in = fma(df, df, in);
        }
    }
    for (; t < seg_end; ++t) {
      const float* r = X + (size_t)P.members[t] * P.DP;
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        const double xd = (double)r[v];
        sum[v] += xd;
        const double df = xd - cold[v];
        in = fma(df, df, in);
      }
    }
    if (act) {
#pragma unroll
      for (int v = 0; v < VW; ++v) P.tpart[(size_t)(c + q) * P.DP + j0 + v] = sum[v];
    }
    pos = seg_end;
  }
  if (!act) in = 0.0;
  in = wave_sum_f64(in);
  if (lane == 0) P.ipart[q] = in;
}

// ---------------------------------------------------------------------------
// k_finish: one wave per centroid. With the tree update it first forms
// c_new = (partials summed in chunk order) / count, rounded once (an empty
// cluster becomes the zero vector). Then the float shift in the reference's
// order on lane 0 (comparisons/CPU_Kmeans/src/kmeans.cpp:225-234), the FP64
// shift, and C <- C_new, lane-parallel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_finish(FinishParams P) {
  if (*P.done) return;
  __shared__ float new_s[4][kMaxDP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x * 4 + wave;
  if (c >= P.K) return;
  float* cr = P.C + (size_t)c * P.DP;
  float* nrow = P.Cnext + (size_t)c * P.DP;
  float* ns = new_s[wave];
  const int cnt = P.count[c];
  if (P.tree) {
    long long q0 = 0, q1 = -1;
    if (cnt > 0) {
      q0 = (long long)P.start[c] / kTreeRows;
      q1 = ((long long)P.start[c] + cnt - 1) / kTreeRows;
    }
    for (int j = lane; j < P.D; j += 64) {
      // partials summed in chunk order; 8 loads in flight per step
      const double* tp = P.tpart + (size_t)c * P.DP + j;
      double sum = 0.0;
      long long q = q0;
      for (; q + 8 <= q1 + 1; q += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = tp[(size_t)(q + u) * P.DP];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum += v[u];
      }
      for (; q <= q1; ++q) sum += tp[(size_t)q * P.DP];
      const float v = (float)(sum / (double)(cnt > 1 ? cnt : 1));
      ns[j] = v;
      nrow[j] = v;
    }
  } else {
    for (int j = lane; j < P.D; j += 64) ns[j] = nrow[j];
  }
  __builtin_amdgcn_wave_barrier();
  double s64 = 0.0;
  for (int j = lane; j < P.D; j += 64) {
    const double td = (double)ns[j] - (double)cr[j];
    s64 = fma(td, td, s64);
  }
  s64 = wave_sum_f64(s64);
  if (lane == 0) {
    float s = 0.0f;
    for (int j = 0; j < P.D; ++j) {
#pragma clang fp contract(off)
      const float t = ns[j] - cr[j];
      const float q = t * t;
      s = s + q;
    }
    P.conv[c] = (__fsqrt_rn(s) > P.tol) ? 0 : 1;
    P.shift2[c] = s64;
    if (P.exclude_empty && cnt == 0) P.alive[c] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  for (int j = lane; j < P.D; j += 64) cr[j] = ns[j];
}

// ---------------------------------------------------------------------------
// k_stats: one workgroup. Deterministic FP64 reductions and the stop flag.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stats(StatsParams P) {
  if (*P.done) return;
  __shared__ double sd[256];
  __shared__ double sm[256];
  __shared__ int sc[256];
  __shared__ unsigned long long sq2[256], sqa[256];
  double in = 0.0, mx = 0.0;
  int all = 1;
  unsigned long long r2 = 0, ra = 0;
  for (int r = threadIdx.x; r < P.nregions; r += 256) {
    r2 += P.q2cnt[r];
    ra += P.qacnt[r];
  }
  for (int i = threadIdx.x; i < P.nparts; i += 256) in += P.ipart[i];
  for (int c = threadIdx.x; c < P.K; c += 256) {
    mx = fmax(mx, P.shift2[c]);
    all &= P.conv[c];
  }
  sd[threadIdx.x] = in;
  sm[threadIdx.x] = mx;
  sc[threadIdx.x] = all;
  sq2[threadIdx.x] = r2;
  sqa[threadIdx.x] = ra;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sd[threadIdx.x] += sd[threadIdx.x + o];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
      sc[threadIdx.x] &= sc[threadIdx.x + o];
      sq2[threadIdx.x] += sq2[threadIdx.x + o];
      sqa[threadIdx.x] += sqa[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    DevState* s = P.state;
    const unsigned long long it = s->iter + 1;
    s->iter = it;
    s->inertia = sd[0];
    s->max_shift2 = sm[0];
    s->refined_two = sq2[0];
    s->refined_all = sqa[0];
    const double score = sd[0] - s->sum_x2;
    if (P.mode == CONV_SHIFT && it >= 2 && sc[0]) {
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
    if (it >= s->maxiter) s->done = 1;
  }
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
  float s = 0.0f;
  double s64 = 0.0;
  for (int j = lane; j < P.DP; j += 64) {
    const float v = j < P.D ? src[j] : 0.0f;
    P.X[(size_t)i * P.DP + j] = v;
    s = fmaf(v, v, s);
    s64 += (double)v * (double)v;
    if (P.Xhi) {
      const uint16_t hi = bf16_rne(v);
      P.Xhi[(size_t)i * P.DP + j] = hi;
      if (P.Xlo) P.Xlo[(size_t)i * P.DP + j] = bf16_rne(v - bf16_to_f32(hi));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    s64 += __shfl_xor(s64, o, 64);
  }
  if (lane == 0) {
    P.xn2[i] = s;
    P.xn2_64[i] = s64;
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

// Launch variants of k_assign: (waves per workgroup, prefetch).
struct VariantShape {
  int nw;
  bool pf;
};
constexpr VariantShape kVariants[] = {{8, true}, {8, false}, {12, false}, {16, false}, {4, true}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

template <int DP, int PREC, int V>
static const void* kernel_of() {
  return reinterpret_cast<const void*>(k_assign<DP, PREC, kVariants[V].nw, kVariants[V].pf>);
}

template <int DP, int PREC>
static const void* kernel_of_v(int v) {
  switch (v) {
    case 0: return kernel_of<DP, PREC, 0>();
    case 1: return kernel_of<DP, PREC, 1>();
    case 2: return kernel_of<DP, PREC, 2>();
    case 3: return kernel_of<DP, PREC, 3>();
    case 4: return kernel_of<DP, PREC, 4>();
    default: return nullptr;
  }
}

template <int PREC>
static const void* kernel_of_dv(int DP, int v) {
  switch (DP) {
    case 16: return kernel_of_v<16, PREC>(v);
    case 32: return kernel_of_v<32, PREC>(v);
    case 64: return kernel_of_v<64, PREC>(v);
    case 128: return kernel_of_v<128, PREC>(v);
    case 256: return kernel_of_v<256, PREC>(v);
    default: return nullptr;
  }
}

const void* assign_kernel_ptr(int DP, int prec, int v) {
  switch (prec) {
    case PREC_FP32: return kernel_of_dv<PREC_FP32>(DP, v);
    case PREC_BF16X3: return kernel_of_dv<PREC_BF16X3>(DP, v);
    case PREC_BF16: return kernel_of_dv<PREC_BF16>(DP, v);
    default: return nullptr;
  }
}

int assign_variants() { return kNumVariants; }
int assign_threads(int v) { return 64 * kVariants[v].nw; }
int assign_tile(int v) { return 32 * kVariants[v].nw; }

hipError_t launch_assign(int DP, int prec, int v, const AssignParams& P, int grid, size_t lds,
                         hipStream_t s) {
  if (v < 0 || v >= kNumVariants) return hipErrorInvalidValue;
  const void* fn = assign_kernel_ptr(DP, prec, v);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<AssignParams*>(&P)};
  return hipLaunchKernel(fn, dim3(grid), dim3(assign_threads(v)), args, lds, s);
}

hipError_t launch_prep(const PrepParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_prep, dim3((P.Kp + 3) / 4), dim3(256), 0, s, P);
  return hipGetLastError();
}
hipError_t launch_refine(const RefineParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_refine2, dim3(P.nregions), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_refine_all, dim3(P.nregions * kRefineSplit), dim3(256), 0, s, P);
  return hipGetLastError();
}
hipError_t launch_sort(const SortParams& P, hipStream_t s) {
  const size_t lds = (size_t)P.K * sizeof(int);
  hipLaunchKernelGGL(k_hist, dim3(P.nblk), dim3(64), lds, s, P);
  hipLaunchKernelGGL(k_scan_blocks, dim3((P.K + 3) / 4), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_scan_clusters, dim3(1), dim3(1024), 0, s, P);
  hipLaunchKernelGGL(k_scatter, dim3(P.nblk), dim3(64), lds, s, P);
  return hipGetLastError();
}
hipError_t launch_update(const UpdateParams& P, int mode, hipStream_t s) {
  const int S = (P.D + 63) / 64;
  if (mode == UPDATE_TREE) {
    const long long waves = (P.n + kTreeRows - 1) / kTreeRows;
    const dim3 g((unsigned)((waves + 3) / 4));
    if (P.DP >= 256) hipLaunchKernelGGL(k_update_tree<4>, g, dim3(256), 0, s, P);
    else if (P.DP >= 128) hipLaunchKernelGGL(k_update_tree<2>, g, dim3(256), 0, s, P);
    else hipLaunchKernelGGL(k_update_tree<1>, g, dim3(256), 0, s, P);
  } else {
    hipLaunchKernelGGL(k_update, dim3((P.K * S + 3) / 4), dim3(256), 0, s, P);
  }
  return hipGetLastError();
}
hipError_t launch_finish(const FinishParams& F, const StatsParams& S, hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3((F.K + 3) / 4), dim3(256), 0, s, F);
  hipLaunchKernelGGL(k_stats, dim3(1), dim3(256), 0, s, S);
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
