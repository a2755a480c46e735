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
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    P.qcnt[0] = 0u;
    P.qcnt[1] = 0u;
  }
  if (c >= P.Kp) return;
  const size_t row = (size_t)c * P.DP;
  float s = 0.0f;
  for (int j = lane; j < P.DP; j += 64) {
    const float v = (c < P.K) ? P.C[row + j] : 0.0f;
    const float w = -2.0f * v;  // exact scaling
    s = fmaf(v, v, s);
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
  if (lane == 0) P.cn[c] = (c < P.K && P.alive[c]) ? s : __builtin_inff();
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

template <int DP, int PREC>
__global__ __launch_bounds__(256) void k_assign(AssignParams P) {
  if (*P.done) return;
  using T = AssignTraits<DP, PREC>;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ float red_s[4];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p = lane & 31;
  const int h = lane >> 5;
  constexpr int kTile = 128;  // points per workgroup tile

  // max ||c|| over the real centroids (for the error bound)
  float cm = 0.0f;
  for (int c = threadIdx.x; c < P.K; c += blockDim.x) {
    const float v = P.cn[c];
    if (v < __builtin_inff()) cm = fmaxf(cm, v);  // excluded clusters are +inf
  }
  cm = wave_max_f32(cm);
  if (lane == 0) red_s[wave] = cm;
  __syncthreads();
  const float cmax = sqrtf(fmaxf(fmaxf(red_s[0], red_s[1]), fmaxf(red_s[2], red_s[3])));
  const float Ecm = P.beta * cmax * cmax;

  const bool resident = P.kchunk >= P.Kp;
  if (resident) {
    stage_centroids<DP, PREC>(P, smem, 0, P.Kp);
    __syncthreads();
  }

  const long long ntiles = (P.n + kTile - 1) / kTile;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long prow = tile * kTile + wave * 32 + p;
    const bool valid = prow < P.n;

    // ---- B operand fragments (this lane's point, its k-slice) ----
    constexpr int kFragsF = DP / 2;   // fp32 steps
    constexpr int kFragsB = DP / 16;  // bf16 steps
    float xf[PREC == PREC_FP32 ? kFragsF : 1];
    uint4 xh[PREC == PREC_FP32 ? 1 : kFragsB];
    uint4 xl[PREC == PREC_BF16X3 ? kFragsB : 1];
    if (PREC == PREC_FP32) {
      const float* xr = P.X + (size_t)(valid ? prow : 0) * DP;
#pragma unroll
      for (int q = 0; q < DP / 8; ++q) {
        float4 v = valid ? *reinterpret_cast<const float4*>(xr + 8 * q + 4 * h)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
        xf[4 * q + 0] = v.x;
        xf[4 * q + 1] = v.y;
        xf[4 * q + 2] = v.z;
        xf[4 * q + 3] = v.w;
      }
    } else {
      const uint16_t* xr = P.Xhi + (size_t)(valid ? prow : 0) * DP;
#pragma unroll
      for (int s = 0; s < kFragsB; ++s)
        xh[s] = valid ? *reinterpret_cast<const uint4*>(xr + 16 * s + 8 * h)
                      : make_uint4(0u, 0u, 0u, 0u);
      if (PREC == PREC_BF16X3) {
        const uint16_t* xq = P.Xlo + (size_t)(valid ? prow : 0) * DP;
#pragma unroll
        for (int s = 0; s < kFragsB; ++s)
          xl[s] = valid ? *reinterpret_cast<const uint4*>(xq + 16 * s + 8 * h)
                        : make_uint4(0u, 0u, 0u, 0u);
      }
    }

    Top3 st;
    st.init();

    for (int c0 = 0; c0 < P.Kp; c0 += P.kchunk) {
      const int cnt = min(P.kchunk, P.Kp - c0);
      if (!resident) {
        __syncthreads();
        stage_centroids<DP, PREC>(P, smem, c0, cnt);
        __syncthreads();
      }
      const float* cn_s = reinterpret_cast<const float*>(smem);
      const unsigned char* rows = smem + (size_t)cnt * 4;

      for (int cb = 0; cb < cnt; cb += 32) {
        kmc_f32x16 acc;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const float4 n4 = *reinterpret_cast<const float4*>(cn_s + cb + 8 * g + 4 * h);
          acc[4 * g + 0] = n4.x;
          acc[4 * g + 1] = n4.y;
          acc[4 * g + 2] = n4.z;
          acc[4 * g + 3] = n4.w;
        }
        const unsigned char* arow = rows + (size_t)(cb + p) * T::kRowBytes;
        if (PREC == PREC_FP32) {
          const float* a = reinterpret_cast<const float*>(arow);
#pragma unroll
          for (int q = 0; q < DP / 8; ++q) {
            const float4 a4 = *reinterpret_cast<const float4*>(a + 8 * q + 4 * h);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.x, xf[4 * q + 0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.y, xf[4 * q + 1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.z, xf[4 * q + 2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.w, xf[4 * q + 3], acc, 0, 0, 0);
          }
        } else {
          const uint16_t* ah = reinterpret_cast<const uint16_t*>(arow);
          const uint16_t* al = reinterpret_cast<const uint16_t*>(
              rows + (size_t)cnt * T::kRowBytes + (size_t)(cb + p) * T::kRowBytes);
#pragma unroll
          for (int s = 0; s < kFragsB; ++s) {
            const uint4 a_hi = *reinterpret_cast<const uint4*>(ah + 16 * s + 8 * h);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(kmc_bf16x8, a_hi), __builtin_bit_cast(kmc_bf16x8, xh[s]),
                acc, 0, 0, 0);
            if (PREC == PREC_BF16X3) {
              const uint4 a_lo = *reinterpret_cast<const uint4*>(al + 16 * s + 8 * h);
              acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(kmc_bf16x8, a_hi), __builtin_bit_cast(kmc_bf16x8, xl[s]),
                  acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(kmc_bf16x8, a_lo), __builtin_bit_cast(kmc_bf16x8, xh[s]),
                  acc, 0, 0, 0);
            }
          }
        }
        const int cbase = c0 + cb + 4 * h;
#pragma unroll
        for (int r = 0; r < 16; ++r) st.push(acc[r], cbase + 8 * (r >> 2) + (r & 3));
      }
    }

    // ---- merge the two k-slices' halves of the centroid rows ----
    {
      const float o1 = __shfl_xor(st.b1, 32, 64);
      const float o2 = __shfl_xor(st.b2, 32, 64);
      const float o3 = __shfl_xor(st.b3, 32, 64);
      const int j1 = __shfl_xor(st.i1, 32, 64);
      const int j2 = __shfl_xor(st.i2, 32, 64);
      if (h == 0) {
        st.push(o1, j1);
        st.push(o2, j2);
        st.push(o3, 0x7fffffff);
      }
    }

    // ---- margin band test and output ----
    bool flag2 = false, flagA = false;
    if (h == 0 && valid) {
      const float x2 = P.xn2[prow];
      const float E = P.alpha * sqrtf(x2) * cmax + Ecm;
      const float twoE = 2.0f * E;
      const float b1p = fmaxf(st.b1 + x2 + E, 0.0f);
      const bool uniq = (st.b2 == __builtin_inff()) ||
                        (st.b2 - st.b1 > twoE + P.gref * (fmaxf(st.b2 + x2 + E, 0.0f) + b1p));
      const bool two = (st.b3 == __builtin_inff()) ||
                       (st.b3 - st.b1 > twoE + P.gref * (fmaxf(st.b3 + x2 + E, 0.0f) + b1p));
      // A point whose distances are all non-finite keeps the sentinel index;
      // it goes to the all-k pass, which applies the first-min rule.
      const bool sane = st.i1 < P.K;
      P.labels[prow] = sane ? st.i1 : 0;
      flag2 = sane && !uniq && two;
      flagA = !sane || (!uniq && !two);
    }
    const uint64_t m2 = __ballot(flag2);
    const uint64_t mA = __ballot(flagA);
    if (m2) {
      const int leader = __ffsll((unsigned long long)m2) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(&P.qcnt[0], (unsigned)__popcll(m2));
      base = __shfl(base, leader, 64);
      if (flag2) {
        const unsigned pos = base + mbcnt64(m2);
        P.q2[3 * (size_t)pos + 0] = (int)prow;
        P.q2[3 * (size_t)pos + 1] = st.i1;
        P.q2[3 * (size_t)pos + 2] = st.i2;
      }
    }
    if (mA) {
      const int leader = __ffsll((unsigned long long)mA) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(&P.qcnt[1], (unsigned)__popcll(mA));
      base = __shfl(base, leader, 64);
      if (flagA) P.qa[base + mbcnt64(mA)] = (int)prow;
    }
  }
}

// ---------------------------------------------------------------------------
// Refinement: reference float arithmetic, first-min rule.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_refine2(RefineParams P) {
  if (*P.done) return;
  const unsigned cnt = P.qcnt[0];
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += gridDim.x * blockDim.x) {
    const int i = P.q2[3 * (size_t)e + 0];
    const int a = P.q2[3 * (size_t)e + 1];
    const int b = P.q2[3 * (size_t)e + 2];
    const float* x = P.X + (size_t)i * P.DP;
    const float ra = direct_dist_f32(x, P.C + (size_t)a * P.DP, P.D);
    const float rb = direct_dist_f32(x, P.C + (size_t)b * P.DP, P.D);
    P.labels[i] = (rb < ra || (rb == ra && b < a)) ? b : a;
  }
}

__global__ __launch_bounds__(256) void k_refine_all(RefineParams P) {
  if (*P.done) return;
  const unsigned cnt = P.qcnt[1];
  const int lane = threadIdx.x & 63;
  const unsigned nw = gridDim.x * (blockDim.x >> 6);
  for (unsigned e = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); e < cnt; e += nw) {
    const int i = P.qa[e];
    const float* x = P.X + (size_t)i * P.DP;
    float best = __builtin_inff();
    int bi = 0x7fffffff;
    for (int c = lane; c < P.K; c += 64) {
      if (!P.alive[c]) continue;
      const float r = direct_dist_f32(x, P.C + (size_t)c * P.DP, P.D);
      if (r < best) { best = r; bi = c; }
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
// visited in index order, 64 at a time, and equal labels inside one batch are
// ranked by lane order, so the sort is stable.
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
    const int lab = valid ? P.labels[i] : -1;
    uint64_t todo = __ballot(valid);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int L = __builtin_amdgcn_readlane(lab, leader);
      const bool mine = valid && lab == L;
      const uint64_t m = __ballot(mine);
      if (mine) P.members[cur_s[L] + __popcll(m & lt)] = (int)i;
      __builtin_amdgcn_wave_barrier();
      if (lane == leader) cur_s[L] += __popcll(m);
      __builtin_amdgcn_wave_barrier();
      todo &= ~m;
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
  const int cnt = P.count[c];
  const int* mem = P.members + P.start[c];
  const float cold = act ? P.C[(size_t)c * P.DP + j] : 0.0f;
  const float* X = P.X + (act ? j : 0);
  float sum = 0.0f;
  double in = 0.0;
  int t = 0;
  constexpr int U = 16;
  for (; t + U <= cnt; t += U) {
    float xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = X[(size_t)mem[t + u] * P.DP];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma clang fp contract(off)
      sum = sum + xv[u];
      const float df = xv[u] - cold;
      in += (double)df * (double)df;
    }
  }
  for (; t < cnt; ++t) {
#pragma clang fp contract(off)
    const float xv = X[(size_t)mem[t] * P.DP];
    sum = sum + xv;
    const float df = xv - cold;
    in += (double)df * (double)df;
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
// k_finish: one thread per centroid. Shift in the reference's float order
// (comparisons/CPU_Kmeans/src/kmeans.cpp:225-234) and in FP64; C <- C_next.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_finish(FinishParams P) {
  if (*P.done) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.K) return;
  float* cr = P.C + (size_t)c * P.DP;
  const float* nr = P.Cnext + (size_t)c * P.DP;
  float s = 0.0f;
  double s64 = 0.0;
  for (int j = 0; j < P.D; ++j) {
#pragma clang fp contract(off)
    const float a = nr[j], b = cr[j];
    const float t = a - b;
    const float q = t * t;
    s = s + q;
    const double td = (double)a - (double)b;
    s64 += td * td;
  }
  P.conv[c] = (__fsqrt_rn(s) > P.tol) ? 0 : 1;
  P.shift2[c] = s64;
  for (int j = 0; j < P.D; ++j) cr[j] = nr[j];
  if (P.exclude_empty && P.count[c] == 0) P.alive[c] = 0;
}

// ---------------------------------------------------------------------------
// k_stats: one workgroup. Deterministic FP64 reductions and the stop flag.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stats(StatsParams P) {
  if (*P.done) return;
  __shared__ double sd[256];
  __shared__ double sm[256];
  __shared__ int sc[256];
  double in = 0.0, mx = 0.0;
  int all = 1;
  for (int i = threadIdx.x; i < P.nparts; i += 256) in += P.ipart[i];
  for (int c = threadIdx.x; c < P.K; c += 256) {
    mx = fmax(mx, P.shift2[c]);
    all &= P.conv[c];
  }
  sd[threadIdx.x] = in;
  sm[threadIdx.x] = mx;
  sc[threadIdx.x] = all;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sd[threadIdx.x] += sd[threadIdx.x + o];
      sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + o]);
      sc[threadIdx.x] &= sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    DevState* s = P.state;
    const unsigned long long it = s->iter + 1;
    s->iter = it;
    s->inertia = sd[0];
    s->max_shift2 = sm[0];
    s->refined_two = P.qcnt[0];
    s->refined_all = P.qcnt[1];
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

// Sum of FP64 values (one workgroup, fixed order).
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
template <int DP, int PREC>
static hipError_t launch_assign_t(const AssignParams& P, int grid, size_t lds, hipStream_t s) {
  auto fn = k_assign<DP, PREC>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(256), lds, s, P);
  return hipGetLastError();
}

template <int PREC>
static hipError_t launch_assign_p(int DP, const AssignParams& P, int grid, size_t lds,
                                  hipStream_t s) {
  switch (DP) {
    case 16: return launch_assign_t<16, PREC>(P, grid, lds, s);
    case 32: return launch_assign_t<32, PREC>(P, grid, lds, s);
    case 64: return launch_assign_t<64, PREC>(P, grid, lds, s);
    case 128: return launch_assign_t<128, PREC>(P, grid, lds, s);
    case 256: return launch_assign_t<256, PREC>(P, grid, lds, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_assign(int DP, int prec, const AssignParams& P, int grid, size_t lds,
                         hipStream_t s) {
  switch (prec) {
    case PREC_FP32: return launch_assign_p<PREC_FP32>(DP, P, grid, lds, s);
    case PREC_BF16X3: return launch_assign_p<PREC_BF16X3>(DP, P, grid, lds, s);
    case PREC_BF16: return launch_assign_p<PREC_BF16>(DP, P, grid, lds, s);
    default: return hipErrorInvalidValue;
  }
}

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

const void* assign_kernel_ptr(int DP, int prec) {
#define KMC_PTR_CASE(dp)                                                                      \
  case dp:                                                                                    \
    return prec == PREC_FP32     ? reinterpret_cast<const void*>(k_assign<dp, PREC_FP32>)     \
           : prec == PREC_BF16X3 ? reinterpret_cast<const void*>(k_assign<dp, PREC_BF16X3>)   \
                                 : reinterpret_cast<const void*>(k_assign<dp, PREC_BF16>);
  switch (DP) {
    KMC_PTR_CASE(16)
    KMC_PTR_CASE(32)
    KMC_PTR_CASE(64)
    KMC_PTR_CASE(128)
    KMC_PTR_CASE(256)
    default: return nullptr;
  }
#undef KMC_PTR_CASE
}

hipError_t launch_prep(const PrepParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_prep, dim3((P.Kp + 3) / 4), dim3(256), 0, s, P);
  return hipGetLastError();
}
hipError_t launch_refine(const RefineParams& P, int grid2, int gridA, hipStream_t s) {
  hipLaunchKernelGGL(k_refine2, dim3(grid2), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_refine_all, dim3(gridA), dim3(256), 0, s, P);
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
hipError_t launch_update(const UpdateParams& P, hipStream_t s) {
  const int S = (P.D + 63) / 64;
  hipLaunchKernelGGL(k_update, dim3((P.K * S + 3) / 4), dim3(256), 0, s, P);
  return hipGetLastError();
}
hipError_t launch_finish(const FinishParams& F, const StatsParams& S, hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3((F.K + 255) / 256), dim3(256), 0, s, F);
  hipLaunchKernelGGL(k_stats, dim3(1), dim3(256), 0, s, S);
  return hipGetLastError();
}
hipError_t launch_load_points(const LoadParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_load_points, dim3((unsigned)((P.n + 3) / 4)), dim3(256), 0, s, P);
  return hipGetLastError();
}
hipError_t launch_sum_f64(const double* v, long long n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s, v, n, out);
  return hipGetLastError();
}
hipError_t launch_fill_i32(int* p, long long n, int v, hipStream_t s) {
  const long long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096)),
                     dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace kmc
