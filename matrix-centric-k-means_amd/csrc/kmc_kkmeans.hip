// kmc_kkmeans.hip -- gfx950 kernels of kernel k-means (Popcorn formulation).
//
// Set-up:
//   k_kk_norms   ||x_i||^2 (for the RBF kernel)
//   k_kk_build   K = f(X X^T) on MFMA (fp32 32x32x2, or bf16 32x32x16 as one
//                product or hi+lo split), lower-triangle tiles only, each
//                stored with its transpose; kernel function fused in the
//                epilogue, stored fp32 or bf16, row stride Np (zero pad)
// Per iteration (labels -> labels):
//   k_kk_onehot  H^T[c][j] = (label_j == c), fp32 or bf16 (exact 0/1)
//   k_kk_count   cluster sizes
//   k_kk_spmm    S = K H on MFMA (16x16x4 f32 or 16x16x32 bf16): every row of
//                K streamed from HBM once; per-block partial sums of
//                S[i, label_i] for the cluster self-terms
//   k_kk_gsum    g_c = sum_{i in c} S[i,c] (fixed order, FP64)
//   k_kk_assign  d(i,c) = K_ii - 2 S[i,c]/|c| + g_c/|c|^2, first minimum,
//                empty cluster (or g_c == 0) -> +inf; FP64 objective
//   k_kk_stats   iteration counter, score test, stop flag
// The cluster-sorted product (any number of clusters) replaces onehot,
// count and spmm:
//   launch_sort  stable counting sort of the points by label (the Lloyd
//                path's kernels): counts, offsets, sorted member list
//   k_kk_rowsum  K is symmetric, so S[:, c] = sum_{j in c} K[j, :]: rows of
//                K in sorted order, summed per (chunk of 256 sorted rows,
//                cluster) segment in fp32, one partial row per segment
//   k_kk_colsum  S[c, :] = the cluster's chunk partials in chunk order
//                (FP64), and the per-tile partials of g_c
//   k_kk_gfin    g_c = its tile partials in order
#include "kmc_device.h"
#include "kmc_kk.h"

namespace kmc {

constexpr int kKBTile = 128;  // kernel-matrix build: output tile edge

__global__ __launch_bounds__(256) void k_kk_norms(KKParams P) {
  const int lane = threadIdx.x & 63;
  const long long i = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= P.n) return;
  float s = 0.0f;
  for (int j = lane; j < P.d; j += 64) {
    const float v = P.X[(size_t)i * P.ldx + j];
    s = fmaf(v, v, s);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) P.xn2[i] = s;
}

__device__ __forceinline__ float kernel_fn(int kind, float v, float xi2, float xj2, float gamma,
                                           bool diag) {
  switch (kind) {
    case KK_LINEAR: return v;
    case KK_POLY: { const float t = v + 1.0f; return t * t; }
    case KK_SIGMOID: return tanhf(v + 1.0f);
    default: {
      const float sq = diag ? 0.0f : fmaxf(xi2 + xj2 - 2.0f * v, 0.0f);
      return expf(-gamma * sq);
    }
  }
}

// K = f(X X^T) is symmetric: only the tiles on and below the diagonal are
// computed (a 1-D grid over the T(T+1)/2 pairs ti >= tj of tiles),
// and an off-diagonal tile is stored twice, as itself and transposed.
// Workgroup of 4 waves over a 128 x 128 tile, each wave a 64 x 64 block (2 x 2
// MFMA blocks: every operand fragment read from LDS feeds two MFMAs); X
// staged through LDS in k-chunks of 32. The Gram product runs on MFMA in one of three forms:
//   KB_FP32    v_mfma_f32_32x32x2_f32 on the fp32 values
//   KB_BF16X3  x = hi + lo (bf16, round to nearest even), three
//              v_mfma_f32_32x32x16_bf16 per k-step: hi.hi + hi.lo + lo.hi
//   KB_BF16    one bf16 product
// all accumulated in fp32; the kernel function is applied in the epilogue.
__device__ __forceinline__ void tri_tile(long long b, long long& ti, long long& tj) {
  long long t = (long long)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while (t * (t + 1) / 2 > b) --t;
  while ((t + 1) * (t + 2) / 2 <= b) ++t;
  ti = t;
  tj = b - t * (t + 1) / 2;
}

// bf16 hi (and lo) planes of X, n x dpad, zero padded: x = hi + lo with hi
// = bf16(x), lo = bf16(x - hi) (round to nearest even)
__global__ __launch_bounds__(256) void k_kk_split(KKParams P) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= P.n * (long long)P.dpad) return;
  const long long i = t / P.dpad;
  const int k = (int)(t - i * P.dpad);
  const float x = k < P.d ? P.X[(size_t)i * P.ldx + k] : 0.0f;
  const uint16_t hi = bf16_rne(x);
  const_cast<uint16_t*>(P.Xh)[t] = hi;
  if (P.Xl) const_cast<uint16_t*>(P.Xl)[t] = bf16_rne(x - bf16_to_f32(hi));
}

template <int STORE, int BUILD>
__global__ __launch_bounds__(256) void k_kk_build(KKParams P) {
  constexpr int TT = kKBTile;  // output tile (points) per workgroup
  constexpr int TK = 32;
  constexpr int RS = TK + 4;   // fp32 row stride (floats)
  constexpr int RH = TK + 8;   // bf16 row stride (bf16 elements)
  constexpr bool kF32 = BUILD == KB_FP32;
  constexpr int kPl = BUILD == KB_BF16X3 ? 2 : 1;  // bf16 planes (hi, lo)
  __shared__ __attribute__((aligned(16))) float As[kF32 ? TT * RS : 1];
  __shared__ __attribute__((aligned(16))) float Bs[kF32 ? TT * RS : 1];
  __shared__ __attribute__((aligned(16))) uint16_t Ah[kF32 ? 1 : kPl * TT * RH];
  __shared__ __attribute__((aligned(16))) uint16_t Bh[kF32 ? 1 : kPl * TT * RH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long ti, tj;
  tri_tile((long long)blockIdx.x, ti, tj);
  const long long r0 = ti * TT, c0 = tj * TT;
  // wave: a 64 x 64 block of the tile, as 2 x 2 MFMA blocks of 32 x 32
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int p = lane & 31, h = lane >> 5;
  kmc_f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  // staging: each thread moves its share of each operand per k-chunk (fp32:
  // kSt floats; bf16: kSv 16-byte pieces per plane of the pre-split planes);
  // the next chunk is loaded into registers while the current one feeds the
  // MFMAs
  constexpr int kSt = kF32 ? TT * TK / 256 : 1;
  constexpr int kSv = kF32 ? 1 : TT * TK / 8 / 256;
  float ra[kSt], rb[kSt];
  uint4 va[kPl][kSv], vb[kPl][kSv];
  auto load_chunk = [&](int k0) {
    if constexpr (kF32) {
#pragma unroll
      for (int u = 0; u < kSt; ++u) {
        const int v = threadIdx.x + 256 * u;
        const int r = v / TK, q = v - r * TK;
        const long long gi = r0 + r, gj = c0 + r;
        const int kk = k0 + q;
        ra[u] = (gi < P.n && kk < P.d) ? P.X[(size_t)gi * P.ldx + kk] : 0.0f;
        rb[u] = (gj < P.n && kk < P.d) ? P.X[(size_t)gj * P.ldx + kk] : 0.0f;
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSv; ++u) {
        const int v = threadIdx.x + 256 * u;
        const int r = v / (TK / 8), q = 8 * (v - r * (TK / 8));
        const long long gi = r0 + r, gj = c0 + r;
#pragma unroll
        for (int pl = 0; pl < kPl; ++pl) {
          const uint16_t* src = pl == 0 ? P.Xh : P.Xl;
          va[pl][u] = gi < P.n ? *reinterpret_cast<const uint4*>(src + (size_t)gi * P.dpad + k0 + q)
                               : make_uint4(0u, 0u, 0u, 0u);
          vb[pl][u] = gj < P.n ? *reinterpret_cast<const uint4*>(src + (size_t)gj * P.dpad + k0 + q)
                               : make_uint4(0u, 0u, 0u, 0u);
        }
      }
    }
  };
  load_chunk(0);
  for (int k0 = 0; k0 < P.d; k0 += TK) {
    __syncthreads();
    if constexpr (kF32) {
#pragma unroll
      for (int u = 0; u < kSt; ++u) {
        const int v = threadIdx.x + 256 * u;
        const int r = v / TK, q = v - r * TK;
        As[r * RS + q] = ra[u];
        Bs[r * RS + q] = rb[u];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSv; ++u) {
        const int v = threadIdx.x + 256 * u;
        const int r = v / (TK / 8), q = 8 * (v - r * (TK / 8));
#pragma unroll
        for (int pl = 0; pl < kPl; ++pl) {
          *reinterpret_cast<uint4*>(&Ah[pl * TT * RH + r * RH + q]) = va[pl][u];
          *reinterpret_cast<uint4*>(&Bh[pl * TT * RH + r * RH + q]) = vb[pl][u];
        }
      }
    }
    __syncthreads();
    if (k0 + TK < P.d) load_chunk(k0 + TK);
    if constexpr (kF32) {
      // lane half h takes k = 2*step + h
#pragma unroll 4
      for (int s = 0; s < TK / 2; ++s) {
        float a[2], b[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          a[u] = As[(wr + 32 * u + p) * RS + 2 * s + h];
          b[u] = Bs[(wc + 32 * u + p) * RS + 2 * s + h];
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
            acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[x], b[y], acc[x][y], 0, 0, 0);
      }
    } else {
      // k-step s: dimensions 16 s + 8 h + [0, 8) of every row
#pragma unroll
      for (int s = 0; s < TK / 16; ++s) {
        uint4 a[2][kPl], b[2][kPl];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int pl = 0; pl < kPl; ++pl) {
            a[u][pl] = *reinterpret_cast<const uint4*>(
                &Ah[pl * TT * RH + (wr + 32 * u + p) * RH + 16 * s + 8 * h]);
            b[u][pl] = *reinterpret_cast<const uint4*>(
                &Bh[pl * TT * RH + (wc + 32 * u + p) * RH + 16 * s + 8 * h]);
          }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(kmc_bf16x8, a[x][0]), __builtin_bit_cast(kmc_bf16x8, b[y][0]),
                acc[x][y], 0, 0, 0);
            if constexpr (kPl == 2) {
              acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(kmc_bf16x8, a[x][0]),
                  __builtin_bit_cast(kmc_bf16x8, b[y][kPl - 1]), acc[x][y], 0, 0, 0);
              acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(kmc_bf16x8, a[x][kPl - 1]),
                  __builtin_bit_cast(kmc_bf16x8, b[y][0]), acc[x][y], 0, 0, 0);
            }
          }
      }
    }
  }
  // acc[x][y][r]: row (i) = 32 x + 8*(r>>2) + 4*h + (r&3), column (j) = 32 y + p
  const bool mirror = ti != tj;
#pragma unroll
  for (int y = 0; y < 2; ++y) {
    const long long gj = c0 + wc + 32 * y + p;
    const float xj2 = (P.kind == KK_RBF && gj < P.n) ? P.xn2[gj] : 0.0f;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        float v[4];
        const long long gi0 = r0 + wr + 32 * x + 8 * g + 4 * h;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const long long gi = gi0 + c;
          v[c] = 0.0f;
          if (gi < P.n && gj < P.n) {
            const float xi2 = P.kind == KK_RBF ? P.xn2[gi] : 0.0f;
            v[c] = kernel_fn(P.kind, acc[x][y][4 * g + c], xi2, xj2, P.gamma, gi == gj);
          }
          // a diagonal tile keeps its lower triangle and mirrors it (the
          // hi+lo product is not symmetric in the order of its terms)
          if (gi < P.Np && gj < P.Np && (mirror || gi >= gj)) {
            const size_t off = (size_t)gi * P.Np + gj;
            if (STORE == KS_FP32) reinterpret_cast<float*>(P.Kmat)[off] = v[c];
            else reinterpret_cast<uint16_t*>(P.Kmat)[off] = bf16_rne(v[c]);
            if (!mirror && gi > gj) {
              const size_t moff = (size_t)gj * P.Np + gi;
              if (STORE == KS_FP32) reinterpret_cast<float*>(P.Kmat)[moff] = v[c];
              else reinterpret_cast<uint16_t*>(P.Kmat)[moff] = bf16_rne(v[c]);
            }
          }
        }
        // the transposed copy: row gj, columns gi0 .. gi0 + 3 (one 16- or
        // 8-byte store per lane; a row's pieces meet in the L2 as whole lines)
        if (mirror && gj < P.Np && gi0 + 3 < P.Np) {
          const size_t off = (size_t)gj * P.Np + gi0;
          if (STORE == KS_FP32) {
            *reinterpret_cast<float4*>(reinterpret_cast<float*>(P.Kmat) + off) =
                make_float4(v[0], v[1], v[2], v[3]);
          } else {
            const uint32_t lo = (uint32_t)bf16_rne(v[0]) | ((uint32_t)bf16_rne(v[1]) << 16);
            const uint32_t hi = (uint32_t)bf16_rne(v[2]) | ((uint32_t)bf16_rne(v[3]) << 16);
            *reinterpret_cast<uint2*>(reinterpret_cast<uint16_t*>(P.Kmat) + off) =
                make_uint2(lo, hi);
          }
        }
      }
  }
}

// Diagonal of K (as stored), FP64 sum of it goes to the state.
template <int STORE>
__global__ __launch_bounds__(256) void k_kk_diag(KKParams P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const size_t off = (size_t)i * P.Np + i;
  P.kdiag[i] = STORE == KS_FP32 ? reinterpret_cast<const float*>(P.Kmat)[off]
                                : bf16_to_f32(reinterpret_cast<const uint16_t*>(P.Kmat)[off]);
}

template <int STORE>
__global__ __launch_bounds__(256) void k_kk_onehot(KKParams P) {
  if (*P.done) return;
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.Np) return;
  const int lab = j < P.n ? P.labels[j] : -1;
  for (int c = 0; c < P.Kp; ++c) {
    const bool on = (lab == c);
    const size_t off = (size_t)c * P.Np + j;
    if (STORE == KS_FP32) reinterpret_cast<float*>(P.HT)[off] = on ? 1.0f : 0.0f;
    else reinterpret_cast<uint16_t*>(P.HT)[off] = on ? (uint16_t)0x3f80 : (uint16_t)0;
  }
}

__global__ __launch_bounds__(256) void k_kk_count(KKParams P) {
  if (*P.done) return;
  __shared__ int cnt_s[KK_MAX_K];
  for (int c = threadIdx.x; c < P.K; c += 256) cnt_s[c] = 0;
  __syncthreads();
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < P.n;
       i += (long long)gridDim.x * 256)
    atomicAdd(&cnt_s[P.labels[i]], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < P.K; c += 256)
    if (cnt_s[c]) atomicAdd(&P.count[c], cnt_s[c]);
}

// One wave per 16-row block of K. MFMA 16x16: A = 16 rows x k columns of K,
// B = k columns x 16 clusters of H; D[row][cluster]. Lane l: row/cluster
// (l & 15), k-group g = l >> 4.
//   fp32 (16x16x4): a 16-byte load K[row][j0 + 16q + 4g + 0..3] feeds steps
//                   4q..4q+3 (column j0 + 16q + 4g + t at step t)
//   bf16 (16x16x32): K[row][j0 + 8g + 0..7] per step, 32 columns per step
// acc[r] of lane l holds S[row 4g + r][cluster l & 15].
template <int STORE, int NB>
__global__ __launch_bounds__(256) void k_kk_spmm(KKParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  const long long blk = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long nblk = (P.n + 15) / 16;
  if (blk >= nblk) return;
  const int m = lane & 15, g = lane >> 4;
  const long long row = blk * 16 + m;  // < Np (K rows are padded to Np)
  kmc_f32x4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = kmc_f32x4{0.f, 0.f, 0.f, 0.f};
  if (STORE == KS_FP32) {
    const float* krow = reinterpret_cast<const float*>(P.Kmat) + (size_t)row * P.Np + 4 * g;
    const float* H = reinterpret_cast<const float*>(P.HT) + 4 * g;
    for (long long j0 = 0; j0 < P.Np; j0 += 16) {
      const float4 a = *reinterpret_cast<const float4*>(krow + j0);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const float4 hb =
            *reinterpret_cast<const float4*>(H + (size_t)(16 * b + m) * P.Np + j0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, hb.x, acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, hb.y, acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, hb.z, acc[b], 0, 0, 0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, hb.w, acc[b], 0, 0, 0);
      }
    }
  } else {
    const uint16_t* krow = reinterpret_cast<const uint16_t*>(P.Kmat) + (size_t)row * P.Np + 8 * g;
    const uint16_t* H = reinterpret_cast<const uint16_t*>(P.HT) + 8 * g;
    for (long long j0 = 0; j0 < P.Np; j0 += 32) {
      const uint4 a = *reinterpret_cast<const uint4*>(krow + j0);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const uint4 hb = *reinterpret_cast<const uint4*>(H + (size_t)(16 * b + m) * P.Np + j0);
        acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(kmc_bf16x8, a),
                                                        __builtin_bit_cast(kmc_bf16x8, hb),
                                                        acc[b], 0, 0, 0);
      }
    }
  }
  // write S and the partial self-terms of this block
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int c = 16 * b + m;
    float part = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long long i = blk * 16 + 4 * g + r;
      if (i < P.n) {
        P.S[(size_t)i * P.Kp + c] = acc[b][r];
        if (P.labels[i] == c) part += acc[b][r];
      }
    }
    // fixed-order combine of the 4 row groups
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    if (g == 0) P.gpart[(size_t)blk * P.Kp + c] = part;
  }
}

// ---------------------------------------------------------------------------
// Cluster-sorted product (the reference's REORDER level, src/kmeans.cu:
// 1125-1261, puts K's columns in cluster order so that V K has contiguous
// blocks; here K stays in place): K is symmetric, so the sum over a
// cluster's columns of row i equals the sum over the cluster's rows at
// column i. Workgroup (column tile x, chunk q) stages the member ids and
// labels of sorted positions [256 q, 256 q + 256) in LDS, then walks them
// segment by segment (a segment = one cluster's part of the chunk), adding
// the rows' 16-byte column slices in sorted order in fp32 (eight rows' loads
// issued together); each segment leaves one partial row at index c + q,
// injective because clusters occupy increasing chunk ranges. Every element
// of K is read once, in whole-row runs, whatever the number of clusters.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kk_seg_end(const int* lab_s, int r, int L, int c, int lane) {
  for (int w = r + 1; w < L; w += 64) {
    const int i = w + lane;
    const uint64_t m = __ballot(i >= L || lab_s[i < L ? i : 0] != c);
    if (m) return w + __ffsll((unsigned long long)m) - 1;
  }
  return L;
}

template <int STORE>
__global__ __launch_bounds__(256) void k_kk_rowsum(KKParams P) {
  if (*P.done) return;
  __shared__ int mem_s[kKKRows], lab_s[kKKRows];
  constexpr int ES = STORE == KS_FP32 ? 4 : 2;
  constexpr int CPT = 16 / ES;  // columns per thread: one 16-byte load
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.y;
  const long long p0 = (long long)q * kKKRows;
  const int L = (int)min((long long)kKKRows, P.n - p0);
  for (int t = threadIdx.x; t < kKKRows; t += 256) {
    const int id = t < L ? P.members[p0 + t] : 0;
    mem_s[t] = id;
    lab_s[t] = t < L ? P.labels[id] : -1;
  }
  __syncthreads();
  const long long col = ((long long)blockIdx.x * 256 + threadIdx.x) * CPT;
  const bool act = col < P.Np;  // Np is a multiple of 64: whole loads in range
  const unsigned char* Kb =
      reinterpret_cast<const unsigned char*>(P.Kmat) + (size_t)(act ? col : 0) * ES;
  const size_t rowb = (size_t)P.Np * ES;
  int r = 0;
  while (r < L) {
    const int c = lab_s[r];
    const int e = kk_seg_end(lab_s, r, L, c, lane);
    float sum[CPT];
#pragma unroll
    for (int v = 0; v < CPT; ++v) sum[v] = 0.0f;
    constexpr int U = 8;
    for (int t = r; t < e; t += U) {
      uint4 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int tt = min(t + u, e - 1);
        x[u] = *reinterpret_cast<const uint4*>(Kb + (size_t)mem_s[tt] * rowb);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (t + u < e) {  // segment-uniform
          const uint32_t w[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
          if (STORE == KS_FP32) {
#pragma unroll
            for (int v = 0; v < 4; ++v) sum[v % CPT] += __uint_as_float(w[v]);
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              sum[(2 * v) % CPT] += __uint_as_float(w[v] << 16);
              sum[(2 * v + 1) % CPT] += __uint_as_float(w[v] & 0xffff0000u);
            }
          }
        }
      }
    }
    if (act) {
      float* dst = P.tpart + (size_t)(c + q) * P.Np + col;
#pragma unroll
      for (int v = 0; v < CPT; v += 4)
        *reinterpret_cast<float4*>(dst + v) = make_float4(sum[v], sum[v + 1 < CPT ? v + 1 : v],
                                                          sum[v + 2 < CPT ? v + 2 : v],
                                                          sum[v + 3 < CPT ? v + 3 : v]);
    }
    r = e;
  }
}

// S[c, col] = sum of cluster c's chunk partials in chunk order (FP64, rounded
// once); the columns of c's own points add S[c, i] to the tile's partial of
// g_c (fixed-order block reduction). Workgroup (column tile, cluster), four
// columns per thread.
__global__ __launch_bounds__(256) void k_kk_colsum(KKParams P) {
  if (*P.done) return;
  __shared__ double red_s[256];
  const int c = blockIdx.y;
  const long long col = ((long long)blockIdx.x * 256 + threadIdx.x) * 4;
  const int cnt = P.count[c];
  double gs = 0.0;
  if (col < P.Np) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (cnt > 0) {
      const long long q0 = (long long)P.start[c] / kKKRows;
      const long long q1 = ((long long)P.start[c] + cnt - 1) / kKKRows;
      for (long long q = q0; q <= q1; ++q) {
        const float4 v = *reinterpret_cast<const float4*>(P.tpart + (size_t)(c + q) * P.Np + col);
        s[0] += v.x;
        s[1] += v.y;
        s[2] += v.z;
        s[3] += v.w;
      }
    }
    float out[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = (float)s[v];
    *reinterpret_cast<float4*>(P.S + (size_t)c * P.Np + col) =
        make_float4(out[0], out[1], out[2], out[3]);
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (col + v < P.n && P.labels[col + v] == c) gs += (double)out[v];
  }
  red_s[threadIdx.x] = gs;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red_s[threadIdx.x] += red_s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) P.gtile[(size_t)c * P.ntile + blockIdx.x] = red_s[0];
}

__global__ __launch_bounds__(256) void k_kk_gfin(KKParams P) {
  if (*P.done) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= P.K) return;
  double g = 0.0;
  for (int t = 0; t < P.ntile; ++t) g += P.gtile[(size_t)c * P.ntile + t];
  P.g[c] = g;
}

// g_c = sum over row blocks (in order) of the partial self-terms, FP64.
__global__ __launch_bounds__(256) void k_kk_gsum(KKParams P) {
  if (*P.done) return;
  const int c = blockIdx.x;
  __shared__ double sd[256];
  const long long nblk = (P.n + 15) / 16;
  double s = 0.0;
  for (long long b = threadIdx.x; b < nblk; b += 256) s += (double)P.gpart[(size_t)b * P.Kp + c];
  sd[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sd[threadIdx.x] += sd[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) P.g[c] = sd[0];
}

// One thread per point.
__global__ __launch_bounds__(256) void k_kk_assign(KKParams P) {
  if (*P.done) return;
  __shared__ float cnorm_s[KK_MAX_K], inv_s[KK_MAX_K];
  __shared__ double red_s[256];
  __shared__ int chg_s[256];
  for (int c = threadIdx.x; c < P.K; c += 256) {
    const int n_c = P.count[c];
    const double gc = P.g[c];
    // empty cluster, or a zero centroid norm, is never chosen
    // (src/kernels/distances.cu:588-598)
    const bool dead = (n_c == 0) || (gc == 0.0);
    cnorm_s[c] = dead ? __builtin_inff() : (float)(gc / ((double)n_c * (double)n_c));
    inv_s[c] = n_c ? (float)(2.0 / (double)n_c) : 0.0f;
  }
  __syncthreads();
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  double obj = 0.0;
  int changed = 0;
  if (i < P.n) {
    const float* s = P.S + (size_t)i * P.sS_i;
    float best = __builtin_inff();
    int bi = 0;
    for (int c = 0; c < P.K; ++c) {
      const float v = cnorm_s[c] - inv_s[c] * s[(size_t)c * P.sS_c];
      if (v < best) { best = v; bi = c; }
    }
    obj = (double)P.kdiag[i] + (double)best;
    changed = (bi != P.labels[i]);
    P.labels[i] = bi;
  }
  red_s[threadIdx.x] = obj;
  chg_s[threadIdx.x] = changed;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red_s[threadIdx.x] += red_s[threadIdx.x + o];
      chg_s[threadIdx.x] += chg_s[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    P.opart[blockIdx.x] = red_s[0];
    P.cpart[blockIdx.x] = chg_s[0];
  }
}

__global__ __launch_bounds__(256) void k_kk_stats(KKParams P) {
  if (*P.done) return;
  __shared__ double sd[256];
  __shared__ long long sc[256];
  double s = 0.0;
  long long ch = 0;
  for (int b = threadIdx.x; b < P.nassign; b += 256) {
    s += P.opart[b];
    ch += P.cpart[b];
  }
  sd[threadIdx.x] = s;
  sc[threadIdx.x] = ch;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sd[threadIdx.x] += sd[threadIdx.x + o];
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    KKState* st = P.state;
    const unsigned long long it = st->iter + 1;
    st->iter = it;
    st->objective = sd[0];
    st->changed = (unsigned long long)sc[0];
    const double score = sd[0] - st->sum_diag;
    if (P.mode == CONV_SCORE && it >= 2 && it < st->maxiter &&
        fabsf((float)score - (float)st->last_score) < P.tol) {
      st->converged_at = it;
      st->done = 1;
    }
    if (P.mode == CONV_STABLE && sc[0] == 0) {
      st->converged_at = it;
      st->done = 1;
    }
    st->last_score = score;
    if (it >= st->maxiter) st->done = 1;
  }
}

// Distances as the reference materialises them (src/kernels/distances.cu:
// 895-1002, without the K_ii term): D[i,c] = g_c/|c|^2 - 2 S[i,c]/|c|,
// +inf for an empty cluster or a zero norm.
__global__ __launch_bounds__(256) void k_kk_dist_out(KKParams P) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n) return;
  for (int c = 0; c < P.K; ++c) {
    const int n_c = P.count[c];
    const double gc = P.g[c];
    float v = __builtin_inff();
    if (n_c != 0 && gc != 0.0)
      v = (float)(gc / ((double)n_c * (double)n_c)) -
          (float)(2.0 / (double)n_c) * P.S[(size_t)i * P.sS_i + (size_t)c * P.sS_c];
    P.Dout[(size_t)i * P.K + c] = v;
  }
}

__global__ void k_kk_zero(KKParams P) {
  if (*P.done) return;
  for (int c = threadIdx.x; c < P.K; c += blockDim.x) P.count[c] = 0;
}

// ---------------------------------------------------------------------------
hipError_t kk_launch_setup(const KKParams& P, hipStream_t s) {
  hipLaunchKernelGGL(k_kk_norms, dim3((unsigned)((P.n + 3) / 4)), dim3(256), 0, s, P);
  if (P.build != KB_FP32)
    hipLaunchKernelGGL(k_kk_split, dim3((unsigned)((P.n * (long long)P.dpad + 255) / 256)),
                       dim3(256), 0, s, P);
  const long long T = (P.Np + kKBTile - 1) / kKBTile;
  const dim3 grid((unsigned)(T * (T + 1) / 2));  // tiles on and below the diagonal
#define KK_BUILD_CASE(st, bd) \
  if (P.store == st && P.build == bd) hipLaunchKernelGGL((k_kk_build<st, bd>), grid, dim3(256), 0, s, P);
  KK_BUILD_CASE(KS_FP32, KB_FP32)
  KK_BUILD_CASE(KS_FP32, KB_BF16X3)
  KK_BUILD_CASE(KS_FP32, KB_BF16)
  KK_BUILD_CASE(KS_BF16, KB_FP32)
  KK_BUILD_CASE(KS_BF16, KB_BF16X3)
  KK_BUILD_CASE(KS_BF16, KB_BF16)
#undef KK_BUILD_CASE
  if (P.store == KS_FP32)
    hipLaunchKernelGGL(k_kk_diag<KS_FP32>, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, s, P);
  else
    hipLaunchKernelGGL(k_kk_diag<KS_BF16>, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, s, P);
  return hipGetLastError();
}

template <int STORE>
static void launch_spmm_t(const KKParams& P, unsigned grid, hipStream_t s) {
  switch (P.Kp / 16) {
    case 1: hipLaunchKernelGGL((k_kk_spmm<STORE, 1>), dim3(grid), dim3(256), 0, s, P); break;
    case 2: hipLaunchKernelGGL((k_kk_spmm<STORE, 2>), dim3(grid), dim3(256), 0, s, P); break;
    default: hipLaunchKernelGGL((k_kk_spmm<STORE, 4>), dim3(grid), dim3(256), 0, s, P); break;
  }
}

// the cluster-sorted product: sort, chunk row sums, column sums, g
static hipError_t launch_sorted_product(const KKParams& P, hipStream_t s) {
  SortParams sp{};
  sp.labels = P.labels;
  sp.bcount = P.bcount;
  sp.count = P.count;
  sp.start = P.start;
  sp.members = P.members;
  sp.done = P.done;
  sp.n = P.n;
  sp.K = P.K;
  sp.CH = P.CH;
  sp.nblk = P.nblk;
  sp.lbits = P.lbits;
  hipError_t e = launch_sort(sp, s);
  if (e != hipSuccess) return e;
  return hipSuccess;
}
static hipError_t launch_sorted_sums(const KKParams& P, hipStream_t s) {
  const int cpt = P.store == KS_FP32 ? 4 : 8;
  const dim3 gr((unsigned)((P.Np + 256LL * cpt - 1) / (256LL * cpt)), (unsigned)P.nq);
  if (P.store == KS_FP32) hipLaunchKernelGGL(k_kk_rowsum<KS_FP32>, gr, dim3(256), 0, s, P);
  else hipLaunchKernelGGL(k_kk_rowsum<KS_BF16>, gr, dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_colsum, dim3((unsigned)P.ntile, (unsigned)P.K), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_gfin, dim3((unsigned)((P.K + 255) / 256)), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t kk_launch_distances(const KKParams& P, hipStream_t s) {
  if (P.product == KK_SORTED) {
    hipError_t e = launch_sorted_product(P, s);
    if (e == hipSuccess) e = launch_sorted_sums(P, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_kk_dist_out, dim3(P.nassign), dim3(256), 0, s, P);
    return hipGetLastError();
  }
  const unsigned gj = (unsigned)((P.Np + 255) / 256);
  if (P.store == KS_FP32) hipLaunchKernelGGL(k_kk_onehot<KS_FP32>, dim3(gj), dim3(256), 0, s, P);
  else hipLaunchKernelGGL(k_kk_onehot<KS_BF16>, dim3(gj), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_zero, dim3(1), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_count, dim3(64), dim3(256), 0, s, P);
  const unsigned gs = (unsigned)(((P.n + 15) / 16 + 3) / 4);
  if (P.store == KS_FP32) launch_spmm_t<KS_FP32>(P, gs, s);
  else launch_spmm_t<KS_BF16>(P, gs, s);
  hipLaunchKernelGGL(k_kk_gsum, dim3(P.K), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_dist_out, dim3(P.nassign), dim3(256), 0, s, P);
  return hipGetLastError();
}

hipError_t kk_launch_iter(const KKParams& P, hipEvent_t e0, hipEvent_t e1, hipStream_t s) {
  if (P.product == KK_SORTED) {
    hipError_t e = launch_sorted_product(P, s);
    if (e != hipSuccess) return e;
    if (e0) (void)hipEventRecord(e0, s);
    e = launch_sorted_sums(P, s);
    if (e != hipSuccess) return e;
    if (e1) (void)hipEventRecord(e1, s);
    hipLaunchKernelGGL(k_kk_assign, dim3(P.nassign), dim3(256), 0, s, P);
    hipLaunchKernelGGL(k_kk_stats, dim3(1), dim3(256), 0, s, P);
    return hipGetLastError();
  }
  const unsigned gj = (unsigned)((P.Np + 255) / 256);
  if (P.store == KS_FP32) hipLaunchKernelGGL(k_kk_onehot<KS_FP32>, dim3(gj), dim3(256), 0, s, P);
  else hipLaunchKernelGGL(k_kk_onehot<KS_BF16>, dim3(gj), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_zero, dim3(1), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_count, dim3(64), dim3(256), 0, s, P);
  const unsigned gs = (unsigned)(((P.n + 15) / 16 + 3) / 4);
  if (e0) (void)hipEventRecord(e0, s);
  if (P.store == KS_FP32) launch_spmm_t<KS_FP32>(P, gs, s);
  else launch_spmm_t<KS_BF16>(P, gs, s);
  if (e1) (void)hipEventRecord(e1, s);
  hipLaunchKernelGGL(k_kk_gsum, dim3(P.K), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_assign, dim3(P.nassign), dim3(256), 0, s, P);
  hipLaunchKernelGGL(k_kk_stats, dim3(1), dim3(256), 0, s, P);
  return hipGetLastError();
}

}  // namespace kmc
