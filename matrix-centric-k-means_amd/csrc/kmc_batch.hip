// kmc_batch.hip -- many independent Lloyd problems in one device layout
// (kmc_batch_* of include/kmc.h).
//
// Problems are grouped by padded width DP (d rounded up to 16, 32, 64, 128 or
// 256); each group is packed and launched on its own stream, so the groups'
// iterations run side by side.
//
// Layout of a group: its problems' points are packed into one point space, problem p at
// rows [poff_p, poff_p + n_p) with poff_p a multiple of kBatchAlign; the rows
// up to the next problem are zero and carry the problem's extra "dead"
// cluster (local index k_p), so every stage can run over the whole packed
// space and the stable sort places problem p's members at the same offsets
// (relative to poff_p) as a run of that problem alone. Master clusters of p
// are [coff_p, coff_p + k_p]; the distance operands of its k_p real clusters
// are rows [cpoff_p, cpoff_p + kpad_p) (padding rows carry kFar).
//
// One iteration, one launch per stage for the whole group:
//   k_bprep      distance operands + transposed centroids of every problem
//   k_assign<..., BATCH>  persistent distance kernel over the packed
//                (problem, tile) queue; converged problems are skipped
//   k_refine2 / k_refine_all   exact refinement (the all-k pass restricted
//                to the point's problem)
//   k_bhist / k_bscan_blocks / k_bscatter   stable counting sort, blocks of
//                kBatchAlign points never straddle problems
//   k_update_tree, k_finish    exact update and shift (shared kernels): the
//                64-bit integer sums of round(x * scale) at each problem's
//                fixed-point scale over the sorted members, or, when at most
//                an eighth of the points moved, the moved points' rows
//                subtracted from and added to the kept cluster totals (the
//                sort kernels then return at once)
//   k_bstats     per-problem shift test and iteration counter, stop flags
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kmc.h"
#include "kmc_copy.h"
#include "kmc_device.h"
#include "kmc_kernels.h"

namespace kmc {

struct BPrepParams {
  const float* C;  // master centroids, KT x DP
  float* Cw;
  uint16_t* Chi;
  uint16_t* Clo;
  float* cn;
  float* CT;  // DP x KTp
  const BatchProblem* bp;
  const int* row_prob;  // problem of each operand row
  const int* done;
  int nrows, KTp, DP, prec;
};

// One wave per distance-operand row (the same arithmetic as prep_row, so a
// problem's operands are the bits a run of it alone would use).
__global__ __launch_bounds__(256) void k_bprep(BPrepParams P) {
  if (*P.done) return;
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= P.nrows) return;
  const BatchProblem bpr = P.bp[P.row_prob[r]];
  const int c = r - bpr.cpoff;
  const bool real = c < bpr.k;
  const int g = bpr.coff + c;
  const size_t row = (size_t)r * P.DP;
  float s = 0.0f;
  for (int j = lane; j < P.DP; j += 64) {
    const float v = real ? P.C[(size_t)g * P.DP + j] : 0.0f;
    const float w = -2.0f * v;
    s = fmaf(v, v, s);
    if (real) P.CT[(size_t)j * P.KTp + g] = v;
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
  if (lane == 0) P.cn[r] = real ? s : kFar;
}

struct BSortParams {
  const int* labels;
  int* bcount;
  int* count;
  int* start;
  int* members;
  const BatchProblem* bp;
  const int* blk_prob;   // problem of each kBatchAlign-point block
  const int* clus_prob;  // problem of each master cluster
  const int* done;
  int nblk, KT, lbits;
  // the sort serves the full passes of the exact update only
  const DevState* fx_state;
  long long fx_n;
};

// Block histograms: one wave per kBatchAlign points, local labels, the counts
// stored cluster-major per problem.
__global__ __launch_bounds__(64) void k_bhist(BSortParams P) {
  if (*P.done) return;
  if (!fx_full_pass(P.fx_state, P.fx_n)) return;
  extern __shared__ int hist_s[];
  const int b = blockIdx.x;
  const BatchProblem bpr = P.bp[P.blk_prob[b]];
  const int lb = b - (int)(bpr.poff / kBatchAlign);
  const int kc = bpr.k + 1;
  for (int c = threadIdx.x; c < kc; c += 64) hist_s[c] = 0;
  __syncthreads();
  const long long lo = (long long)b * kBatchAlign;
  int lab[kBatchAlign / 64];
#pragma unroll
  for (int u = 0; u < kBatchAlign / 64; ++u) lab[u] = P.labels[lo + u * 64 + threadIdx.x];
#pragma unroll
  for (int u = 0; u < kBatchAlign / 64; ++u) atomicAdd(&hist_s[lab[u] - bpr.coff], 1);
  __syncthreads();
  for (int c = threadIdx.x; c < kc; c += 64)
    P.bcount[bpr.bcoff + (size_t)c * bpr.nblk + lb] = hist_s[c];
}

// Exclusive scan over the problem's blocks for each master cluster (one
// wave per cluster, 8 counts per lane per round).
__global__ __launch_bounds__(256) void k_bscan_blocks(BSortParams P) {
  if (*P.done) return;
  if (!fx_full_pass(P.fx_state, P.fx_n)) return;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= P.KT) return;
  const BatchProblem bpr = P.bp[P.clus_prob[g]];
  const int c = g - bpr.coff;
  int* a = P.bcount + bpr.bcoff + (size_t)c * bpr.nblk;
  const int nb = bpr.nblk;
  int carry = 0;
  for (int base = 0; base < nb; base += 512) {
    int v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = base + lane * 8 + u;
      v[u] = idx < nb ? a[idx] : 0;
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
      if (idx < nb) a[idx] = off + v[u];
    }
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) P.count[g] = carry;
}

// Stable scatter, one wave per block: the problem's cluster offsets (its
// members start at poff), then 64 points at a time with equal labels found
// by one ballot per label bit.
__global__ __launch_bounds__(64) void k_bscatter(BSortParams P) {
  if (*P.done) return;
  if (!fx_full_pass(P.fx_state, P.fx_n)) return;
  extern __shared__ int cur_s[];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const BatchProblem bpr = P.bp[P.blk_prob[b]];
  const int lb = b - (int)(bpr.poff / kBatchAlign);
  const int kc = bpr.k + 1;
  {
    const int S = (kc + 63) / 64;
    const int c0 = min(kc, lane * S), c1 = min(kc, c0 + S);
    int tot = 0;
    for (int c = c0; c < c1; ++c) tot += P.count[bpr.coff + c];
    int incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    int run = (int)bpr.poff + incl - tot;
    for (int c = c0; c < c1; ++c) {
      if (lb == 0) P.start[bpr.coff + c] = run;
      cur_s[c] = run + P.bcount[bpr.bcoff + (size_t)c * bpr.nblk + lb];
      run += P.count[bpr.coff + c];
    }
  }
  __syncthreads();
  const long long lo = (long long)b * kBatchAlign;
  const uint64_t lt = lanemask_lt();
  int labs[kBatchAlign / 64];
#pragma unroll
  for (int u = 0; u < kBatchAlign / 64; ++u) labs[u] = P.labels[lo + u * 64 + lane] - bpr.coff;
#pragma unroll
  for (int u = 0; u < kBatchAlign / 64; ++u) {
    const long long i = lo + u * 64 + lane;
    const int lab = labs[u];
    uint64_t match = ~0ull;
    for (int bit = 0; bit < P.lbits; ++bit) {
      const bool on = (lab >> bit) & 1;
      const uint64_t B = __ballot(on);
      match &= on ? B : ~B;
    }
    const int rank = __popcll(match & lt);
    const int pos = cur_s[lab] + rank;
    __builtin_amdgcn_wave_barrier();
    if (rank == 0) cur_s[lab] = pos + __popcll(match);
    P.members[pos] = (int)i;
    __builtin_amdgcn_wave_barrier();
  }
}

struct BStatsParams {
  const int* conv;
  const BatchProblem* bp;
  int* pdone;
  unsigned* piter;
  unsigned* pconv;
  DevState* state;
  const int* done;
  int nprob, check;
  unsigned long long maxiter;
  long long n;  // points of the packed space (the full-pass threshold)
};

// Per-problem iteration counter and shift test (reference: tested from
// iteration 2, comparisons/CPU_Kmeans/src/kmeans.cpp:205-208); a problem that
// passes or reaches maxiter is frozen; the batch stops when all are.
constexpr int kBStatsThreads = 1024;
__global__ __launch_bounds__(kBStatsThreads) void k_bstats(BStatsParams P) {
  if (*P.done) return;
  __shared__ int active_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    active_s = 0;
    // the iteration's update is complete: its moved count decides the next
    // one, and the cluster totals now stand for its labels
    if (fx_full_pass(P.state, P.n))
      P.state->full_passes += 1;
    else
      P.state->moved_total += P.state->moved_acc;
    P.state->moved = P.state->moved_acc;
    P.state->moved_acc = 0ull;
    P.state->sums_valid = 1;
  }
  __syncthreads();
  int active = 0;
  for (int pb = wave; pb < P.nprob; pb += kBStatsThreads / 64) {
    if (P.pdone[pb]) continue;
    const BatchProblem bpr = P.bp[pb];
    int all = 1;
    for (int c = lane; c < bpr.k; c += 64) all &= P.conv[bpr.coff + c];
    all = __ballot(!all) == 0ull;
    const unsigned it = P.piter[pb] + 1;
    bool stop = false;
    if (P.check && it >= 2 && all) {
      if (lane == 0) P.pconv[pb] = it;
      stop = true;
    }
    if (it >= P.maxiter && !stop) {  // not converged: maxiter, as Kmeans::run returns
      if (lane == 0) P.pconv[pb] = it;
      stop = true;
    }
    if (lane == 0) {
      P.piter[pb] = it;
      if (stop) P.pdone[pb] = 1;
    }
    if (!stop) ++active;
  }
  if (lane == 0 && active) atomicAdd(&active_s, active);
  __syncthreads();
  if (threadIdx.x == 0) {
    P.state->iter += 1;
    if (active_s == 0) P.state->done = 1;
  }
}

// Every point of the packed space starts in its problem's dead cluster (the
// padding rows stay there; the first assignment moves the real points).
__global__ __launch_bounds__(256) void k_binit_labels(int* labels, const BatchProblem* bp,
                                                      const int* blk_prob, long long npad) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= npad) return;
  const BatchProblem bpr = bp[blk_prob[i / kBatchAlign]];
  labels[i] = bpr.coff + bpr.k;
}

// FP64 inertia of each problem's final state in a fixed order: every
// kBatchAlign-point block sums its points (a lane takes four dimensions of a
// point, a wave DP/4-lane groups of points, four steps' loads issued
// together), then one wave per problem adds its block sums.
constexpr int kBInertiaThreads = 256;
template <int DP>
__global__ __launch_bounds__(kBInertiaThreads) void k_binertia_blocks(
    const float* X, const float* C, const int* labels, const BatchProblem* bp,
    const int* blk_prob, double* bpart) {
  constexpr int LPP = DP / 4;    // lanes per point
  constexpr int PPS = 64 / LPP;  // points per wave step
  constexpr int U = 4;
  constexpr int NW = kBInertiaThreads / 64;
  constexpr int PW = kBatchAlign / NW;  // points per wave
  __shared__ double red_s[NW];
  const int b = blockIdx.x;
  const BatchProblem bpr = bp[blk_prob[b]];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane / LPP, j = 4 * (lane % LPP);
  const long long lo = (long long)b * kBatchAlign + (long long)wave * PW;
  const long long hi = min(lo + (long long)PW, bpr.poff + (long long)bpr.n);
  double s = 0.0;
  for (long long r0 = lo; r0 < hi; r0 += U * PPS) {
    float4 xv[U], cv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long r = r0 + u * PPS + sub;
      const long long rr = r < hi ? r : lo;
      xv[u] = *reinterpret_cast<const float4*>(X + (size_t)rr * DP + j);
      cv[u] = *reinterpret_cast<const float4*>(C + (size_t)labels[rr] * DP + j);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = r0 + u * PPS + sub < hi;
      const double t0 = (double)xv[u].x - (double)cv[u].x;
      const double t1 = (double)xv[u].y - (double)cv[u].y;
      const double t2 = (double)xv[u].z - (double)cv[u].z;
      const double t3 = (double)xv[u].w - (double)cv[u].w;
      const double q = fma(t3, t3, fma(t2, t2, fma(t1, t1, t0 * t0)));
      s = ok ? s + q : s;
    }
  }
  s = wave_sum_f64(s);
  if (lane == 0) red_s[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += red_s[w];
    bpart[b] = t;
  }
}

__global__ __launch_bounds__(256) void k_binertia_sum(const double* bpart, const BatchProblem* bp,
                                                      int nprob, double* out) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= nprob) return;
  const BatchProblem bpr = bp[p];
  const double* a = bpart + bpr.poff / kBatchAlign;
  double s = 0.0;
  for (int i = lane; i < bpr.nblk; i += 64) s += a[i];
  s = wave_sum_f64(s);
  if (lane == 0) out[p] = s;
}

}  // namespace kmc

using kmc::BatchProblem;
using kmc::DevState;

// One width class of a batch: the problems whose d rounds up to the same
// padded width DP share one packed point space and one launch per stage, so
// a problem of d=16 streams 64-byte rows, not the widest problem's.
struct BatchGroup {
  hipStream_t stream = nullptr;  // the group's iterations run on their own stream
  bool own_stream = true;        // false: the batch stream (the largest group)
  hipEvent_t end = nullptr;
  uint32_t nprob = 0;           // problems in this group
  std::vector<uint32_t> ids;    // their indices in the batch
  int DP = 0, prec = 0, idx_bits = 0, lbits = 0, cus = 0;
  float alpha = 0, beta = 0, gref = 0, xpk = 0;
  long long npad = 0;
  int KT = 0, KTp = 0, nops = 0, nitems = 0, nblk = 0, grid = 0;
  long long qstride = 0, nbcount = 0;
  size_t lds = 0;
  std::vector<BatchProblem> hp;
  std::vector<uint32_t> d;
  // device
  float *X = nullptr, *xn2 = nullptr, *C = nullptr, *C0 = nullptr, *Cnext = nullptr;
  float *Cw = nullptr, *cn = nullptr, *CT = nullptr;
  uint16_t *Xhi = nullptr, *Xlo = nullptr, *Chi = nullptr, *Clo = nullptr;
  int *labels = nullptr, *q2 = nullptr, *qa = nullptr;
  unsigned *q2cnt = nullptr, *qacnt = nullptr;
  int *bcount = nullptr, *count = nullptr, *start = nullptr, *members = nullptr;
  double *tpart = nullptr, *ipart = nullptr, *shift2 = nullptr, *inertia = nullptr;
  double* bpart = nullptr;
  int *conv = nullptr, *alive = nullptr;
  DevState* state = nullptr;
  int* pdone = nullptr;
  unsigned *piter = nullptr, *pconv = nullptr;
  BatchProblem* bp = nullptr;
  int *item_prob = nullptr, *blk_prob = nullptr, *clus_prob = nullptr, *row_prob = nullptr;
  // exact update: cluster totals, labels they stand for, per-cluster scales
  long long* Stot = nullptr;
  int *Ntot = nullptr, *prev = nullptr;
  float* cl_scale = nullptr;
  double* cl_inv_scale = nullptr;
};

// A captured chunk of batch iterations: every group's iterations as one
// branch of the graph (the groups are independent).
struct BatchGraph {
  uint64_t iters = 0, maxiter = 0;
  int check = 0;
  hipGraphExec_t exec = nullptr;
};

struct kmc_batch {
  kmc_batch_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;  // capture: the groups' branches start here
  std::string err;
  uint32_t nprob = 0;
  std::vector<BatchGroup*> groups;
  std::vector<int> grp_of, loc_of;  // group of each problem, index inside it
  std::vector<BatchGraph> graphs;
};

namespace {

kmc_status bfail(kmc_batch* b, kmc_status s, const std::string& msg) {
  if (b) b->err = msg;
  return s;
}

#define KMC_BCK(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return bfail(b, e_ == hipErrorOutOfMemory ? KMC_ERR_OOM : KMC_ERR_HIP,        \
                   std::string(#expr) + ": " + hipGetErrorString(e_));              \
  } while (0)

template <typename T>
hipError_t balloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

template <typename T>
hipError_t upload(T** p, const std::vector<T>& v, hipStream_t s) {
  hipError_t e = balloc(p, v.size());
  if (e == hipSuccess && !v.empty())
    e = hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
  // v may be a local of the caller that an early error return destroys
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

void gfree_all(BatchGroup* g) {
  void* ptrs[] = {g->X,     g->xn2,   g->C,       g->C0,      g->Cnext,   g->Cw,     g->cn,
                  g->CT,    g->Xhi,   g->Xlo,     g->Chi,     g->Clo,     g->labels, g->q2,
                  g->qa,    g->q2cnt, g->qacnt,   g->bcount,  g->count,   g->start,  g->members,
                  g->tpart, g->ipart, g->shift2,  g->inertia, g->conv,    g->alive,  g->state,
                  g->pdone, g->piter, g->pconv,   g->bp,      g->item_prob, g->blk_prob,
                  g->clus_prob, g->row_prob, g->bpart, g->Stot, g->Ntot,
                  g->prev,  g->cl_scale, g->cl_inv_scale};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

void bfree_all(kmc_batch* b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (BatchGraph& gr : b->graphs) (void)hipGraphExecDestroy(gr.exec);
  b->graphs.clear();
  if (b->fork) (void)hipEventDestroy(b->fork);
  b->fork = nullptr;
  for (BatchGroup* g : b->groups) {
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    gfree_all(g);
    if (g->end) (void)hipEventDestroy(g->end);
    if (g->stream && g->own_stream) (void)hipStreamDestroy(g->stream);
    delete g;
  }
  b->groups.clear();
}

int dp_of(uint32_t d) {
  const int opts[] = {16, 32, 64, 128, 256};
  for (int o : opts)
    if ((int)d <= o) return o;
  return 0;
}

kmc_status enqueue_iteration(kmc_batch* b, BatchGroup* g, uint64_t maxiter, int check) {
  const int* done = &g->state->done;
  kmc::BPrepParams pp{};
  pp.C = g->C;
  pp.Cw = g->Cw;
  pp.Chi = g->Chi;
  pp.Clo = g->Clo;
  pp.cn = g->cn;
  pp.CT = g->CT;
  pp.bp = g->bp;
  pp.row_prob = g->row_prob;
  pp.done = done;
  pp.nrows = g->nops;
  pp.KTp = g->KTp;
  pp.DP = g->DP;
  pp.prec = g->prec;
  hipLaunchKernelGGL(kmc::k_bprep, dim3((g->nops + 3) / 4), dim3(256), 0, g->stream, pp);
  KMC_BCK(hipGetLastError());

  kmc::AssignParams ap{};
  ap.X = g->X;
  ap.Xhi = g->Xhi;
  ap.Xlo = g->Xlo;
  ap.xn2 = g->xn2;
  ap.n = g->npad;
  ap.Cw = g->Cw;
  ap.Chi = g->Chi;
  ap.Clo = g->Clo;
  ap.cn = g->cn;
  ap.K = g->KT;
  ap.Kp = g->nops;
  ap.kchunk = g->nops;
  ap.idx_mask = (1u << g->idx_bits) - 1u;
  ap.alpha = g->alpha;
  ap.beta = g->beta;
  ap.gref = g->gref;
  ap.xpk = g->xpk;
  ap.labels = g->labels;
  ap.q2 = g->q2;
  ap.qa = g->qa;
  ap.q2cnt = g->q2cnt;
  ap.qacnt = g->qacnt;
  ap.qstride = g->qstride;
  ap.done = done;
  ap.bp = g->bp;
  ap.item_prob = g->item_prob;
  ap.pdone = g->pdone;
  ap.nitems = g->nitems;
  KMC_BCK(kmc::launch_assign_batch(g->DP, g->prec, ap, g->grid, g->lds, g->stream));

  kmc::RefineParams rp{};
  rp.X = g->X;
  rp.C = g->C;
  rp.CT = g->CT;
  rp.q2 = g->q2;
  rp.qa = g->qa;
  rp.q2cnt = g->q2cnt;
  rp.qacnt = g->qacnt;
  rp.qstride = g->qstride;
  rp.nregions = g->grid;
  rp.alive = g->alive;
  rp.labels = g->labels;
  rp.done = done;
  rp.K = g->KT;
  rp.Kp = g->KTp;
  rp.D = g->DP;
  rp.DP = g->DP;
  rp.bp = g->bp;
  rp.blk_prob = g->blk_prob;
  KMC_BCK(kmc::launch_refine(rp, g->stream));

  kmc::BSortParams sp{};
  sp.labels = g->labels;
  sp.bcount = g->bcount;
  sp.count = g->count;
  sp.start = g->start;
  sp.members = g->members;
  sp.bp = g->bp;
  sp.blk_prob = g->blk_prob;
  sp.clus_prob = g->clus_prob;
  sp.done = done;
  sp.nblk = g->nblk;
  sp.KT = g->KT;
  sp.lbits = g->lbits;
  sp.fx_state = g->state;
  sp.fx_n = g->npad;
  const size_t hl = (size_t)(1 << g->lbits) * sizeof(int);
  hipLaunchKernelGGL(kmc::k_bhist, dim3(g->nblk), dim3(64), hl, g->stream, sp);
  hipLaunchKernelGGL(kmc::k_bscan_blocks, dim3((g->KT + 3) / 4), dim3(256), 0, g->stream, sp);
  hipLaunchKernelGGL(kmc::k_bscatter, dim3(g->nblk), dim3(64), hl, g->stream, sp);
  KMC_BCK(hipGetLastError());

  kmc::UpdateParams up{};
  up.X = g->X;
  up.C = g->C;
  up.Cnext = g->Cnext;
  up.labels = g->labels;
  up.members = g->members;
  up.start = g->start;
  up.count = g->count;
  up.ipart = g->ipart;
  up.tpart = g->tpart;
  up.done = done;
  up.n = g->npad;
  up.K = g->KT;
  up.D = g->DP;
  up.DP = g->DP;
  // exact form: full passes over the sorted members, otherwise the moved
  // points only
  up.cl_scale = g->cl_scale;
  up.Stot = g->Stot;
  up.Ntot = g->Ntot;
  up.Qtot = nullptr;
  up.prev = g->prev;
  up.state = g->state;
  KMC_BCK(kmc::launch_update(up, kmc::UPDATE_TREE, g->stream));

  kmc::FinishParams fp{};
  fp.C = g->C;
  fp.Cnext = g->Cnext;
  fp.tpart = g->tpart;
  fp.start = g->start;
  fp.tree = 2;  // exact integer sums (full pass) or the cluster totals
  fp.Stot = g->Stot;
  fp.Ntot = g->Ntot;
  fp.cl_inv_scale = g->cl_inv_scale;
  fp.fx_state = g->state;
  fp.n = g->npad;
  fp.conv = g->conv;
  fp.shift2 = g->shift2;
  fp.count = g->count;
  fp.alive = g->alive;
  fp.done = done;
  fp.tol = b->cfg.tol;
  fp.K = g->KT;
  fp.D = g->DP;
  fp.DP = g->DP;
  fp.exclude_empty = 0;
  fp.do_prep = 0;
  fp.stats.state = nullptr;  // the batch's own stop test follows
  KMC_BCK(kmc::launch_finish(fp, g->stream));

  kmc::BStatsParams st{};
  st.conv = g->conv;
  st.bp = g->bp;
  st.pdone = g->pdone;
  st.piter = g->piter;
  st.pconv = g->pconv;
  st.state = g->state;
  st.done = done;
  st.nprob = (int)g->nprob;
  st.check = check;
  st.maxiter = maxiter;
  st.n = g->npad;
  hipLaunchKernelGGL(kmc::k_bstats, dim3(1), dim3(kmc::kBStatsThreads), 0, g->stream, st);
  KMC_BCK(hipGetLastError());
  return KMC_OK;
}


// Replays (capturing it on first use) a graph of m iterations of every group
// on the batch stream: the capture forks each group's stream from the batch
// stream and joins it back, so the groups' iterations are parallel branches.
// A group whose device stop flag is set makes its remaining kernels no-ops,
// as in the eager launches. KMC_ERR_UNSUPPORTED when the runtime refuses the
// capture (the caller then launches eagerly).
kmc_status launch_batch_graph(kmc_batch* b, uint64_t m, uint64_t maxiter, int check) {
  for (const BatchGraph& gr : b->graphs) {
    if (gr.iters == m && gr.maxiter == maxiter && gr.check == check) {
      KMC_BCK(hipGraphLaunch(gr.exec, b->stream));
      return KMC_OK;
    }
  }
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return KMC_ERR_UNSUPPORTED;
  kmc_status st = KMC_OK;
  if (hipEventRecord(b->fork, b->stream) != hipSuccess) st = KMC_ERR_UNSUPPORTED;
  // every branch is forked and filled before any is joined (one group may
  // run on the batch stream itself)
  for (BatchGroup* g : b->groups) {
    if (st != KMC_OK) break;
    if (hipStreamWaitEvent(g->stream, b->fork, 0) != hipSuccess) {
      st = KMC_ERR_UNSUPPORTED;
      break;
    }
    for (uint64_t i = 0; i < m && st == KMC_OK; ++i) st = enqueue_iteration(b, g, maxiter, check);
  }
  for (BatchGroup* g : b->groups) {
    if (st != KMC_OK) break;
    if (hipEventRecord(g->end, g->stream) != hipSuccess ||
        hipStreamWaitEvent(b->stream, g->end, 0) != hipSuccess)
      st = KMC_ERR_UNSUPPORTED;
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->stream, &graph);
  if (st != KMC_OK || ec != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return KMC_ERR_UNSUPPORTED;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    (void)hipGetLastError();
    return KMC_ERR_UNSUPPORTED;
  }
  b->graphs.push_back({m, maxiter, check, exec});
  KMC_BCK(hipGraphLaunch(exec, b->stream));
  return KMC_OK;
}

// Packs one width class: tables, launch shape, buffers, points and initial
// centroids, norms and bf16 planes.
kmc_status create_group(kmc_batch* b, BatchGroup* g, const std::vector<kmc_problem>& probs) {
  const uint32_t count = (uint32_t)probs.size();
  uint32_t dmax = 0, kmax = 0;
  for (const kmc_problem& q : probs) {
    dmax = std::max(dmax, q.d);
    kmax = std::max(kmax, q.k);
  }
  g->nprob = count;
  g->prec = b->cfg.precision;
  g->DP = dp_of(dmax);
  hipDeviceProp_t prop;
  KMC_BCK(hipGetDeviceProperties(&prop, b->dev));
  g->cus = prop.multiProcessorCount;
  // operand rows of the largest problem must fit the LDS with one copy
  const int kpad_max = (int)((kmax + 63) / 64 * 64);
  g->lds = kmc::assign_lds(g->DP, g->prec, kpad_max);
  if (g->lds > 160 * 1024)
    return bfail(b, KMC_ERR_UNSUPPORTED, "kmc_batch_create: centroid operands of the "
                                         "largest problem exceed the LDS");
  g->idx_bits = 0;
  while ((1 << g->idx_bits) < kpad_max) ++g->idx_bits;
  g->lbits = 0;
  while ((1 << g->lbits) < (int)kmax + 1) ++g->lbits;
  kmc::band_constants(g->prec, (int)dmax, g->idx_bits, &g->alpha, &g->beta, &g->gref, &g->xpk);

  // tables
  g->hp.resize(count);
  g->d.resize(count);
  long long poff = 0, bcoff = 0;
  int coff = 0, cpoff = 0, tile0 = 0;
  for (uint32_t p = 0; p < count; ++p) {
    const kmc_problem& q = probs[p];
    BatchProblem& e = g->hp[p];
    const long long npad = ((long long)q.n + kmc::kBatchAlign - 1) / kmc::kBatchAlign *
                           kmc::kBatchAlign;
    e.poff = poff;
    e.n = (int)q.n;
    e.k = (int)q.k;
    e.kpad = (int)((q.k + 63) / 64 * 64);
    e.coff = coff;
    e.cpoff = cpoff;
    e.tile0 = tile0;
    e.nblk = (int)(npad / kmc::kBatchAlign);
    e.bcoff = bcoff;
    g->d[p] = q.d;
    poff += npad;
    coff += (int)q.k + 1;
    cpoff += e.kpad;
    tile0 += (int)((q.n + kmc::kBatchTile - 1) / kmc::kBatchTile);
    bcoff += (long long)(q.k + 1) * e.nblk;
  }
  g->npad = poff;
  g->KT = coff;
  g->KTp = (coff + 63) / 64 * 64;
  g->nops = cpoff;
  g->nitems = tile0;
  g->nblk = (int)(poff / kmc::kBatchAlign);
  g->nbcount = bcoff;
  std::vector<int> item_prob(g->nitems), blk_prob(g->nblk), clus_prob(g->KT), row_prob(g->nops);
  // each problem's fixed-point scale, from its own n and max|x| exactly as a
  // handle of that problem alone chooses it (UPDATE_FIXED), so the exact sums
  // and the centroids are those of its lone run
  std::vector<float> cl_scale(g->KT);
  std::vector<double> cl_inv(g->KT);
  for (uint32_t p = 0; p < count; ++p) {
    const kmc_problem& q = probs[p];
    float am = 0.0f;
    for (uint64_t i = 0; i < q.n; ++i)
      for (uint32_t j = 0; j < q.d; ++j) am = std::max(am, std::fabs(q.X[i * q.ld + j]));
    const float sc = std::ldexp(1.0f, kmc::fx_scale_exp((size_t)q.n, am));
    for (int c = 0; c <= g->hp[p].k; ++c) {
      cl_scale[g->hp[p].coff + c] = sc;
      cl_inv[g->hp[p].coff + c] = 1.0 / (double)sc;
    }
  }
  for (uint32_t p = 0; p < count; ++p) {
    const BatchProblem& e = g->hp[p];
    const int nt = (e.n + kmc::kBatchTile - 1) / kmc::kBatchTile;
    for (int t = 0; t < nt; ++t) item_prob[e.tile0 + t] = (int)p;
    for (int k = 0; k < e.nblk; ++k) blk_prob[e.poff / kmc::kBatchAlign + k] = (int)p;
    for (int c = 0; c <= e.k; ++c) clus_prob[e.coff + c] = (int)p;
    for (int r = 0; r < e.kpad; ++r) row_prob[e.cpoff + r] = (int)p;
  }

  // launch shape of the distance kernel
  const void* fn = kmc::assign_batch_kernel_ptr(g->DP, g->prec);
  if (!fn) return bfail(b, KMC_ERR_UNSUPPORTED, "kmc_batch_create: no kernel for this shape");
  KMC_BCK(kmc::raise_lds_limit(fn, g->lds));
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kmc::kBatchAssignThreads,
                                                   g->lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  g->grid = (int)std::min<long long>(
      {(long long)g->nitems, (long long)per_cu * g->cus, (long long)kmc::kMaxRegions});
  g->qstride = ((long long)g->nitems + g->grid - 1) / g->grid * kmc::kBatchTile;

  // buffers
  const size_t DP = (size_t)g->DP, np = (size_t)g->npad;
  KMC_BCK(balloc(&g->X, np * DP));
  if (g->prec != kmc::PREC_FP32) KMC_BCK(balloc(&g->Xhi, np * DP));
  if (g->prec == kmc::PREC_BF16X3) KMC_BCK(balloc(&g->Xlo, np * DP));
  KMC_BCK(balloc(&g->xn2, np));
  KMC_BCK(balloc(&g->C, (size_t)g->KT * DP));
  KMC_BCK(balloc(&g->C0, (size_t)g->KT * DP));
  KMC_BCK(balloc(&g->Cnext, (size_t)g->KT * DP));
  if (g->prec == kmc::PREC_FP32) KMC_BCK(balloc(&g->Cw, (size_t)g->nops * DP));
  else KMC_BCK(balloc(&g->Chi, (size_t)g->nops * DP));
  if (g->prec == kmc::PREC_BF16X3) KMC_BCK(balloc(&g->Clo, (size_t)g->nops * DP));
  KMC_BCK(balloc(&g->cn, (size_t)g->nops));
  KMC_BCK(balloc(&g->CT, (size_t)g->KTp * DP));
  KMC_BCK(balloc(&g->labels, np));
  KMC_BCK(balloc(&g->q2, 4 * (size_t)g->grid * (size_t)g->qstride));
  KMC_BCK(balloc(&g->qa, (size_t)g->grid * (size_t)g->qstride));
  KMC_BCK(balloc(&g->q2cnt, (size_t)g->grid));
  KMC_BCK(balloc(&g->qacnt, (size_t)g->grid));
  KMC_BCK(balloc(&g->bcount, (size_t)g->nbcount));
  KMC_BCK(balloc(&g->count, (size_t)g->KT));
  KMC_BCK(balloc(&g->start, (size_t)g->KT));
  KMC_BCK(balloc(&g->members, np));
  KMC_BCK(balloc(&g->tpart, (np / kmc::kTreeRows + (size_t)g->KT) * DP));
  KMC_BCK(balloc(&g->ipart, np / kmc::kTreeRows + 1));
  KMC_BCK(balloc(&g->shift2, (size_t)g->KT));
  KMC_BCK(balloc(&g->conv, (size_t)g->KT));
  KMC_BCK(balloc(&g->alive, (size_t)g->KT));
  KMC_BCK(balloc(&g->inertia, (size_t)count));
  KMC_BCK(balloc(&g->bpart, (size_t)g->nblk));
  KMC_BCK(balloc(&g->state, 1));
  KMC_BCK(balloc(&g->pdone, (size_t)count));
  KMC_BCK(balloc(&g->piter, (size_t)count));
  KMC_BCK(balloc(&g->pconv, (size_t)count));
  KMC_BCK(upload(&g->bp, g->hp, b->stream));
  KMC_BCK(upload(&g->item_prob, item_prob, b->stream));
  KMC_BCK(upload(&g->blk_prob, blk_prob, b->stream));
  KMC_BCK(upload(&g->clus_prob, clus_prob, b->stream));
  KMC_BCK(upload(&g->row_prob, row_prob, b->stream));
  KMC_BCK(upload(&g->cl_scale, cl_scale, b->stream));
  KMC_BCK(upload(&g->cl_inv_scale, cl_inv, b->stream));
  KMC_BCK(balloc(&g->Stot, (size_t)g->KT * DP));
  KMC_BCK(balloc(&g->Ntot, (size_t)g->KT));
  KMC_BCK(balloc(&g->prev, np));
  KMC_BCK(kmc::launch_fill_i32(g->alive, (long long)g->KT, 1, b->stream));
  KMC_BCK(hipMemsetAsync(g->X, 0, np * DP * sizeof(float), b->stream));
  if (g->Xhi) KMC_BCK(hipMemsetAsync(g->Xhi, 0, np * DP * sizeof(uint16_t), b->stream));
  if (g->Xlo) KMC_BCK(hipMemsetAsync(g->Xlo, 0, np * DP * sizeof(uint16_t), b->stream));
  KMC_BCK(hipMemsetAsync(g->xn2, 0, np * sizeof(float), b->stream));
  KMC_BCK(hipMemsetAsync(g->CT, 0, (size_t)g->KTp * DP * sizeof(float), b->stream));
  KMC_BCK(hipMemsetAsync(g->q2cnt, 0, (size_t)g->grid * sizeof(unsigned), b->stream));
  KMC_BCK(hipMemsetAsync(g->qacnt, 0, (size_t)g->grid * sizeof(unsigned), b->stream));
  // initial centroids, packed on the host and copied at once
  std::vector<float> c0((size_t)g->KT * DP, 0.0f);
  for (uint32_t p = 0; p < count; ++p) {
    const kmc_problem& q = probs[p];
    const BatchProblem& e = g->hp[p];
    for (uint32_t c = 0; c < q.k; ++c)
      std::memcpy(&c0[((size_t)e.coff + c) * DP], q.C0 + (size_t)c * q.d, q.d * sizeof(float));
  }
  KMC_BCK(hipMemcpyAsync(g->C0, c0.data(), c0.size() * sizeof(float), hipMemcpyHostToDevice,
                         b->stream));
  KMC_BCK(hipStreamSynchronize(b->stream));  // c0 dies on an early error return
  // points: each problem's rows as one contiguous copy into a staging buffer
  // (a pitched copy from pageable memory goes row by row), then the norms,
  // padded rows and bf16 planes at the problem's offset
  size_t smax = 0;
  for (const kmc_problem& q : probs) smax = std::max(smax, (size_t)((q.n - 1) * q.ld + q.d));
  float* stage = nullptr;
  KMC_BCK(balloc(&stage, smax));
  kmc_status st = KMC_OK;
  for (uint32_t p = 0; p < count && st == KMC_OK; ++p) {
    const kmc_problem& q = probs[p];
    const BatchProblem& e = g->hp[p];
    hipError_t er = hipMemcpyAsync(stage, q.X, ((q.n - 1) * q.ld + q.d) * sizeof(float),
                                   hipMemcpyHostToDevice, b->stream);
    if (er == hipSuccess) {
      kmc::LoadParams lp{};
      lp.src = stage;
      lp.ld = q.ld;
      lp.n = (long long)q.n;
      lp.D = (int)q.d;
      lp.DP = g->DP;
      lp.X = g->X + (size_t)e.poff * DP;
      lp.Xhi = g->Xhi ? g->Xhi + (size_t)e.poff * DP : nullptr;
      lp.Xlo = g->Xlo ? g->Xlo + (size_t)e.poff * DP : nullptr;
      lp.xn2 = g->xn2 + e.poff;
      lp.xn2_64 = nullptr;
      er = kmc::launch_load_points(lp, b->stream);
    }
    if (er != hipSuccess)
      st = bfail(b, KMC_ERR_HIP, std::string("kmc_batch_create: upload: ") +
                                     hipGetErrorString(er));
  }
  const hipError_t es = hipStreamSynchronize(b->stream);
  (void)hipFree(stage);
  if (st != KMC_OK) return st;
  KMC_BCK(es);
  return KMC_OK;
}

}  // namespace

extern "C" {

kmc_status kmc_batch_create(const kmc_problem* probs, uint32_t count,
                            const kmc_batch_config* cfg, kmc_batch** out) {
  if (!probs || !cfg || !out || count == 0) return KMC_ERR_ARG;
  *out = nullptr;
  if (cfg->precision == kmc::PREC_FP8) return KMC_ERR_UNSUPPORTED;
  if (cfg->precision < kmc::PREC_FP32 || cfg->precision > kmc::PREC_BF16) return KMC_ERR_ARG;
  for (uint32_t p = 0; p < count; ++p) {
    const kmc_problem& q = probs[p];
    if (q.n == 0 || q.d == 0 || q.k == 0 || q.k > q.n || !q.X || !q.C0 || q.ld < q.d ||
        q.n > (uint64_t)0x7fffffff)
      return KMC_ERR_ARG;
    if (q.d > 256 || q.k > 256) return KMC_ERR_UNSUPPORTED;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return KMC_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev) return KMC_ERR_ARG;

  kmc_batch* b = new (std::nothrow) kmc_batch();
  if (!b) return KMC_ERR_OOM;
  b->cfg = *cfg;
  b->dev = cfg->device;
  b->nprob = count;
  b->grp_of.assign(count, -1);
  b->loc_of.assign(count, -1);
  auto body = [&]() -> kmc_status {
    KMC_BCK(hipSetDevice(b->dev));
    KMC_BCK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    KMC_BCK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    // one group per padded width, problems in batch order inside a group
    const uint32_t classes[] = {16, 32, 64, 128, 256};
    for (uint32_t dp : classes) {
      std::vector<kmc_problem> sub;
      std::vector<uint32_t> ids;
      for (uint32_t p = 0; p < count; ++p)
        if (dp_of(probs[p].d) == (int)dp) {
          sub.push_back(probs[p]);
          ids.push_back(p);
        }
      if (sub.empty()) continue;
      BatchGroup* g = new (std::nothrow) BatchGroup();
      if (!g) return bfail(b, KMC_ERR_OOM, "kmc_batch_create: out of host memory");
      b->groups.push_back(g);
      KMC_BCK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
      KMC_BCK(hipEventCreateWithFlags(&g->end, hipEventDisableTiming));
      g->ids = ids;
      for (size_t i = 0; i < ids.size(); ++i) {
        b->grp_of[ids[i]] = (int)b->groups.size() - 1;
        b->loc_of[ids[i]] = (int)i;
      }
      const kmc_status s = create_group(b, g, sub);
      if (s != KMC_OK) return s;
    }
    // The device runs a few hardware queues per process (streams beyond
    // that share one, and a shared queue runs its groups one after the
    // other): the group that streams the most bytes runs on the batch stream
    // instead of a stream of its own.
    BatchGroup* big = nullptr;
    for (BatchGroup* g : b->groups)
      if (!big || g->npad * g->DP > big->npad * big->DP) big = g;
    if (big && b->groups.size() > 1) {
      KMC_BCK(hipStreamSynchronize(big->stream));
      KMC_BCK(hipStreamDestroy(big->stream));
      big->stream = b->stream;
      big->own_stream = false;
    }
    return KMC_OK;
  };
  const kmc_status st = body();
  if (st != KMC_OK) {
    fprintf(stderr, "kmc_batch_create: %s\n", b->err.c_str());
    bfree_all(b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return st;
  }
  *out = b;
  return KMC_OK;
}

kmc_status kmc_batch_run(kmc_batch* b, uint64_t maxiter, uint32_t flags, kmc_batch_result* res,
                         double* ms_total) {
  if (!b) return KMC_ERR_ARG;
  KMC_BCK(hipSetDevice(b->dev));
  const int check = (flags & KMC_RUN_CHECK_CONVERGED) ? 1 : 0;
  // every run starts from the initial centroids
  hipEvent_t t0, t1;
  KMC_BCK(hipEventCreate(&t0));
  KMC_BCK(hipEventCreate(&t1));
  KMC_BCK(hipEventRecord(t0, b->stream));
  // the groups are independent: each runs on its own stream from t0 on
  for (BatchGroup* g : b->groups) {
    const size_t DP = (size_t)g->DP;
    KMC_BCK(hipStreamWaitEvent(g->stream, t0, 0));
    KMC_BCK(hipMemcpyAsync(g->C, g->C0, (size_t)g->KT * DP * sizeof(float),
                           hipMemcpyDeviceToDevice, g->stream));
    KMC_BCK(hipMemcpyAsync(g->Cnext, g->C0, (size_t)g->KT * DP * sizeof(float),
                           hipMemcpyDeviceToDevice, g->stream));
    hipLaunchKernelGGL(kmc::k_binit_labels, dim3((unsigned)((g->npad + 255) / 256)), dim3(256),
                       0, g->stream, g->labels, g->bp, g->blk_prob, g->npad);
    KMC_BCK(hipGetLastError());
    KMC_BCK(hipMemsetAsync(g->pdone, 0, g->nprob * sizeof(int), g->stream));
    KMC_BCK(hipMemsetAsync(g->piter, 0, g->nprob * sizeof(unsigned), g->stream));
    KMC_BCK(hipMemsetAsync(g->pconv, 0, g->nprob * sizeof(unsigned), g->stream));
    KMC_BCK(hipMemsetAsync(g->state, 0, sizeof(DevState), g->stream));
  }
  kmc_status st = KMC_OK;
  const size_t ng = b->groups.size();
  bool use_graph = (flags & KMC_RUN_GRAPH) != 0;
  if (use_graph) {
    // the graph runs on the batch stream: it starts after every group's set-up
    for (BatchGroup* g : b->groups) {
      KMC_BCK(hipEventRecord(g->end, g->stream));
      KMC_BCK(hipStreamWaitEvent(b->stream, g->end, 0));
    }
  }
  if (maxiter > 0) {
    // chunks of iterations between host looks at the stop flags (a group's
    // device flag makes its queued iterations after the stop no-ops)
    std::vector<DevState> hs(ng);
    std::vector<char> gdone(ng, 0);
    uint64_t launched = 0;
    while (launched < maxiter && st == KMC_OK) {
      const uint64_t m = std::min<uint64_t>(16, maxiter - launched);
      if (use_graph) {
        st = launch_batch_graph(b, m, maxiter, check);
        if (st == KMC_ERR_UNSUPPORTED) {
          // capture refused: launch eagerly from here on, each group on its
          // own stream after the batch stream's work so far
          use_graph = false;
          st = KMC_OK;
          KMC_BCK(hipEventRecord(b->fork, b->stream));
          for (BatchGroup* g : b->groups) KMC_BCK(hipStreamWaitEvent(g->stream, b->fork, 0));
        }
        if (st != KMC_OK) break;
        if (use_graph) {
          for (size_t gi = 0; gi < ng; ++gi)
            KMC_BCK(hipMemcpyAsync(&hs[gi], b->groups[gi]->state, sizeof(DevState),
                                   hipMemcpyDeviceToHost, b->stream));
          KMC_BCK(hipStreamSynchronize(b->stream));
        }
      }
      if (!use_graph) {
        for (uint64_t i = 0; i < m && st == KMC_OK; ++i)
          for (size_t gi = 0; gi < ng && st == KMC_OK; ++gi)
            if (!gdone[gi]) st = enqueue_iteration(b, b->groups[gi], maxiter, check);
        if (st != KMC_OK) break;
        for (size_t gi = 0; gi < ng; ++gi)
          if (!gdone[gi])
            KMC_BCK(hipMemcpyAsync(&hs[gi], b->groups[gi]->state, sizeof(DevState),
                                   hipMemcpyDeviceToHost, b->groups[gi]->stream));
        for (size_t gi = 0; gi < ng; ++gi)
          if (!gdone[gi]) KMC_BCK(hipStreamSynchronize(b->groups[gi]->stream));
      }
      launched += m;
      bool all = true;
      for (size_t gi = 0; gi < ng; ++gi) {
        if (!gdone[gi] && hs[gi].done) gdone[gi] = 1;
        all = all && gdone[gi];
      }
      if (all) break;
    }
  }
  if (st == KMC_OK && use_graph) {
    // the per-group tails below run on the groups' streams after the graphs
    KMC_BCK(hipEventRecord(b->fork, b->stream));
    for (BatchGroup* g : b->groups) KMC_BCK(hipStreamWaitEvent(g->stream, b->fork, 0));
  }
  if (st == KMC_OK) {
    for (BatchGroup* g : b->groups) {
      const dim3 grid(g->nblk), block(kmc::kBInertiaThreads);
      switch (g->DP) {
        case 16: hipLaunchKernelGGL(kmc::k_binertia_blocks<16>, grid, block, 0, g->stream,
                                    g->X, g->C, g->labels, g->bp, g->blk_prob, g->bpart); break;
        case 32: hipLaunchKernelGGL(kmc::k_binertia_blocks<32>, grid, block, 0, g->stream,
                                    g->X, g->C, g->labels, g->bp, g->blk_prob, g->bpart); break;
        case 64: hipLaunchKernelGGL(kmc::k_binertia_blocks<64>, grid, block, 0, g->stream,
                                    g->X, g->C, g->labels, g->bp, g->blk_prob, g->bpart); break;
        case 128: hipLaunchKernelGGL(kmc::k_binertia_blocks<128>, grid, block, 0, g->stream,
                                     g->X, g->C, g->labels, g->bp, g->blk_prob, g->bpart); break;
        default: hipLaunchKernelGGL(kmc::k_binertia_blocks<256>, grid, block, 0, g->stream,
                                    g->X, g->C, g->labels, g->bp, g->blk_prob, g->bpart); break;
      }
      hipLaunchKernelGGL(kmc::k_binertia_sum, dim3((g->nprob + 3) / 4), dim3(256), 0,
                         g->stream, g->bpart, g->bp, (int)g->nprob, g->inertia);
      KMC_BCK(hipGetLastError());
    }
    for (BatchGroup* g : b->groups) {
      KMC_BCK(hipEventRecord(g->end, g->stream));
      KMC_BCK(hipStreamWaitEvent(b->stream, g->end, 0));
    }
    KMC_BCK(hipEventRecord(t1, b->stream));
    KMC_BCK(hipEventSynchronize(t1));
    float ms = 0.f;
    KMC_BCK(hipEventElapsedTime(&ms, t0, t1));
    if (ms_total) *ms_total = ms;
    if (res) {
      for (BatchGroup* g : b->groups) {
        std::vector<unsigned> it(g->nprob), cv(g->nprob);
        std::vector<double> in(g->nprob);
        KMC_BCK(hipMemcpy(it.data(), g->piter, g->nprob * sizeof(unsigned),
                          hipMemcpyDeviceToHost));
        KMC_BCK(hipMemcpy(cv.data(), g->pconv, g->nprob * sizeof(unsigned),
                          hipMemcpyDeviceToHost));
        KMC_BCK(hipMemcpy(in.data(), g->inertia, g->nprob * sizeof(double),
                          hipMemcpyDeviceToHost));
        for (uint32_t p = 0; p < g->nprob; ++p) {
          kmc_batch_result& r = res[g->ids[p]];
          r.iterations = it[p];
          r.converged_at = cv[p];
          r.inertia = in[p];
        }
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return st;
}

kmc_status kmc_batch_get_labels(kmc_batch* b, uint32_t problem, int32_t* labels) {
  if (!b || !labels || problem >= b->nprob) return bfail(b, KMC_ERR_ARG, "kmc_batch_get_labels");
  KMC_BCK(hipSetDevice(b->dev));
  const BatchGroup* g = b->groups[b->grp_of[problem]];
  const BatchProblem& e = g->hp[b->loc_of[problem]];
  KMC_BCK(hipMemcpy(labels, g->labels + e.poff, (size_t)e.n * sizeof(int32_t),
                    hipMemcpyDeviceToHost));
  for (int i = 0; i < e.n; ++i) labels[i] -= e.coff;
  return KMC_OK;
}

kmc_status kmc_batch_get_centroids(kmc_batch* b, uint32_t problem, float* C) {
  if (!b || !C || problem >= b->nprob) return bfail(b, KMC_ERR_ARG, "kmc_batch_get_centroids");
  KMC_BCK(hipSetDevice(b->dev));
  const BatchGroup* g = b->groups[b->grp_of[problem]];
  const int loc = b->loc_of[problem];
  const BatchProblem& e = g->hp[loc];
  const size_t d = g->d[loc];
  KMC_BCK(kmc::rows_to_host(C, d, g->C + (size_t)e.coff * g->DP, (size_t)g->DP, d,
                            (size_t)e.k, b->stream));
  return KMC_OK;
}

const char* kmc_batch_last_error(const kmc_batch* b) { return b ? b->err.c_str() : ""; }

kmc_status kmc_batch_destroy(kmc_batch* b) {
  if (!b) return KMC_ERR_ARG;
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  bfree_all(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return KMC_OK;
}

}  // extern "C"
