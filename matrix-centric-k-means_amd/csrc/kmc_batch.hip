// kmc_batch.hip -- many independent Lloyd problems in one device layout
// (kmc_batch_* of include/kmc.h).
//
// Layout: the problems' points are packed into one point space, problem p at
// rows [poff_p, poff_p + n_p) with poff_p a multiple of kBatchAlign; the rows
// up to the next problem are zero and carry the problem's extra "dead"
// cluster (local index k_p), so every stage can run over the whole packed
// space and the stable sort places problem p's members at the same offsets
// (relative to poff_p) as a run of that problem alone. Master clusters of p
// are [coff_p, coff_p + k_p]; the distance operands of its k_p real clusters
// are rows [cpoff_p, cpoff_p + kpad_p) (padding rows carry kFar).
//
// One iteration, one launch per stage for the whole batch:
//   k_bprep      distance operands + transposed centroids of every problem
//   k_assign<..., BATCH>  persistent distance kernel over the packed
//                (problem, tile) queue; converged problems are skipped
//   k_refine2 / k_refine_all   exact refinement (the all-k pass restricted
//                to the point's problem)
//   k_bhist / k_bscan_blocks / k_bscatter   stable counting sort, blocks of
//                kBatchAlign points never straddle problems
//   k_update_tree, k_finish    FP64 tree update and shift (shared kernels)
//   k_bstats     per-problem shift test and iteration counter, stop flags
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kmc.h"
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
};

// Block histograms: one wave per kBatchAlign points, local labels, the counts
// stored cluster-major per problem.
__global__ __launch_bounds__(64) void k_bhist(BSortParams P) {
  if (*P.done) return;
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
};

// Per-problem iteration counter and shift test (reference: tested from
// iteration 2, comparisons/CPU_Kmeans/src/kmeans.cpp:205-208); a problem that
// passes or reaches maxiter is frozen; the batch stops when all are.
constexpr int kBStatsThreads = 1024;
__global__ __launch_bounds__(kBStatsThreads) void k_bstats(BStatsParams P) {
  if (*P.done) return;
  __shared__ int active_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) active_s = 0;
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

// FP64 inertia of each problem's final state, one workgroup per problem,
// fixed summation order.
__global__ __launch_bounds__(256) void k_binertia(const float* X, const float* C,
                                                  const int* labels, const BatchProblem* bp,
                                                  int DP, double* out) {
  __shared__ double red_s[256];
  const BatchProblem bpr = bp[blockIdx.x];
  double s = 0.0;
  for (long long i = threadIdx.x; i < bpr.n; i += 256) {
    const long long r = bpr.poff + i;
    const float* x = X + (size_t)r * DP;
    const float* c = C + (size_t)labels[r] * DP;
    for (int j = 0; j < DP; ++j) {
      const double t = (double)x[j] - (double)c[j];
      s = fma(t, t, s);
    }
  }
  red_s[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red_s[threadIdx.x] += red_s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red_s[0];
}

}  // namespace kmc

using kmc::BatchProblem;
using kmc::DevState;

struct kmc_batch {
  kmc_batch_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  std::string err;
  uint32_t nprob = 0;
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
  int *conv = nullptr, *alive = nullptr;
  DevState* state = nullptr;
  int* pdone = nullptr;
  unsigned *piter = nullptr, *pconv = nullptr;
  BatchProblem* bp = nullptr;
  int *item_prob = nullptr, *blk_prob = nullptr, *clus_prob = nullptr, *row_prob = nullptr;
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
  return e;
}

void bfree_all(kmc_batch* b) {
  void* ptrs[] = {b->X,     b->xn2,   b->C,       b->C0,      b->Cnext,   b->Cw,     b->cn,
                  b->CT,    b->Xhi,   b->Xlo,     b->Chi,     b->Clo,     b->labels, b->q2,
                  b->qa,    b->q2cnt, b->qacnt,   b->bcount,  b->count,   b->start,  b->members,
                  b->tpart, b->ipart, b->shift2,  b->inertia, b->conv,    b->alive,  b->state,
                  b->pdone, b->piter, b->pconv,   b->bp,      b->item_prob, b->blk_prob,
                  b->clus_prob, b->row_prob};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

int dp_of(uint32_t d) {
  const int opts[] = {16, 32, 64, 128, 256};
  for (int o : opts)
    if ((int)d <= o) return o;
  return 0;
}

kmc_status enqueue_iteration(kmc_batch* b, uint64_t maxiter, int check) {
  const int* done = &b->state->done;
  kmc::BPrepParams pp{};
  pp.C = b->C;
  pp.Cw = b->Cw;
  pp.Chi = b->Chi;
  pp.Clo = b->Clo;
  pp.cn = b->cn;
  pp.CT = b->CT;
  pp.bp = b->bp;
  pp.row_prob = b->row_prob;
  pp.done = done;
  pp.nrows = b->nops;
  pp.KTp = b->KTp;
  pp.DP = b->DP;
  pp.prec = b->prec;
  hipLaunchKernelGGL(kmc::k_bprep, dim3((b->nops + 3) / 4), dim3(256), 0, b->stream, pp);
  KMC_BCK(hipGetLastError());

  kmc::AssignParams ap{};
  ap.X = b->X;
  ap.Xhi = b->Xhi;
  ap.Xlo = b->Xlo;
  ap.xn2 = b->xn2;
  ap.n = b->npad;
  ap.Cw = b->Cw;
  ap.Chi = b->Chi;
  ap.Clo = b->Clo;
  ap.cn = b->cn;
  ap.K = b->KT;
  ap.Kp = b->nops;
  ap.kchunk = b->nops;
  ap.idx_mask = (1u << b->idx_bits) - 1u;
  ap.alpha = b->alpha;
  ap.beta = b->beta;
  ap.gref = b->gref;
  ap.xpk = b->xpk;
  ap.labels = b->labels;
  ap.q2 = b->q2;
  ap.qa = b->qa;
  ap.q2cnt = b->q2cnt;
  ap.qacnt = b->qacnt;
  ap.qstride = b->qstride;
  ap.done = done;
  ap.bp = b->bp;
  ap.item_prob = b->item_prob;
  ap.pdone = b->pdone;
  ap.nitems = b->nitems;
  KMC_BCK(kmc::launch_assign_batch(b->DP, b->prec, ap, b->grid, b->lds, b->stream));

  kmc::RefineParams rp{};
  rp.X = b->X;
  rp.C = b->C;
  rp.CT = b->CT;
  rp.q2 = b->q2;
  rp.qa = b->qa;
  rp.q2cnt = b->q2cnt;
  rp.qacnt = b->qacnt;
  rp.qstride = b->qstride;
  rp.nregions = b->grid;
  rp.alive = b->alive;
  rp.labels = b->labels;
  rp.done = done;
  rp.K = b->KT;
  rp.Kp = b->KTp;
  rp.D = b->DP;
  rp.DP = b->DP;
  rp.bp = b->bp;
  rp.blk_prob = b->blk_prob;
  KMC_BCK(kmc::launch_refine(rp, b->stream));

  kmc::BSortParams sp{};
  sp.labels = b->labels;
  sp.bcount = b->bcount;
  sp.count = b->count;
  sp.start = b->start;
  sp.members = b->members;
  sp.bp = b->bp;
  sp.blk_prob = b->blk_prob;
  sp.clus_prob = b->clus_prob;
  sp.done = done;
  sp.nblk = b->nblk;
  sp.KT = b->KT;
  sp.lbits = b->lbits;
  const size_t hl = (size_t)(1 << b->lbits) * sizeof(int);
  hipLaunchKernelGGL(kmc::k_bhist, dim3(b->nblk), dim3(64), hl, b->stream, sp);
  hipLaunchKernelGGL(kmc::k_bscan_blocks, dim3((b->KT + 3) / 4), dim3(256), 0, b->stream, sp);
  hipLaunchKernelGGL(kmc::k_bscatter, dim3(b->nblk), dim3(64), hl, b->stream, sp);
  KMC_BCK(hipGetLastError());

  kmc::UpdateParams up{};
  up.X = b->X;
  up.C = b->C;
  up.Cnext = b->Cnext;
  up.labels = b->labels;
  up.members = b->members;
  up.start = b->start;
  up.count = b->count;
  up.ipart = b->ipart;
  up.tpart = b->tpart;
  up.done = done;
  up.n = b->npad;
  up.K = b->KT;
  up.D = b->DP;
  up.DP = b->DP;
  KMC_BCK(kmc::launch_update(up, kmc::UPDATE_TREE, b->stream));

  kmc::FinishParams fp{};
  fp.C = b->C;
  fp.Cnext = b->Cnext;
  fp.tpart = b->tpart;
  fp.start = b->start;
  fp.tree = 1;
  fp.conv = b->conv;
  fp.shift2 = b->shift2;
  fp.count = b->count;
  fp.alive = b->alive;
  fp.done = done;
  fp.tol = b->cfg.tol;
  fp.K = b->KT;
  fp.D = b->DP;
  fp.DP = b->DP;
  fp.exclude_empty = 0;
  fp.do_prep = 0;
  fp.stats.state = nullptr;  // the batch's own stop test follows
  KMC_BCK(kmc::launch_finish(fp, b->stream));

  kmc::BStatsParams st{};
  st.conv = b->conv;
  st.bp = b->bp;
  st.pdone = b->pdone;
  st.piter = b->piter;
  st.pconv = b->pconv;
  st.state = b->state;
  st.done = done;
  st.nprob = (int)b->nprob;
  st.check = check;
  st.maxiter = maxiter;
  hipLaunchKernelGGL(kmc::k_bstats, dim3(1), dim3(kmc::kBStatsThreads), 0, b->stream, st);
  KMC_BCK(hipGetLastError());
  return KMC_OK;
}

}  // namespace

extern "C" {

kmc_status kmc_batch_create(const kmc_problem* probs, uint32_t count,
                            const kmc_batch_config* cfg, kmc_batch** out) {
  if (!probs || !cfg || !out || count == 0) return KMC_ERR_ARG;
  *out = nullptr;
  if (cfg->precision < kmc::PREC_FP32 || cfg->precision > kmc::PREC_BF16) return KMC_ERR_ARG;
  uint32_t dmax = 0, kmax = 0;
  for (uint32_t p = 0; p < count; ++p) {
    const kmc_problem& q = probs[p];
    if (q.n == 0 || q.d == 0 || q.k == 0 || q.k > q.n || !q.X || !q.C0 || q.ld < q.d ||
        q.n > (uint64_t)0x7fffffff)
      return KMC_ERR_ARG;
    if (q.d > 256 || q.k > 256) return KMC_ERR_UNSUPPORTED;
    dmax = std::max(dmax, q.d);
    kmax = std::max(kmax, q.k);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return KMC_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev) return KMC_ERR_ARG;

  kmc_batch* b = new (std::nothrow) kmc_batch();
  if (!b) return KMC_ERR_OOM;
  b->cfg = *cfg;
  b->dev = cfg->device;
  b->nprob = count;
  b->prec = cfg->precision;
  b->DP = dp_of(dmax);
  kmc_status st = KMC_OK;
  auto body = [&]() -> kmc_status {
    KMC_BCK(hipSetDevice(b->dev));
    hipDeviceProp_t prop;
    KMC_BCK(hipGetDeviceProperties(&prop, b->dev));
    b->cus = prop.multiProcessorCount;
    KMC_BCK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    // operand rows of the largest problem must fit the LDS with one copy
    const int kpad_max = (int)((kmax + 63) / 64 * 64);
    b->lds = kmc::assign_lds(b->DP, b->prec, kpad_max);
    if (b->lds > 160 * 1024)
      return bfail(b, KMC_ERR_UNSUPPORTED, "kmc_batch_create: centroid operands of the "
                                           "largest problem exceed the LDS");
    b->idx_bits = 0;
    while ((1 << b->idx_bits) < kpad_max) ++b->idx_bits;
    b->lbits = 0;
    while ((1 << b->lbits) < (int)kmax + 1) ++b->lbits;
    kmc::band_constants(b->prec, (int)dmax, b->idx_bits, &b->alpha, &b->beta, &b->gref,
                        &b->xpk);

    // tables
    b->hp.resize(count);
    b->d.resize(count);
    long long poff = 0, bcoff = 0;
    int coff = 0, cpoff = 0, tile0 = 0;
    for (uint32_t p = 0; p < count; ++p) {
      const kmc_problem& q = probs[p];
      BatchProblem& e = b->hp[p];
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
      b->d[p] = q.d;
      poff += npad;
      coff += (int)q.k + 1;
      cpoff += e.kpad;
      tile0 += (int)((q.n + kmc::kBatchTile - 1) / kmc::kBatchTile);
      bcoff += (long long)(q.k + 1) * e.nblk;
    }
    b->npad = poff;
    b->KT = coff;
    b->KTp = (coff + 63) / 64 * 64;
    b->nops = cpoff;
    b->nitems = tile0;
    b->nblk = (int)(poff / kmc::kBatchAlign);
    b->nbcount = bcoff;
    std::vector<int> item_prob(b->nitems), blk_prob(b->nblk), clus_prob(b->KT),
        row_prob(b->nops);
    for (uint32_t p = 0; p < count; ++p) {
      const BatchProblem& e = b->hp[p];
      const int nt = (e.n + kmc::kBatchTile - 1) / kmc::kBatchTile;
      for (int t = 0; t < nt; ++t) item_prob[e.tile0 + t] = (int)p;
      for (int k = 0; k < e.nblk; ++k) blk_prob[e.poff / kmc::kBatchAlign + k] = (int)p;
      for (int c = 0; c <= e.k; ++c) clus_prob[e.coff + c] = (int)p;
      for (int r = 0; r < e.kpad; ++r) row_prob[e.cpoff + r] = (int)p;
    }

    // launch shape of the distance kernel
    const void* fn = kmc::assign_batch_kernel_ptr(b->DP, b->prec);
    if (!fn) return bfail(b, KMC_ERR_UNSUPPORTED, "kmc_batch_create: no kernel for this shape");
    KMC_BCK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds));
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kmc::kBatchAssignThreads,
                                                     b->lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    b->grid = (int)std::min<long long>(
        {(long long)b->nitems, (long long)per_cu * b->cus, (long long)kmc::kMaxRegions});
    b->qstride = ((long long)b->nitems + b->grid - 1) / b->grid * kmc::kBatchTile;

    // buffers
    const size_t DP = (size_t)b->DP, np = (size_t)b->npad;
    KMC_BCK(balloc(&b->X, np * DP));
    if (b->prec != kmc::PREC_FP32) KMC_BCK(balloc(&b->Xhi, np * DP));
    if (b->prec == kmc::PREC_BF16X3) KMC_BCK(balloc(&b->Xlo, np * DP));
    KMC_BCK(balloc(&b->xn2, np));
    KMC_BCK(balloc(&b->C, (size_t)b->KT * DP));
    KMC_BCK(balloc(&b->C0, (size_t)b->KT * DP));
    KMC_BCK(balloc(&b->Cnext, (size_t)b->KT * DP));
    if (b->prec == kmc::PREC_FP32) KMC_BCK(balloc(&b->Cw, (size_t)b->nops * DP));
    else KMC_BCK(balloc(&b->Chi, (size_t)b->nops * DP));
    if (b->prec == kmc::PREC_BF16X3) KMC_BCK(balloc(&b->Clo, (size_t)b->nops * DP));
    KMC_BCK(balloc(&b->cn, (size_t)b->nops));
    KMC_BCK(balloc(&b->CT, (size_t)b->KTp * DP));
    KMC_BCK(balloc(&b->labels, np));
    KMC_BCK(balloc(&b->q2, 4 * (size_t)b->grid * (size_t)b->qstride));
    KMC_BCK(balloc(&b->qa, (size_t)b->grid * (size_t)b->qstride));
    KMC_BCK(balloc(&b->q2cnt, (size_t)b->grid));
    KMC_BCK(balloc(&b->qacnt, (size_t)b->grid));
    KMC_BCK(balloc(&b->bcount, (size_t)b->nbcount));
    KMC_BCK(balloc(&b->count, (size_t)b->KT));
    KMC_BCK(balloc(&b->start, (size_t)b->KT));
    KMC_BCK(balloc(&b->members, np));
    KMC_BCK(balloc(&b->tpart, (np / kmc::kTreeRows + (size_t)b->KT) * DP));
    KMC_BCK(balloc(&b->ipart, np / kmc::kTreeRows + 1));
    KMC_BCK(balloc(&b->shift2, (size_t)b->KT));
    KMC_BCK(balloc(&b->conv, (size_t)b->KT));
    KMC_BCK(balloc(&b->alive, (size_t)b->KT));
    KMC_BCK(balloc(&b->inertia, (size_t)count));
    KMC_BCK(balloc(&b->state, 1));
    KMC_BCK(balloc(&b->pdone, (size_t)count));
    KMC_BCK(balloc(&b->piter, (size_t)count));
    KMC_BCK(balloc(&b->pconv, (size_t)count));
    KMC_BCK(upload(&b->bp, b->hp, b->stream));
    KMC_BCK(upload(&b->item_prob, item_prob, b->stream));
    KMC_BCK(upload(&b->blk_prob, blk_prob, b->stream));
    KMC_BCK(upload(&b->clus_prob, clus_prob, b->stream));
    KMC_BCK(upload(&b->row_prob, row_prob, b->stream));
    KMC_BCK(kmc::launch_fill_i32(b->alive, (long long)b->KT, 1, b->stream));
    KMC_BCK(hipMemsetAsync(b->X, 0, np * DP * sizeof(float), b->stream));
    KMC_BCK(hipMemsetAsync(b->C0, 0, (size_t)b->KT * DP * sizeof(float), b->stream));
    KMC_BCK(hipMemsetAsync(b->CT, 0, (size_t)b->KTp * DP * sizeof(float), b->stream));
    KMC_BCK(hipMemsetAsync(b->q2cnt, 0, (size_t)b->grid * sizeof(unsigned), b->stream));
    KMC_BCK(hipMemsetAsync(b->qacnt, 0, (size_t)b->grid * sizeof(unsigned), b->stream));
    // points and initial centroids of every problem
    for (uint32_t p = 0; p < count; ++p) {
      const kmc_problem& q = probs[p];
      const BatchProblem& e = b->hp[p];
      KMC_BCK(hipMemcpy2DAsync(b->X + (size_t)e.poff * DP, DP * sizeof(float), q.X,
                               q.ld * sizeof(float), q.d * sizeof(float), q.n,
                               hipMemcpyHostToDevice, b->stream));
      KMC_BCK(hipMemcpy2DAsync(b->C0 + (size_t)e.coff * DP, DP * sizeof(float), q.C0,
                               q.d * sizeof(float), q.d * sizeof(float), q.k,
                               hipMemcpyHostToDevice, b->stream));
    }
    // norms and bf16 planes, in place over the packed rows
    kmc::LoadParams lp{};
    lp.src = b->X;
    lp.ld = DP;
    lp.n = b->npad;
    lp.D = b->DP;
    lp.DP = b->DP;
    lp.X = b->X;
    lp.Xhi = b->Xhi;
    lp.Xlo = b->Xlo;
    lp.xn2 = b->xn2;
    lp.xn2_64 = nullptr;
    KMC_BCK(kmc::launch_load_points(lp, b->stream));
    KMC_BCK(hipStreamSynchronize(b->stream));
    return KMC_OK;
  };
  st = body();
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
  const size_t DP = (size_t)b->DP;
  // every run starts from the initial centroids
  KMC_BCK(hipMemcpyAsync(b->C, b->C0, (size_t)b->KT * DP * sizeof(float),
                         hipMemcpyDeviceToDevice, b->stream));
  KMC_BCK(hipMemcpyAsync(b->Cnext, b->C0, (size_t)b->KT * DP * sizeof(float),
                         hipMemcpyDeviceToDevice, b->stream));
  hipLaunchKernelGGL(kmc::k_binit_labels, dim3((unsigned)((b->npad + 255) / 256)), dim3(256), 0,
                     b->stream, b->labels, b->bp, b->blk_prob, b->npad);
  KMC_BCK(hipGetLastError());
  KMC_BCK(hipMemsetAsync(b->pdone, 0, b->nprob * sizeof(int), b->stream));
  KMC_BCK(hipMemsetAsync(b->piter, 0, b->nprob * sizeof(unsigned), b->stream));
  KMC_BCK(hipMemsetAsync(b->pconv, 0, b->nprob * sizeof(unsigned), b->stream));
  KMC_BCK(hipMemsetAsync(b->state, 0, sizeof(DevState), b->stream));
  hipEvent_t t0, t1;
  KMC_BCK(hipEventCreate(&t0));
  KMC_BCK(hipEventCreate(&t1));
  KMC_BCK(hipEventRecord(t0, b->stream));
  kmc_status st = KMC_OK;
  if (maxiter > 0) {
    // chunks of iterations between host looks at the stop flag (the device
    // flag makes the queued iterations after the stop no-ops)
    uint64_t launched = 0;
    while (launched < maxiter && st == KMC_OK) {
      const uint64_t m = std::min<uint64_t>(16, maxiter - launched);
      for (uint64_t i = 0; i < m && st == KMC_OK; ++i) st = enqueue_iteration(b, maxiter, check);
      launched += m;
      if (st != KMC_OK) break;
      DevState s;
      KMC_BCK(hipMemcpyAsync(&s, b->state, sizeof(DevState), hipMemcpyDeviceToHost, b->stream));
      KMC_BCK(hipStreamSynchronize(b->stream));
      if (s.done) break;
    }
  }
  if (st == KMC_OK) {
    hipLaunchKernelGGL(kmc::k_binertia, dim3(b->nprob), dim3(256), 0, b->stream, b->X, b->C,
                       b->labels, b->bp, b->DP, b->inertia);
    KMC_BCK(hipGetLastError());
    KMC_BCK(hipEventRecord(t1, b->stream));
    KMC_BCK(hipEventSynchronize(t1));
    float ms = 0.f;
    KMC_BCK(hipEventElapsedTime(&ms, t0, t1));
    if (ms_total) *ms_total = ms;
    if (res) {
      std::vector<unsigned> it(b->nprob), cv(b->nprob);
      std::vector<double> in(b->nprob);
      KMC_BCK(hipMemcpy(it.data(), b->piter, b->nprob * sizeof(unsigned), hipMemcpyDeviceToHost));
      KMC_BCK(hipMemcpy(cv.data(), b->pconv, b->nprob * sizeof(unsigned), hipMemcpyDeviceToHost));
      KMC_BCK(hipMemcpy(in.data(), b->inertia, b->nprob * sizeof(double), hipMemcpyDeviceToHost));
      for (uint32_t p = 0; p < b->nprob; ++p) {
        res[p].iterations = it[p];
        res[p].converged_at = cv[p];
        res[p].inertia = in[p];
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
  const BatchProblem& e = b->hp[problem];
  KMC_BCK(hipMemcpy(labels, b->labels + e.poff, (size_t)e.n * sizeof(int32_t),
                    hipMemcpyDeviceToHost));
  for (int i = 0; i < e.n; ++i) labels[i] -= e.coff;
  return KMC_OK;
}

kmc_status kmc_batch_get_centroids(kmc_batch* b, uint32_t problem, float* C) {
  if (!b || !C || problem >= b->nprob) return bfail(b, KMC_ERR_ARG, "kmc_batch_get_centroids");
  KMC_BCK(hipSetDevice(b->dev));
  const BatchProblem& e = b->hp[problem];
  const size_t d = b->d[problem];
  KMC_BCK(hipMemcpy2D(C, d * sizeof(float), b->C + (size_t)e.coff * b->DP,
                      (size_t)b->DP * sizeof(float), d * sizeof(float), (size_t)e.k,
                      hipMemcpyDeviceToHost));
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
