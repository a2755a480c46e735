// kmc_api.hip -- the C ABI (include/kmc.h): handle, buffers, Lloyd driver.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/kmc.h"
#include "kmc_copy.h"
#include "kmc_kernels.h"

using kmc::DevState;

struct kmc_handle {
  kmc_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  std::string err;

  size_t n = 0;
  int D = 0, K = 0, DP = 0, Kp = 0, prec = 0;
  int cus = 0;

  // points
  float* X = nullptr;
  uint16_t* Xhi = nullptr;
  uint16_t* Xlo = nullptr;
  uint8_t* Xq = nullptr;  // fp8 form: e4m3 plane and per-row E8M0 scales
  int* xsc = nullptr;
  float* xdn = nullptr;   // single-product forms: measured operand rounding
  float* cdn = nullptr;
  float hop = 0.0f;
  float* xn2 = nullptr;
  // centroids
  float* C = nullptr;
  float* Cnext = nullptr;
  float* Cw = nullptr;
  uint16_t* Chi = nullptr;
  uint16_t* Clo = nullptr;
  uint8_t* Cq = nullptr;
  float* cn = nullptr;
  float* CT = nullptr;
  // assignment
  int* labels = nullptr;
  int* q2 = nullptr;
  int* qa = nullptr;
  unsigned* q2cnt = nullptr;
  uint16_t* Phi = nullptr;  // bf16 hi / lo of -2c for the pruned all-k passes
  uint16_t* Plo = nullptr;  // (the bf16x3 form uses Chi / Clo)
  unsigned* qacnt = nullptr;
  long long qstride = 0;
  // sort / update
  int CH = 0, nblk = 0, lbits = 0;
  int* bcount = nullptr;
  int* count = nullptr;
  int* start = nullptr;
  int* members = nullptr;
  double* ipart = nullptr;
  double* tpart = nullptr;
  int nparts = 0;
  int update_mode = 0;
  // fixed-point streaming update
  double* xn64 = nullptr;
  long long* fpart = nullptr;
  // incremental fixed-point update: labels of the last update, exact totals
  int* prev = nullptr;
  long long* Stot = nullptr;
  long long* Qtot = nullptr;
  int* Ntot = nullptr;
  long long* qpart = nullptr;
  int* cpart = nullptr;
  // the moved points' changes to the totals, kDeltaCopies copies
  long long* Sdel = nullptr;
  long long* Qdel = nullptr;
  int* Ndel = nullptr;
  int fx_ds = 0, fx_nslice = 0, fx_wgps = 0;
  float fx_scale = 1.0f;
  int fx_magic = 0;  // max |x| * fx_scale < 2^50 (k_update_fx's fast conversion)
  double fx_qscale = 1.0;
  int* conv = nullptr;
  double* shift2 = nullptr;
  int* alive = nullptr;
  DevState* state = nullptr;

  // launch configuration
  int kchunk = 0;
  size_t assign_lds = 0;
  int assign_grid = 0;
  int variant = 0;  // k_assign launch variant (kmc_autotune)
  int vgrid[KMC_MAX_VARIANTS] = {0};
  long long vqstride[KMC_MAX_VARIANTS] = {0};
  int max_regions = 0;
  float alpha = 0, beta = 0, gref = 0, xpk = 0;
  float alpha3 = 0, beta3 = 0, gref3 = 0;  // bf16x3 constants (pruned all-k pass)
  int prune_min = kmc::kPruneMinPoints;
  bool short_prune = true;  // shorter all-k queues take the single-sweep pruned pass
  // launch the pruned all-k pass in this run: pruning is on and the last run
  // ended with an all-k queue within a quarter of the threshold (or none
  // ran yet); otherwise the exact pass takes every queue (same labels)
  bool prune_launch() const {
    return prune_min > 0 && (last_all_k < 0 || last_all_k >= prune_min / 4);
  }
  int idx_bits = 0;

  bool have_points = false, have_centroids = false;
  // per timed iteration kEvPerIter events: distance stage start and end,
  // refinement end, update kernel start and end, iteration end
  std::vector<hipEvent_t> events;
  uint64_t timed_iters = 0;  // iterations of the last run with stage events
  kmc_setup_times setup{};

  // captured iteration chunks (KMC_RUN_GRAPH), reused by later runs
  struct GraphEntry {
    uint64_t iters, ev_off;
    int mode, rules, variant, timed, prune;
    hipGraphExec_t exec;
  };
  // all-k queue length of the last assignment of the last run (-1: none yet)
  long long last_all_k = -1;
  std::vector<GraphEntry> graphs;
};

namespace kmc {

// gamma_m = m u / (1 - m u), the classical bound for m rounded operations.
double gam(double m, double u) { return m * u / (1.0 - m * u); }

// Margin-band constants. The MFMA value of centroid j is
//   b_j = fl( fl(||c_j||^2 + ||x||^2) + sum_k x_k (-2 c_jk) ),
// the full squared distance d_j with its centroid index packed into the low
// mantissa bits. Its error splits into a part common to all centroids of the
// point (the rounding of ||x||^2), which cancels in every margin, and a part
// eta_j with
//   |eta_j| <= (u + gA)(||c||^2 + ||x||^2 + 2||x||c) + gD c^2
//              + 2 (t_in + gA (1 + t_in)) ||x|| c          (operands, products)
// where c = cmax, u = 2^-24, gA = accumulation bound, gD = gamma_D, t_in =
// operand rounding of the bf16 and fp8 forms. With ||x||^2 <= 2 d_j + 2 c^2:
//   |eta_j| <= alpha ||x|| c + beta c^2 + g1 d_j
//   alpha = 2(u + gA) + 2(t_in + gA(1 + t_in)),  beta = 3(u + gA) + gD,
//   g1 = 2(u + gA).
// The reference's direct float form r_j satisfies |r_j - d_j| <= g_{D+3} d_j,
// and the packed index moves b_j by < 2^(idx_bits-23) b_j. All terms relative
// to the distance go into gref; the band test evaluates them at the upper
// bound b + E of each distance. Every constant carries a factor 1.25 of
// headroom.
void band_constants(int prec, int D, int idx_bits, float* alpha, float* beta, float* gref,
                    float* xpk, bool measured, float* hop) {
  const double u = std::ldexp(1.0, -24);   // fp32 unit roundoff
  const double ub = std::ldexp(1.0, -8);   // bf16 unit roundoff (8-bit significand)
  const double u8 = std::ldexp(1.0, -4);   // e4m3 unit roundoff (4-bit significand)
  double t_in, gA;
  if (prec == kmc::PREC_FP32) {
    t_in = 0.0;             // operands exact
    gA = gam(D + 2, u);     // product rounding + D-term accumulation
  } else if (prec == kmc::PREC_BF16X3) {
    // x = hi + lo + r with |r| <= ub^2 |x| (likewise -2c); the dropped
    // lo*lo and residual products sum to at most 3.2 ub^2 |x||c|
    t_in = 3.2 * ub * ub;
    gA = gam(3 * D + 1, u); // products exact in fp32, 3D-term accumulation
  } else if (prec == kmc::PREC_FP8) {
    // each operand row is scaled by 2^-e so that its largest entry lands in
    // [224, 448]: normal entries round with relative error u8, subnormal
    // ones with absolute error 2^-10 * 2^e <= eta * max|row| <= eta * ||row||
    const double eta = std::ldexp(1.0, -10) / 224.0;
    t_in = 2.0 * u8 + u8 * u8 + 2.0 * (1.0 + u8) * std::sqrt((double)D) * eta +
           (double)D * eta * eta;
    // the scaled fp8 MFMA does not accumulate like sequential fp32: its error
    // measured on random e4m3 operands (wide exponent spread, 64-term dot
    // products) stays below 2^-10.7 of the largest product; 2^-8 of the
    // absolute sum is taken
    gA = std::ldexp(1.0, -8);
  } else {
    t_in = 2.0 * ub + ub * ub;
    gA = gam(D + 1, u);
  }
  const double gD = gam(D, u);
  const double pk = std::ldexp(1.0, idx_bits - 23);
  // measured: the single-product forms know each stored operand's rounding
  // error norm; the distance kernel adds ||dc|| ||x|| + ||c'|| ||dx|| +
  // ||dc|| ||dx|| (c' = -2c, Cauchy-Schwarz) in place of 2 t_in ||x|| c
  const bool meas = measured && (prec == kmc::PREC_BF16 || prec == kmc::PREC_FP8);
  const double a = 2.0 * (u + gA) + 2.0 * ((meas ? 0.0 : t_in) + gA * (1.0 + t_in));
  const double b = 3.0 * (u + gA) + gD;
  const double g = gam(D + 3, u) + pk + 2.0 * (u + gA);
  const double H = 1.25;  // headroom over the bound
  *alpha = (float)(H * a);
  *beta = (float)(H * b);
  *gref = (float)(H * g);
  // single-product bf16 and fp8 rank ||c||^2 - 2 x.c and add ||x||^2
  // afterwards: packing perturbation and the final add are then relative to
  // d^2 + ||x||^2
  *xpk = (prec == kmc::PREC_BF16 || prec == kmc::PREC_FP8) ? (float)(H * (pk + 2.0 * u)) : 0.0f;
  if (hop) *hop = meas ? (float)H : 0.0f;
}

}  // namespace kmc

namespace {

constexpr int kEvPerIter = 6;

const char* kStatusText[] = {"ok",
                             "invalid or missing argument",
                             "no HIP device available",
                             "index out of bounds",
                             "invalid input file",
                             "HIP runtime error",
                             "out of device memory",
                             "call out of order",
                             "unsupported configuration"};

kmc_status fail(kmc_handle* h, kmc_status s, const std::string& msg) {
  if (h) h->err = msg;
  return s;
}

#define KMC_CK(expr)                                                                \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(h, e_ == hipErrorOutOfMemory ? KMC_ERR_OOM : KMC_ERR_HIP,         \
                  std::string(#expr) + ": " + hipGetErrorString(e_));               \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

int pick_dp(int d) {
  const int opts[] = {16, 32, 64, 128, 256};
  for (int o : opts)
    if (d <= o) return o;
  return 0;
}


kmc_status alloc_all(kmc_handle* h) {
  const size_t n = h->n;
  const size_t DP = h->DP, K = h->K, Kp = h->Kp;
  KMC_CK(dalloc(&h->X, n * DP));
  if (h->prec == kmc::PREC_BF16 || h->prec == kmc::PREC_BF16X3)
    KMC_CK(dalloc(&h->Xhi, n * DP));
  if (h->prec == kmc::PREC_BF16X3) KMC_CK(dalloc(&h->Xlo, n * DP));
  if (h->prec == kmc::PREC_FP8) {
    KMC_CK(dalloc(&h->Xq, n * DP));
    KMC_CK(dalloc(&h->xsc, n));
  }
  if (h->prec == kmc::PREC_FP8 || h->prec == kmc::PREC_BF16) {
    KMC_CK(dalloc(&h->xdn, n));
    KMC_CK(dalloc(&h->cdn, Kp));
  }
  KMC_CK(dalloc(&h->xn2, n));
  KMC_CK(dalloc(&h->C, K * DP));
  KMC_CK(dalloc(&h->Cnext, K * DP));
  if (h->prec == kmc::PREC_FP32) KMC_CK(dalloc(&h->Cw, Kp * DP));
  else if (h->prec == kmc::PREC_FP8) KMC_CK(dalloc(&h->Cq, Kp * (DP + kmc::kFp8RowPad)));
  else KMC_CK(dalloc(&h->Chi, Kp * DP));
  if (h->prec == kmc::PREC_BF16X3) KMC_CK(dalloc(&h->Clo, Kp * DP));
  KMC_CK(dalloc(&h->cn, Kp));
  if (h->prec != kmc::PREC_BF16X3) {
    KMC_CK(dalloc(&h->Phi, Kp * DP));
    KMC_CK(dalloc(&h->Plo, Kp * DP));
  }
  KMC_CK(dalloc(&h->CT, Kp * DP));
  KMC_CK(dalloc(&h->labels, n));
  size_t qtot = 0;
  for (int v = 0; v < kmc::assign_variants(); ++v)
    qtot = std::max(qtot, (size_t)h->vgrid[v] * (size_t)h->vqstride[v]);
  KMC_CK(dalloc(&h->q2, 4 * qtot));
  KMC_CK(dalloc(&h->qa, qtot));
  KMC_CK(dalloc(&h->q2cnt, (size_t)h->max_regions));
  KMC_CK(dalloc(&h->qacnt, (size_t)h->max_regions));
  KMC_CK(dalloc(&h->bcount, K * (size_t)h->nblk));
  KMC_CK(dalloc(&h->count, K));
  KMC_CK(dalloc(&h->start, K));
  KMC_CK(dalloc(&h->members, n));
  KMC_CK(dalloc(&h->ipart, (size_t)h->nparts));
  if (h->update_mode == kmc::UPDATE_TREE)
    KMC_CK(dalloc(&h->tpart, ((n + kmc::kTreeRows - 1) / kmc::kTreeRows + K) * DP));
  if (h->update_mode == kmc::UPDATE_FIXED) {
    const size_t Kp = (size_t)h->Kp, ws = (size_t)h->fx_wgps;
    KMC_CK(dalloc(&h->xn64, n));
    KMC_CK(dalloc(&h->fpart, (size_t)h->fx_nslice * ws * Kp * (size_t)h->fx_ds));
    KMC_CK(dalloc(&h->prev, n));
    KMC_CK(dalloc(&h->Stot, Kp * DP));
    KMC_CK(dalloc(&h->Qtot, Kp));
    KMC_CK(dalloc(&h->Ntot, Kp));
    KMC_CK(dalloc(&h->qpart, ws * Kp));
    KMC_CK(dalloc(&h->cpart, ws * Kp));
    const size_t nc = (size_t)kmc::kDeltaCopies;
    KMC_CK(dalloc(&h->Sdel, nc * Kp * DP));
    KMC_CK(dalloc(&h->Qdel, nc * Kp));
    KMC_CK(dalloc(&h->Ndel, nc * Kp));
    KMC_CK(hipMemsetAsync(h->Sdel, 0, nc * Kp * DP * sizeof(long long), h->stream));
    KMC_CK(hipMemsetAsync(h->Qdel, 0, nc * Kp * sizeof(long long), h->stream));
    KMC_CK(hipMemsetAsync(h->Ndel, 0, nc * Kp * sizeof(int), h->stream));
  }
  KMC_CK(dalloc(&h->conv, K));
  KMC_CK(dalloc(&h->shift2, K));
  KMC_CK(dalloc(&h->alive, K));
  KMC_CK(dalloc(&h->state, 1));
  KMC_CK(kmc::launch_fill_i32(h->alive, (long long)K, 1, h->stream));
  KMC_CK(hipMemsetAsync(h->C, 0, K * DP * sizeof(float), h->stream));
  KMC_CK(hipMemsetAsync(h->Cnext, 0, K * DP * sizeof(float), h->stream));
  KMC_CK(hipMemsetAsync(h->state, 0, sizeof(DevState), h->stream));
  KMC_CK(hipMemsetAsync(h->q2cnt, 0, h->max_regions * sizeof(unsigned), h->stream));
  KMC_CK(hipMemsetAsync(h->qacnt, 0, h->max_regions * sizeof(unsigned), h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  return KMC_OK;
}

void free_all(kmc_handle* h) {
  void* ptrs[] = {h->X, h->Xhi, h->Xlo, h->xn2, h->C, h->Cnext, h->Cw, h->Chi,
                  h->Clo, h->cn, h->CT, h->labels, h->q2, h->qa, h->q2cnt, h->qacnt,
                  h->bcount,
                  h->count, h->start, h->members, h->ipart, h->tpart, h->conv,
                  h->shift2, h->alive, h->state, h->xn64, h->fpart, h->qpart, h->cpart,
                  h->prev, h->Stot, h->Qtot, h->Ntot,
                  h->Xq, h->xsc, h->Cq, h->xdn, h->cdn, h->Phi, h->Plo,
                  h->Sdel, h->Qdel, h->Ndel};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

// Device stop flag lives inside the state block.
const int* done_ptr(kmc_handle* h) { return &h->state->done; }

kmc_status reset_state(kmc_handle* h, uint64_t maxiter) {
  DevState s{};
  // keep sum_x2
  KMC_CK(hipMemcpyAsync(&s.sum_x2, &h->state->sum_x2, sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  s.iter = 0;
  s.maxiter = maxiter;
  s.converged_at = maxiter;
  s.done = 0;
  // s lives on this stack frame: the copy must be done before it returns
  KMC_CK(hipMemcpyAsync(h->state, &s, sizeof(DevState), hipMemcpyHostToDevice, h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  return KMC_OK;
}

// Record an event on the handle's stream. Under stream capture an event
// recorded with hipEventRecord comes back unusable, so the record node is
// added to the capturing graph explicitly and made the stream's dependency.
hipError_t record_event(hipStream_t s, hipEvent_t e) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t err = hipStreamIsCapturing(s, &cs);
  if (err != hipSuccess) return err;
  if (cs != hipStreamCaptureStatusActive) return hipEventRecord(e, s);
  unsigned long long id = 0;
  hipGraph_t g = nullptr;
  const hipGraphNode_t* deps = nullptr;
  size_t nd = 0;
  err = hipStreamGetCaptureInfo_v2(s, &cs, &id, &g, &deps, &nd);
  if (err != hipSuccess) return err;
  hipGraphNode_t node;
  err = hipGraphAddEventRecordNode(&node, g, deps, nd, e);
  if (err != hipSuccess) return err;
  return hipStreamUpdateCaptureDependencies(s, &node, 1, hipStreamSetCaptureDependencies);
}

kmc::PrepParams prep_params(kmc_handle* h) {
  kmc::PrepParams pp{};
  pp.C = h->C;
  pp.Cw = h->Cw;
  pp.Chi = h->Chi;
  pp.Clo = h->Clo;
  pp.Cq = h->Cq;
  pp.cdn = h->cdn;
  pp.cn = h->cn;
  pp.Phi = h->Phi;
  pp.Plo = h->Plo;
  pp.CT = h->CT;
  pp.alive = h->alive;
  pp.done = done_ptr(h);
  pp.K = h->K;
  pp.Kp = h->Kp;
  pp.DP = h->DP;
  pp.prec = h->prec;
  return pp;
}

// prep: form the distance operands from C first (the first iteration of a
// run; later iterations get them from the previous iteration's k_finish).
// e0/e1 bracket the distance stage, r1 marks the end of the refinement.
kmc_status enqueue_assign(kmc_handle* h, hipEvent_t e0, hipEvent_t e1, bool prep = true,
                          hipEvent_t r1 = nullptr) {
  if (e0) KMC_CK(record_event(h->stream, e0));
  if (prep) KMC_CK(kmc::launch_prep(prep_params(h), h->stream));

  kmc::AssignParams ap{};
  ap.X = h->X;
  ap.Xhi = h->Xhi;
  ap.Xq = h->Xq;
  ap.xsc = h->xsc;
  ap.Cq = h->Cq;
  ap.xdn = h->xdn;
  ap.cdn = h->cdn;
  ap.hop = h->hop;
  ap.Xlo = h->Xlo;
  ap.xn2 = h->xn2;
  ap.n = (long long)h->n;
  ap.Cw = h->Cw;
  ap.Chi = h->Chi;
  ap.Clo = h->Clo;
  ap.cn = h->cn;
  ap.K = h->K;
  ap.Kp = h->Kp;
  ap.kchunk = h->kchunk;
  ap.idx_mask = (1u << h->idx_bits) - 1u;
  ap.alpha = h->alpha;
  ap.beta = h->beta;
  ap.gref = h->gref;
  ap.xpk = h->xpk;
  ap.labels = h->labels;
  ap.q2 = h->q2;
  ap.qa = h->qa;
  ap.q2cnt = h->q2cnt;
  ap.qacnt = h->qacnt;
  ap.qstride = h->qstride;
  ap.done = done_ptr(h);
  KMC_CK(kmc::launch_assign(h->DP, h->prec, h->variant, ap, h->assign_grid, h->assign_lds,
                            h->stream));
  if (e1) KMC_CK(record_event(h->stream, e1));

  kmc::RefineParams rp{};
  rp.X = h->X;
  rp.C = h->C;
  rp.CT = h->CT;
  rp.q2 = h->q2;
  rp.qa = h->qa;
  rp.q2cnt = h->q2cnt;
  rp.qacnt = h->qacnt;
  rp.qstride = h->qstride;
  rp.nregions = h->assign_grid;
  rp.alive = h->alive;
  rp.labels = h->labels;
  rp.done = done_ptr(h);
  rp.K = h->K;
  rp.Kp = h->Kp;
  rp.D = h->D;
  rp.DP = h->DP;
  rp.cn = h->cn;
  rp.xn2 = h->xn2;
  rp.Phi = h->prec == kmc::PREC_BF16X3 ? h->Chi : h->Phi;
  rp.Plo = h->prec == kmc::PREC_BF16X3 ? h->Clo : h->Plo;
  rp.alpha3 = h->alpha3;
  rp.beta3 = h->beta3;
  rp.gref3 = h->gref3;
  rp.pq = h->short_prune ? 1 : 0;
  // with the fused shape the single-sweep groups take all-k queues of any
  // length; k_refine_prune serves the other shapes, or with the fused form off
  rp.prune_min = (rp.pq && kmc::pq_shape(h->K, h->DP)) ? 0
                 : h->prune_launch()                  ? h->prune_min
                                                      : 0;
  KMC_CK(kmc::launch_refine(rp, h->stream));
  if (r1) KMC_CK(record_event(h->stream, r1));
  return KMC_OK;
}

// u0/u1: events around the centroid-update kernel (timed iterations); f1:
// after the finishing kernels (end of the iteration)
kmc_status enqueue_update(kmc_handle* h, int mode, int rules, hipEvent_t u0 = nullptr,
                          hipEvent_t u1 = nullptr, hipEvent_t f1 = nullptr) {
  kmc::SortParams sp{};
  sp.labels = h->labels;
  sp.bcount = h->bcount;
  sp.count = h->count;
  sp.start = h->start;
  sp.members = h->members;
  sp.done = done_ptr(h);
  sp.n = (long long)h->n;
  sp.K = h->K;
  sp.CH = h->CH;
  sp.nblk = h->nblk;
  sp.lbits = h->lbits;
  const bool fixed = h->update_mode == kmc::UPDATE_FIXED;
  if (!fixed) KMC_CK(kmc::launch_sort(sp, h->stream));

  kmc::UpdateParams up{};
  up.X = h->X;
  up.C = h->C;
  up.Cnext = h->Cnext;
  up.labels = h->labels;
  up.members = h->members;
  up.start = h->start;
  up.count = h->count;
  up.ipart = h->ipart;
  up.tpart = h->tpart;
  up.done = done_ptr(h);
  up.n = (long long)h->n;
  up.K = h->K;
  up.D = h->D;
  up.DP = h->DP;
  up.xn64 = h->xn64;
  up.fpart = h->fpart;
  up.prev = h->prev;
  up.Stot = h->Stot;
  up.Qtot = h->Qtot;
  up.Ntot = h->Ntot;
  up.state = h->state;
  up.qpart = h->qpart;
  up.cpart = h->cpart;
  up.fscale = h->fx_scale;
  up.magic = h->fx_magic;
  up.qscale = h->fx_qscale;
  up.Kp = h->Kp;
  up.nslice = h->fx_nslice;
  up.wgps = h->fx_wgps;
  up.ds = h->fx_ds;
  up.Sdel = h->Sdel;
  up.Qdel = h->Qdel;
  up.Ndel = h->Ndel;
  if (u0) KMC_CK(record_event(h->stream, u0));
  KMC_CK(kmc::launch_update(up, h->update_mode, h->stream));
  if (u1) KMC_CK(record_event(h->stream, u1));

  kmc::FinishParams fp{};
  fp.C = h->C;
  fp.Cnext = h->Cnext;
  fp.tpart = h->tpart;
  fp.start = h->start;
  fp.tree = h->update_mode == kmc::UPDATE_TREE ? 1 : 0;
  fp.conv = h->conv;
  fp.shift2 = h->shift2;
  fp.count = h->count;
  fp.alive = h->alive;
  fp.done = done_ptr(h);
  fp.tol = h->cfg.tol;
  fp.K = h->K;
  fp.D = h->D;
  fp.DP = h->DP;
  fp.exclude_empty = (rules & kmc::RULE_EXCLUDE_EMPTY) ? 1 : 0;
  fp.keep_empty = (rules & kmc::RULE_KEEP_EMPTY) ? 1 : 0;
  fp.do_prep = 1;
  fp.prep = prep_params(h);
  fp.fixed = fixed ? 1 : 0;
  fp.fpart = h->fpart;
  fp.Stot = h->Stot;
  fp.Qtot = h->Qtot;
  fp.Ntot = h->Ntot;
  fp.n = (long long)h->n;
  fp.qpart = h->qpart;
  fp.cpart = h->cpart;
  fp.count_out = h->count;
  fp.ipart_out = h->ipart;
  fp.Kp = h->Kp;
  fp.nslice = h->fx_nslice;
  fp.wgps = h->fx_wgps;
  fp.ds = h->fx_ds;
  fp.inv_fscale = 1.0 / (double)h->fx_scale;
  fp.inv_qscale = 1.0 / h->fx_qscale;
  fp.Sdel = h->Sdel;
  fp.Qdel = h->Qdel;
  fp.Ndel = h->Ndel;
  fp.labels = h->labels;
  fp.prev = h->prev;
  kmc::StatsParams& st = fp.stats;
  st.ipart = h->ipart;
  st.nparts = h->nparts;
  st.shift2 = h->shift2;
  st.conv = h->conv;
  st.q2cnt = h->q2cnt;
  st.qacnt = h->qacnt;
  st.nregions = h->assign_grid;
  st.state = h->state;
  st.done = done_ptr(h);
  st.tol = h->cfg.tol;
  st.K = h->K;
  st.mode = mode;
  st.cost_ratio = (rules & kmc::RULE_COST_RATIO) ? 1 : 0;
  st.incremental = h->Stot ? 1 : 0;
  st.n = (long long)h->n;
  st.fx_modes = fixed ? 1 : 0;
  KMC_CK(kmc::launch_finish(fp, h->stream));
  if (f1) KMC_CK(record_event(h->stream, f1));
  return KMC_OK;
}

// m iterations starting at global iteration `first`; events around the
// distance kernel and the update kernel for iterations below ntimed.
kmc_status enqueue_chunk(kmc_handle* h, uint64_t first, uint64_t m, int mode, int rules,
                         uint64_t ntimed) {
  for (uint64_t it = 0; it < m; ++it) {
    const uint64_t gi = first + it;
    const bool tm = gi < ntimed;
    hipEvent_t* ev = tm ? &h->events[kEvPerIter * gi] : nullptr;
    kmc_status st = enqueue_assign(h, tm ? ev[0] : nullptr, tm ? ev[1] : nullptr, it == 0,
                                   tm ? ev[2] : nullptr);
    if (st == KMC_OK)
      st = enqueue_update(h, mode, rules, tm ? ev[3] : nullptr, tm ? ev[4] : nullptr,
                          tm ? ev[5] : nullptr);
    if (st != KMC_OK) return st;
  }
  return KMC_OK;
}

// Launch m iterations as one graph; the graph is captured on first use and
// kept for later runs with the same shape. Returns KMC_ERR_UNSUPPORTED when
// the stream cannot be captured (the caller then launches eagerly).
kmc_status launch_chunk_graph(kmc_handle* h, uint64_t first, uint64_t m, int mode, int rules,
                              uint64_t ntimed) {
  const int timed = first < ntimed ? 1 : 0;
  for (auto& g : h->graphs) {
    if (g.iters == m && g.ev_off == (timed ? first : 0) && g.mode == mode &&
        g.rules == rules && g.variant == h->variant && g.timed == timed &&
        g.prune == (h->prune_launch() ? 1 : 0)) {
      KMC_CK(hipGraphLaunch(g.exec, h->stream));
      return KMC_OK;
    }
  }
  if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return KMC_ERR_UNSUPPORTED;
  kmc_status st = enqueue_chunk(h, first, m, mode, rules, timed ? ntimed : 0);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->stream, &graph);
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
  h->graphs.push_back({m, timed ? first : 0, mode, rules, h->variant, timed,
                       h->prune_launch() ? 1 : 0, exec});
  KMC_CK(hipGraphLaunch(exec, h->stream));
  return KMC_OK;
}

kmc_status read_state(kmc_handle* h, DevState* s) {
  KMC_CK(hipMemcpyAsync(s, h->state, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  return KMC_OK;
}

void fill_result(const kmc_handle* h, const DevState& s, kmc_result* r) {
  if (!r) return;
  r->iterations = s.iter;
  r->converged_at = s.converged_at;
  r->inertia = s.inertia;
  r->score = s.inertia - s.sum_x2;
  r->max_shift = std::sqrt(s.max_shift2);
  r->refined_two = s.refined_two;
  r->refined_full = s.refined_all;
  r->full_passes = s.full_passes;
  r->moved_points = s.moved_total;
  (void)h;
}

// Milliseconds between two events recorded on the handle's stream around a
// synchronous set-up step (0 when timing fails: timing never fails a call).
struct StepTimer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit StepTimer(hipStream_t st) : s(st) {
    if (hipEventCreate(&a) != hipSuccess) a = nullptr;
    if (hipEventCreate(&b) != hipSuccess) b = nullptr;
    if (a) (void)hipEventRecord(a, s);
  }
  double stop() {
    float ms = 0.f;
    if (a && b && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess &&
        hipEventElapsedTime(&ms, a, b) == hipSuccess)
      return ms;
    return 0.0;
  }
  ~StepTimer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

kmc_status load_points(kmc_handle* h, const float* dsrc, size_t ld) {
  StepTimer tm(h->stream);
  const bool keep64 = h->update_mode == kmc::UPDATE_FIXED;
  double* xn64 = keep64 ? h->xn64 : nullptr;
  if (!keep64) KMC_CK(dalloc(&xn64, h->n + 256));
  kmc::LoadParams lp{};
  lp.src = dsrc;
  lp.ld = ld;
  lp.n = (long long)h->n;
  lp.D = h->D;
  lp.DP = h->DP;
  lp.X = h->X;
  lp.Xhi = h->Xhi;
  lp.Xq = h->Xq;
  lp.xsc = h->xsc;
  lp.xdn = h->xdn;
  lp.Xlo = h->Xlo;
  lp.xn2 = h->xn2;
  lp.xn2_64 = xn64;
  hipError_t e = kmc::launch_load_points(lp, h->stream);
  double* part = nullptr;
  if (e == hipSuccess) e = dalloc(&part, 257 + 2);
  if (e == hipSuccess)
    e = kmc::launch_sum_f64(xn64, (long long)h->n, &h->state->sum_x2, part, h->stream);
  if (e == hipSuccess && keep64) {
    // fixed-point scales: powers of two that keep n * max|x| * scale and
    // n * max||x||^2 * qscale below 2^61
    unsigned* amax = reinterpret_cast<unsigned*>(part + 257);
    unsigned long long* qmax = reinterpret_cast<unsigned long long*>(part + 258);
    e = hipMemsetAsync(part + 257, 0, 2 * sizeof(double), h->stream);
    if (e == hipSuccess)
      e = kmc::launch_absmax(h->X, (long long)h->n * h->DP, xn64, (long long)h->n, amax, qmax,
                             h->stream);
    unsigned long long hv[2] = {0, 0};
    if (e == hipSuccess)
      e = hipMemcpyAsync(hv, part + 257, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) {
      float am;
      const unsigned ab = (unsigned)hv[0];
      std::memcpy(&am, &ab, sizeof(am));
      double qm;
      std::memcpy(&qm, &hv[1], sizeof(qm));
      const double nn = (double)std::max<size_t>(h->n, 1);
      const int s1 = kmc::fx_scale_exp(h->n, am);
      const int s2 = std::max(-900, std::min(900, (int)std::floor(
                                                      61.0 - std::log2(nn * std::max(qm, 1e-300)))));
      h->fx_scale = std::ldexp(1.0f, s1);
      h->fx_qscale = std::ldexp(1.0, s2);
      h->fx_magic = std::ldexp((double)am, s1) < std::ldexp(1.0, 50) ? 1 : 0;
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (part) (void)hipFree(part);
  if (!keep64) (void)hipFree(xn64);
  KMC_CK(e);
  h->setup.operands_ms = tm.stop();
  h->have_points = true;
  return KMC_OK;
}

}  // namespace

extern "C" {

const char* kmc_status_string(kmc_status s) {
  const int i = (int)s;
  if (i < 0 || i >= (int)(sizeof(kStatusText) / sizeof(kStatusText[0]))) return "unknown status";
  return kStatusText[i];
}

const char* kmc_last_error(const kmc_handle* h) { return h ? h->err.c_str() : ""; }

int kmc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

kmc_status kmc_memory_footprint(int device, size_t* used_bytes) {
  if (!used_bytes) return KMC_ERR_ARG;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return KMC_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return KMC_ERR_ARG;
  size_t fr = 0, tot = 0;
  const hipError_t e = hipMemGetInfo(&fr, &tot);
  (void)hipSetDevice(cur);
  if (e != hipSuccess) return KMC_ERR_HIP;
  *used_bytes = tot - fr;
  return KMC_OK;
}

kmc_status kmc_create(const kmc_config* cfg, kmc_handle** out) {
  if (!cfg || !out) return KMC_ERR_ARG;
  *out = nullptr;
  if (cfg->n == 0 || cfg->d == 0 || cfg->k == 0) return KMC_ERR_ARG;
  if (cfg->k > cfg->n) return KMC_ERR_ARG;
  if (cfg->n > (size_t)0x7fffffff) return KMC_ERR_UNSUPPORTED;
  if (cfg->k > 16384u) return KMC_ERR_UNSUPPORTED;
  if (cfg->precision < 0 || cfg->precision > 3) return KMC_ERR_ARG;
  if (cfg->update != KMC_UPDATE_ORDERED && cfg->update != KMC_UPDATE_TREE &&
      cfg->update != KMC_UPDATE_FIXED)
    return KMC_ERR_ARG;
  if (!(cfg->tol >= 0.0f)) return KMC_ERR_ARG;
  // one fp8 MFMA step covers 64 dimensions
  const int DP = cfg->precision == KMC_PREC_FP8 && pick_dp((int)cfg->d) > 0
                     ? std::max(64, pick_dp((int)cfg->d))
                     : pick_dp((int)cfg->d);
  if (DP == 0) return KMC_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return KMC_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev) return KMC_ERR_ARG;

  kmc_handle* h = new (std::nothrow) kmc_handle();
  if (!h) return KMC_ERR_OOM;
  h->cfg = *cfg;
  h->dev = cfg->device;
  h->n = cfg->n;
  h->D = (int)cfg->d;
  h->K = (int)cfg->k;
  h->DP = DP;
  h->Kp = (h->K + 63) / 64 * 64;  // centroid blocks are processed in pairs
  h->prec = cfg->precision;
  h->update_mode = cfg->update == KMC_UPDATE_TREE    ? kmc::UPDATE_TREE
                   : cfg->update == KMC_UPDATE_FIXED ? kmc::UPDATE_FIXED
                                                     : kmc::UPDATE_ORDERED;
  kmc_status st = KMC_OK;
  do {
    if (hipSetDevice(h->dev) != hipSuccess) { st = fail(h, KMC_ERR_HIP, "hipSetDevice"); break; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->dev) != hipSuccess) {
      st = fail(h, KMC_ERR_HIP, "hipGetDeviceProperties");
      break;
    }
    h->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
      st = fail(h, KMC_ERR_HIP, "hipStreamCreate");
      break;
    }
    // sort blocking: ~4 points per cluster counter per block, 1024..16384
    int ch = ((h->K + 255) / 256) * 256 * 4;
    h->CH = std::max(1024, std::min(16384, ch));
    h->nblk = (int)((h->n + h->CH - 1) / h->CH);
    h->lbits = 0;
    while ((1 << h->lbits) < h->K) ++h->lbits;
    if (h->update_mode == kmc::UPDATE_FIXED) {
      // slice width: the widest whose per-workgroup accumulators fit the
      // LDS; none (very large K): the tree update instead
      const size_t kFxLdsMax = 160 * 1024 - 256;  // the kernel's static LDS beside it
      h->fx_ds = 0;
      for (int ds : {64, 32, 16})
        if (ds <= DP && kmc::update_fx_lds(h->Kp, ds) <= kFxLdsMax) {
          h->fx_ds = ds;
          break;
        }
      if (h->fx_ds == 0) {
        h->update_mode = kmc::UPDATE_TREE;
      } else {
        h->fx_nslice = DP / h->fx_ds;
        h->fx_wgps = std::max(1, h->cus / h->fx_nslice);
      }
    }
    const long long S = (h->D + 63) / 64;
    h->nparts = h->update_mode == kmc::UPDATE_TREE
                    ? (int)(((long long)h->n + kmc::kTreeRows - 1) / kmc::kTreeRows)
                    : h->update_mode == kmc::UPDATE_FIXED ? h->K : (int)(h->K * S);

    // assign launch shape: all centroids resident in LDS when they fit
    const size_t kResidentMax = 144 * 1024, kChunkMax = 144 * 1024;
    h->kchunk = h->Kp;
    h->assign_lds = kmc::assign_lds(DP, h->prec, h->kchunk);
    if (h->assign_lds > kResidentMax) {
      int kc = 64;
      while (kc + 64 <= h->Kp && kmc::assign_lds(DP, h->prec, kc + 64) <= kChunkMax) kc += 64;
      h->kchunk = kc;
      h->assign_lds = kmc::assign_lds(DP, h->prec, kc);
    }
    bool ok = true;
    for (int v = 0; v < kmc::assign_variants() && ok; ++v) {
      const void* fn = kmc::assign_kernel_ptr(DP, h->prec, v);
      if (!fn) continue;  // variant not built for this (DP, precision)
      if (kmc::assign_variant_resident_only(v) && h->kchunk < h->Kp) continue;
      // a variant whose registers spill to scratch memory is not offered
      hipFuncAttributes fa{};
      if (hipFuncGetAttributes(&fa, fn) != hipSuccess) {
        st = fail(h, KMC_ERR_HIP, "hipFuncGetAttributes(assign)");
        ok = false;
        break;
      }
      if (fa.localSizeBytes > 0) continue;
      if (kmc::raise_lds_limit(fn, h->assign_lds) != hipSuccess) {
        st = fail(h, KMC_ERR_HIP, "hipFuncSetAttribute(assign)");
        ok = false;
        break;
      }
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kmc::assign_threads(v),
                                                       h->assign_lds) != hipSuccess ||
          per_cu < 1)
        per_cu = 1;
      const int tile = kmc::assign_tile(v);
      const long long tiles = (long long)((h->n + tile - 1) / tile);
      h->vgrid[v] = (int)std::min<long long>(
          {tiles, (long long)per_cu * h->cus, (long long)kmc::kMaxRegions});
      // each assign workgroup owns a queue region large enough for all its points
      h->vqstride[v] = (tiles + h->vgrid[v] - 1) / h->vgrid[v] * tile;
      h->max_regions = std::max(h->max_regions, h->vgrid[v]);
    }
    if (!ok) break;
    h->variant = 0;  // the first variant offered for this shape
    while (h->variant + 1 < kmc::assign_variants() && h->vgrid[h->variant] == 0) ++h->variant;
    if (h->vgrid[h->variant] == 0) {
      st = fail(h, KMC_ERR_UNSUPPORTED, "no distance-kernel variant for this shape");
      break;
    }
    h->assign_grid = h->vgrid[h->variant];
    h->qstride = h->vqstride[h->variant];
    h->idx_bits = 0;
    while ((1 << h->idx_bits) < h->Kp) ++h->idx_bits;
    kmc::band_constants(h->prec, h->D, h->idx_bits, &h->alpha, &h->beta, &h->gref,
                        &h->xpk, true, &h->hop);
    float xpk3 = 0.0f;
    kmc::band_constants(kmc::PREC_BF16X3, h->D, h->idx_bits, &h->alpha3, &h->beta3, &h->gref3,
                        &xpk3);
    st = alloc_all(h);
  } while (0);
  if (st != KMC_OK) {
    fprintf(stderr, "kmc_create: %s\n", h->err.c_str());
    free_all(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return st;
  }
  *out = h;
  return KMC_OK;
}

kmc_status kmc_destroy(kmc_handle* h) {
  if (!h) return KMC_ERR_ARG;
  (void)hipSetDevice(h->dev);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  for (hipEvent_t e : h->events) (void)hipEventDestroy(e);
  free_all(h);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return KMC_OK;
}

kmc_status kmc_set_points(kmc_handle* h, const float* X, size_t ld) {
  if (!h || !X || ld < (size_t)h->D) return fail(h, KMC_ERR_ARG, "kmc_set_points: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  const size_t bytes = ((h->n - 1) * ld + h->D) * sizeof(float);
  float* tmp = nullptr;
  KMC_CK(hipMalloc(reinterpret_cast<void**>(&tmp), bytes));
  StepTimer tm(h->stream);
  hipError_t e = hipMemcpyAsync(tmp, X, bytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) h->setup.memcpy_ms = tm.stop();
  kmc_status st = KMC_OK;
  if (e == hipSuccess) st = load_points(h, tmp, ld);
  else st = fail(h, KMC_ERR_HIP, std::string("H2D copy: ") + hipGetErrorString(e));
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  return st;
}

kmc_status kmc_set_points_device(kmc_handle* h, const float* dX, size_t ld) {
  if (!h || !dX || ld < (size_t)h->D)
    return fail(h, KMC_ERR_ARG, "kmc_set_points_device: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  h->setup.memcpy_ms = 0.0;
  return load_points(h, dX, ld);
}

kmc_status kmc_set_centroids(kmc_handle* h, const float* C0) {
  if (!h || !C0) return fail(h, KMC_ERR_ARG, "kmc_set_centroids: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  StepTimer tm(h->stream);
  KMC_CK(kmc::rows_to_device(h->C, (size_t)h->DP, C0, (size_t)h->D, (size_t)h->D,
                             (size_t)h->K, h->stream));
  KMC_CK(kmc::launch_fill_i32(h->alive, (long long)h->K, 1, h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  h->setup.init_ms = tm.stop();
  h->have_centroids = true;
  return KMC_OK;
}

// splitmix64 (Steele, Lea, Flood 2014; Vigna's constants): the uniforms of
// the k-means++ draws, u = top 53 bits / 2^53
static double splitmix_uniform(uint64_t& state) {
  uint64_t z = (state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

kmc_status kmc_init_plusplus(kmc_handle* h, uint64_t seed, int64_t* chosen) {
  return kmc_init_plusplus_trials(h, seed, 1, chosen);
}

kmc_status kmc_init_plusplus_trials(kmc_handle* h, uint64_t seed, uint32_t trials,
                                    int64_t* chosen) {
  if (!h) return KMC_ERR_ARG;
  if (!h->have_points) return fail(h, KMC_ERR_STATE, "kmc_init_plusplus: points not set");
  if (trials == 0) trials = 2u + (uint32_t)std::ceil(std::log((double)h->K));
  if (trials > 16) return fail(h, KMC_ERR_ARG, "kmc_init_plusplus_trials: at most 16 trials");
  KMC_CK(hipSetDevice(h->dev));
  StepTimer tm(h->stream);
  const size_t nu = 1 + (size_t)(h->K - 1) * trials;
  std::vector<double> uni(nu);
  uint64_t st = seed;
  for (size_t r = 0; r < nu; ++r) uni[r] = splitmix_uniform(st);
  const long long n = (long long)h->n;
  const size_t nb = (size_t)((n + 1023) / 1024);
  double *du = nullptr, *D2 = nullptr, *bsum = nullptr, *tsum = nullptr;
  long long *dch = nullptr, *cand = nullptr;
  hipError_t e = dalloc(&du, nu);
  if (e == hipSuccess) e = dalloc(&D2, (size_t)n);
  if (e == hipSuccess) e = dalloc(&bsum, nb);
  if (e == hipSuccess) e = dalloc(&dch, (size_t)h->K);
  if (e == hipSuccess && trials > 1) e = dalloc(&cand, (size_t)trials);
  if (e == hipSuccess && trials > 1) e = dalloc(&tsum, (size_t)trials * nb);
  if (e == hipSuccess)
    e = hipMemcpyAsync(du, uni.data(), nu * sizeof(double), hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = kmc::launch_plusplus(h->X, n, h->DP, h->C, h->K, du, D2, bsum, dch, h->stream,
                             (int)trials, cand, tsum);
  if (e == hipSuccess && chosen)
    e = hipMemcpyAsync(chosen, dch, h->K * sizeof(long long), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = kmc::launch_fill_i32(h->alive, (long long)h->K, 1, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(du);
  (void)hipFree(D2);
  (void)hipFree(bsum);
  (void)hipFree(dch);
  if (cand) (void)hipFree(cand);
  if (tsum) (void)hipFree(tsum);
  KMC_CK(e);
  h->setup.init_ms = tm.stop();
  h->have_centroids = true;
  return KMC_OK;
}

kmc_status kmc_set_labels(kmc_handle* h, const int32_t* labels0) {
  if (!h || !labels0) return fail(h, KMC_ERR_ARG, "kmc_set_labels: bad argument");
  if (!h->have_points) return fail(h, KMC_ERR_STATE, "kmc_set_labels: points not set");
  for (size_t i = 0; i < h->n; ++i)
    if (labels0[i] < 0 || labels0[i] >= h->K)
      return fail(h, KMC_ERR_OUT_OF_BOUNDS, "kmc_set_labels: label out of range");
  KMC_CK(hipSetDevice(h->dev));
  StepTimer tm(h->stream);
  KMC_CK(hipMemcpyAsync(h->labels, labels0, h->n * sizeof(int32_t), hipMemcpyHostToDevice,
                        h->stream));
  kmc_status st = reset_state(h, 1);
  if (st != KMC_OK) return st;
  KMC_CK(kmc::launch_fill_i32(h->alive, (long long)h->K, 1, h->stream));
  // C <- means of the label groups; clusters empty under these labels are
  // marked for KMC_RUN_EXCLUDE_EMPTY
  st = enqueue_update(h, kmc::CONV_NONE, 1);
  if (st != KMC_OK) return st;
  KMC_CK(hipStreamSynchronize(h->stream));
  h->setup.init_ms = tm.stop();
  h->have_centroids = true;
  return KMC_OK;
}

kmc_status kmc_run(kmc_handle* h, uint64_t maxiter, uint32_t flags, kmc_result* res) {
  if (!h) return KMC_ERR_ARG;
  if (!h->have_points || !h->have_centroids)
    return fail(h, KMC_ERR_STATE, "kmc_run: points and initial centroids must be set");
  if (res) std::memset(res, 0, sizeof(*res));
  KMC_CK(hipSetDevice(h->dev));
  kmc_status st = reset_state(h, maxiter);
  if (st != KMC_OK) return st;
  if (maxiter == 0) {
    DevState s;
    st = read_state(h, &s);
    if (st == KMC_OK) fill_result(h, s, res);
    return st;
  }
  const bool score_mode = (flags & KMC_RUN_CONVERGE_SCORE) != 0;
  const bool check = (flags & KMC_RUN_CHECK_CONVERGED) != 0;
  const bool timing = (flags & KMC_RUN_TIME_KERNELS) != 0;
  const bool exclude = (flags & KMC_RUN_EXCLUDE_EMPTY) != 0;
  const bool sum_mode = (flags & KMC_RUN_CONVERGE_SHIFT_SUM) != 0;
  const int mode = !check     ? kmc::CONV_NONE
                   : sum_mode ? kmc::CONV_SHIFT_SUM
                   : score_mode ? kmc::CONV_SCORE
                                : kmc::CONV_SHIFT;
  const int rules = (exclude ? kmc::RULE_EXCLUDE_EMPTY : 0) |
                    ((flags & KMC_RUN_KEEP_EMPTY) ? kmc::RULE_KEEP_EMPTY : 0) |
                    ((flags & KMC_RUN_COST_RATIO) ? kmc::RULE_COST_RATIO : 0);
  // the cost-ratio test needs the host to look at the stop flag as well
  const bool watch = check || (flags & KMC_RUN_COST_RATIO) != 0;
  if (!exclude) KMC_CK(kmc::launch_fill_i32(h->alive, (long long)h->K, 1, h->stream));
  const uint64_t kMaxTimed = 4096;
  const uint64_t ntimed = timing ? std::min<uint64_t>(maxiter, kMaxTimed) : 0;
  while (h->events.size() < kEvPerIter * ntimed) {
    hipEvent_t e;
    KMC_CK(hipEventCreate(&e));
    h->events.push_back(e);
  }
  hipEvent_t t0, t1;
  KMC_CK(hipEventCreate(&t0));
  KMC_CK(hipEventCreate(&t1));
  KMC_CK(hipEventRecord(t0, h->stream));
  uint64_t launched = 0, chunk = 1;
  DevState s{};
  bool use_graph = (flags & KMC_RUN_GRAPH) != 0;
  while (launched < maxiter) {
    const uint64_t m = watch ? std::min<uint64_t>(chunk, maxiter - launched) : maxiter - launched;
    if (use_graph) {
      st = launch_chunk_graph(h, launched, m, mode, rules, ntimed);
      if (st == KMC_ERR_UNSUPPORTED) {
        use_graph = false;  // capture refused: launch eagerly from here on
        st = KMC_OK;
      }
      if (st != KMC_OK) break;
    }
    if (!use_graph) st = enqueue_chunk(h, launched, m, mode, rules, ntimed);
    if (st != KMC_OK) break;
    launched += m;
    if (watch) {
      st = read_state(h, &s);
      if (st != KMC_OK) break;
      if (s.done) break;
      chunk = std::min<uint64_t>(chunk * 2, 16);
    }
  }
  if (st == KMC_OK) {
    hipError_t e = hipEventRecord(t1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (e != hipSuccess) st = fail(h, KMC_ERR_HIP, std::string("run timing: ") + hipGetErrorString(e));
    if (st == KMC_OK) st = read_state(h, &s);
    if (st == KMC_OK) h->last_all_k = (long long)s.refined_all;
    if (st == KMC_OK) h->timed_iters = std::min<uint64_t>(s.iter, ntimed);
    if (st == KMC_OK) {
      fill_result(h, s, res);
      if (res) {
        res->ms_total = ms;
        const uint64_t ran = std::min<uint64_t>(s.iter, ntimed);
        double acc = 0.0, accu = 0.0;
        for (uint64_t i = 0; i < ran; ++i) {
          float a = 0.f, u = 0.f;
          const hipEvent_t* ev = &h->events[kEvPerIter * i];
          if (hipEventElapsedTime(&a, ev[0], ev[1]) == hipSuccess) acc += a;
          if (hipEventElapsedTime(&u, ev[3], ev[4]) == hipSuccess) accu += u;
        }
        res->ms_assign = acc;
        res->assign_launches = ran;
        res->ms_update = accu;
        res->update_launches = ran;
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return st;
}

kmc_status kmc_set_assign_variant(kmc_handle* h, int variant) {
  if (!h || variant < 0 || variant >= kmc::assign_variants() || h->vgrid[variant] == 0)
    return fail(h, KMC_ERR_ARG, "kmc_set_assign_variant: no such variant for this problem");
  h->variant = variant;
  h->assign_grid = h->vgrid[variant];
  h->qstride = h->vqstride[variant];
  return KMC_OK;
}

kmc_status kmc_set_prune_min(kmc_handle* h, int32_t min_points) {
  if (!h || min_points < 0) return fail(h, KMC_ERR_ARG, "kmc_set_prune_min: bad argument");
  h->prune_min = min_points;
  h->last_all_k = -1;
  // captured graphs carry the old value in their refinement launches
  KMC_CK(hipSetDevice(h->dev));
  KMC_CK(hipStreamSynchronize(h->stream));
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
  return KMC_OK;
}

kmc_status kmc_set_short_prune(kmc_handle* h, int32_t enable) {
  if (!h) return fail(h, KMC_ERR_ARG, "kmc_set_short_prune: null handle");
  h->short_prune = enable != 0;
  // captured graphs carry the old choice in their refinement launches
  KMC_CK(hipSetDevice(h->dev));
  KMC_CK(hipStreamSynchronize(h->stream));
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
  return KMC_OK;
}

kmc_status kmc_autotune(kmc_handle* h, kmc_tuning* out) {
  if (!h) return KMC_ERR_ARG;
  if (!h->have_points || !h->have_centroids)
    return fail(h, KMC_ERR_STATE, "kmc_autotune: points and centroids must be set");
  KMC_CK(hipSetDevice(h->dev));
  kmc_status st = reset_state(h, 1);
  if (st != KMC_OK) return st;
  constexpr int kTimed = 3, kRounds = 2;
  hipEvent_t e[2 * kTimed];
  for (int i = 0; i < 2 * kTimed; ++i) KMC_CK(hipEventCreate(&e[i]));
  const int nv = std::min(kmc::assign_variants(), KMC_MAX_VARIANTS);
  kmc_tuning t{};
  t.nvariants = nv;
  for (int v = 0; v < nv; ++v) t.ms[v] = h->vgrid[v] == 0 ? -1.0f : 1e30f;
  // every variant is timed in each of kRounds rounds over the variants (so a
  // change of the clock during the tuning does not favour one variant); a
  // variant's time is its best round's mean of kTimed assignments after an
  // untimed one. A timed assignment is the distance kernel and the
  // refinement of its band points, since variants that keep fewer values per
  // lane leave more points to refine.
  for (int r = 0; r < kRounds && st == KMC_OK; ++r) {
    for (int v = 0; v < nv; ++v) {
      if (h->vgrid[v] == 0) continue;  // not built for this (DP, precision)
      st = kmc_set_assign_variant(h, v);
      if (st != KMC_OK) break;
      st = enqueue_assign(h, nullptr, nullptr);
      for (int i = 0; i < kTimed && st == KMC_OK; ++i)
        st = enqueue_assign(h, e[2 * i], nullptr, false, e[2 * i + 1]);
      if (st != KMC_OK) break;
      KMC_CK(hipEventSynchronize(e[2 * kTimed - 1]));
      float sum = 0.f;
      for (int i = 0; i < kTimed; ++i) {
        float a = 0.f;
        KMC_CK(hipEventElapsedTime(&a, e[2 * i], e[2 * i + 1]));
        sum += a;
      }
      t.ms[v] = std::min(t.ms[v], sum / kTimed);
      t.threads[v] = kmc::assign_threads(v);
      t.tile[v] = kmc::assign_tile(v);
    }
  }
  for (int i = 0; i < 2 * kTimed; ++i) (void)hipEventDestroy(e[i]);
  if (st != KMC_OK) return st;
  int best_v = 0;
  float best = 1e30f;
  for (int v = 0; v < nv; ++v)
    if (t.ms[v] >= 0.f && t.ms[v] < best) {
      best = t.ms[v];
      best_v = v;
    }
  t.chosen = best_v;
  st = kmc_set_assign_variant(h, best_v);
  if (out) *out = t;
  return st;
}

kmc_status kmc_get_labels(kmc_handle* h, int32_t* labels) {
  if (!h || !labels) return fail(h, KMC_ERR_ARG, "kmc_get_labels: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  KMC_CK(hipMemcpyAsync(labels, h->labels, h->n * sizeof(int32_t), hipMemcpyDeviceToHost,
                        h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  return KMC_OK;
}

kmc_status kmc_get_centroids(kmc_handle* h, float* C) {
  if (!h || !C) return fail(h, KMC_ERR_ARG, "kmc_get_centroids: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  KMC_CK(kmc::rows_to_host(C, (size_t)h->D, h->C, (size_t)h->DP, (size_t)h->D, (size_t)h->K,
                           h->stream));
  return KMC_OK;
}

kmc_status kmc_assign(kmc_handle* h, const float* C, int32_t* labels, kmc_result* res) {
  if (!h || !C || !labels) return fail(h, KMC_ERR_ARG, "kmc_assign: bad argument");
  if (!h->have_points) return fail(h, KMC_ERR_STATE, "kmc_assign: points not set");
  if (res) std::memset(res, 0, sizeof(*res));
  kmc_status st = kmc_set_centroids(h, C);
  if (st != KMC_OK) return st;
  st = reset_state(h, 1);
  if (st != KMC_OK) return st;
  st = enqueue_assign(h, nullptr, nullptr);
  if (st != KMC_OK) return st;
  std::vector<unsigned> q2(h->assign_grid), qa(h->assign_grid);  // regions of this variant
  KMC_CK(hipMemcpyAsync(q2.data(), h->q2cnt, q2.size() * sizeof(unsigned),
                        hipMemcpyDeviceToHost, h->stream));
  KMC_CK(hipMemcpyAsync(qa.data(), h->qacnt, qa.size() * sizeof(unsigned),
                        hipMemcpyDeviceToHost, h->stream));
  st = kmc_get_labels(h, labels);
  if (st != KMC_OK) return st;
  if (res) {
    for (unsigned v : q2) res->refined_two += v;
    for (unsigned v : qa) res->refined_full += v;
  }
  return KMC_OK;
}

kmc_status kmc_get_setup_times(const kmc_handle* h, kmc_setup_times* out) {
  if (!h || !out) return KMC_ERR_ARG;
  *out = h->setup;
  return KMC_OK;
}

kmc_status kmc_get_stage_times(kmc_handle* h, kmc_stage_times* out, uint64_t cap,
                               uint64_t* count) {
  if (!h || (!out && cap > 0)) return fail(h, KMC_ERR_ARG, "kmc_get_stage_times: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  const uint64_t m = std::min<uint64_t>(cap, h->timed_iters);
  for (uint64_t i = 0; i < m; ++i) {
    const hipEvent_t* ev = &h->events[kEvPerIter * i];
    float a = 0.f, r = 0.f, c = 0.f;
    KMC_CK(hipEventElapsedTime(&a, ev[0], ev[1]));
    KMC_CK(hipEventElapsedTime(&r, ev[1], ev[2]));
    KMC_CK(hipEventElapsedTime(&c, ev[2], ev[5]));
    out[i].distances_ms = a;
    out[i].argmin_ms = r;
    out[i].centroids_ms = c;
  }
  if (count) *count = h->timed_iters;
  return KMC_OK;
}

kmc_status kmc_update(kmc_handle* h, const int32_t* labels, const float* C_old, float* C_new,
                      uint32_t* counts, kmc_result* res) {
  if (!h || !labels || !C_old || !C_new) return fail(h, KMC_ERR_ARG, "kmc_update: bad argument");
  if (!h->have_points) return fail(h, KMC_ERR_STATE, "kmc_update: points not set");
  for (size_t i = 0; i < h->n; ++i)
    if (labels[i] < 0 || labels[i] >= h->K)
      return fail(h, KMC_ERR_OUT_OF_BOUNDS, "kmc_update: label out of range");
  if (res) std::memset(res, 0, sizeof(*res));
  kmc_status st = kmc_set_centroids(h, C_old);
  if (st != KMC_OK) return st;
  KMC_CK(hipMemcpyAsync(h->labels, labels, h->n * sizeof(int32_t), hipMemcpyHostToDevice,
                        h->stream));
  st = reset_state(h, 1);
  if (st != KMC_OK) return st;
  st = enqueue_update(h, kmc::CONV_NONE, 0);
  if (st != KMC_OK) return st;
  if (counts)
    KMC_CK(hipMemcpyAsync(counts, h->count, h->K * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          h->stream));
  st = kmc_get_centroids(h, C_new);
  if (st != KMC_OK) return st;
  DevState s;
  st = read_state(h, &s);
  if (st == KMC_OK) fill_result(h, s, res);
  return st;
}

}  // extern "C"
