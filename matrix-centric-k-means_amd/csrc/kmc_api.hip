// kmc_api.hip -- the C ABI (include/kmc.h): handle, buffers, Lloyd driver.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/kmc.h"
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
  float* xn2 = nullptr;
  // centroids
  float* C = nullptr;
  float* Cnext = nullptr;
  float* Cw = nullptr;
  uint16_t* Chi = nullptr;
  uint16_t* Clo = nullptr;
  float* cn = nullptr;
  // assignment
  int* labels = nullptr;
  int* q2 = nullptr;
  int* qa = nullptr;
  unsigned* qcnt = nullptr;
  // sort / update
  int CH = 0, nblk = 0;
  int* bcount = nullptr;
  int* count = nullptr;
  int* start = nullptr;
  int* members = nullptr;
  double* ipart = nullptr;
  int nparts = 0;
  int* conv = nullptr;
  double* shift2 = nullptr;
  int* alive = nullptr;
  DevState* state = nullptr;

  // launch configuration
  int kchunk = 0;
  size_t assign_lds = 0;
  int assign_grid = 0;
  float alpha = 0, beta = 0, gref = 0;

  bool have_points = false, have_centroids = false;
  std::vector<hipEvent_t> events;
};

namespace {

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

// gamma_m = m u / (1 - m u), the classical bound for m rounded operations.
double gam(double m, double u) { return m * u / (1.0 - m * u); }

// Margin-band constants. For every centroid j the MFMA value b_j satisfies
// |b_j - (||c_j||^2 - 2 x.c_j)| <= alpha ||x|| cmax + beta cmax^2, and the
// reference's direct float form r_j satisfies |r_j - d_j| <= gref d_j. The
// constants carry a factor 2 of headroom over the derivation below.
void band_constants(int prec, int D, float* alpha, float* beta, float* gref) {
  const double u = std::ldexp(1.0, -24);  // fp32 unit roundoff
  const double ub = std::ldexp(1.0, -9);  // bf16 unit roundoff
  double t_in, gA;
  if (prec == kmc::PREC_FP32) {
    t_in = 0.0;             // operands exact
    gA = gam(D + 2, u);     // product rounding + D-term accumulation
  } else if (prec == kmc::PREC_BF16X3) {
    t_in = 3.2 * ub * ub;   // hi+lo split: dropped lo*lo and residuals
    gA = gam(3 * D + 1, u); // products exact in fp32, 3D-term accumulation
  } else {
    t_in = 2.0 * ub + ub * ub;
    gA = gam(D + 1, u);
  }
  const double gD = gam(D, u);  // ||c||^2 in fp32
  // dot term: |x.(-2c)| error <= 2 (t_in + gA (1 + t_in)) ||x|| ||c||
  const double a = 2.0 * (t_in + gA * (1.0 + t_in));
  const double b = gD + gA * (1.0 + gD);
  *alpha = (float)(2.0 * a);
  *beta = (float)(2.0 * b);
  *gref = (float)(2.0 * gam(D + 3, u));
}

kmc_status alloc_all(kmc_handle* h) {
  const size_t n = h->n;
  const size_t DP = h->DP, K = h->K, Kp = h->Kp;
  KMC_CK(dalloc(&h->X, n * DP));
  if (h->prec != kmc::PREC_FP32) KMC_CK(dalloc(&h->Xhi, n * DP));
  if (h->prec == kmc::PREC_BF16X3) KMC_CK(dalloc(&h->Xlo, n * DP));
  KMC_CK(dalloc(&h->xn2, n));
  KMC_CK(dalloc(&h->C, K * DP));
  KMC_CK(dalloc(&h->Cnext, K * DP));
  if (h->prec == kmc::PREC_FP32) KMC_CK(dalloc(&h->Cw, Kp * DP));
  else KMC_CK(dalloc(&h->Chi, Kp * DP));
  if (h->prec == kmc::PREC_BF16X3) KMC_CK(dalloc(&h->Clo, Kp * DP));
  KMC_CK(dalloc(&h->cn, Kp));
  KMC_CK(dalloc(&h->labels, n));
  KMC_CK(dalloc(&h->q2, 3 * n));
  KMC_CK(dalloc(&h->qa, n));
  KMC_CK(dalloc(&h->qcnt, 2));
  KMC_CK(dalloc(&h->bcount, K * (size_t)h->nblk));
  KMC_CK(dalloc(&h->count, K));
  KMC_CK(dalloc(&h->start, K));
  KMC_CK(dalloc(&h->members, n));
  KMC_CK(dalloc(&h->ipart, (size_t)h->nparts));
  KMC_CK(dalloc(&h->conv, K));
  KMC_CK(dalloc(&h->shift2, K));
  KMC_CK(dalloc(&h->alive, K));
  KMC_CK(dalloc(&h->state, 1));
  KMC_CK(kmc::launch_fill_i32(h->alive, (long long)K, 1, h->stream));
  KMC_CK(hipMemsetAsync(h->C, 0, K * DP * sizeof(float), h->stream));
  KMC_CK(hipMemsetAsync(h->Cnext, 0, K * DP * sizeof(float), h->stream));
  KMC_CK(hipMemsetAsync(h->state, 0, sizeof(DevState), h->stream));
  KMC_CK(hipMemsetAsync(h->qcnt, 0, 2 * sizeof(unsigned), h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
  return KMC_OK;
}

void free_all(kmc_handle* h) {
  void* ptrs[] = {h->X, h->Xhi, h->Xlo, h->xn2, h->C, h->Cnext, h->Cw, h->Chi,
                  h->Clo, h->cn, h->labels, h->q2, h->qa, h->qcnt, h->bcount,
                  h->count, h->start, h->members, h->ipart, h->conv, h->shift2,
                  h->alive, h->state};
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
  KMC_CK(hipMemcpyAsync(h->state, &s, sizeof(DevState), hipMemcpyHostToDevice, h->stream));
  return KMC_OK;
}

kmc_status enqueue_assign(kmc_handle* h, hipEvent_t e0, hipEvent_t e1) {
  kmc::PrepParams pp{};
  pp.C = h->C;
  pp.Cw = h->Cw;
  pp.Chi = h->Chi;
  pp.Clo = h->Clo;
  pp.cn = h->cn;
  pp.qcnt = h->qcnt;
  pp.alive = h->alive;
  pp.done = done_ptr(h);
  pp.K = h->K;
  pp.Kp = h->Kp;
  pp.DP = h->DP;
  pp.prec = h->prec;
  KMC_CK(kmc::launch_prep(pp, h->stream));

  kmc::AssignParams ap{};
  ap.X = h->X;
  ap.Xhi = h->Xhi;
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
  ap.alpha = h->alpha;
  ap.beta = h->beta;
  ap.gref = h->gref;
  ap.labels = h->labels;
  ap.q2 = h->q2;
  ap.qa = h->qa;
  ap.qcnt = h->qcnt;
  ap.done = done_ptr(h);
  if (e0) KMC_CK(hipEventRecord(e0, h->stream));
  KMC_CK(kmc::launch_assign(h->DP, h->prec, ap, h->assign_grid, h->assign_lds, h->stream));
  if (e1) KMC_CK(hipEventRecord(e1, h->stream));

  kmc::RefineParams rp{};
  rp.X = h->X;
  rp.C = h->C;
  rp.q2 = h->q2;
  rp.qa = h->qa;
  rp.qcnt = h->qcnt;
  rp.alive = h->alive;
  rp.labels = h->labels;
  rp.done = done_ptr(h);
  rp.K = h->K;
  rp.D = h->D;
  rp.DP = h->DP;
  const int g2 = std::max(1, std::min(h->cus * 4, (int)((h->n + 255) / 256)));
  const int gA = std::max(1, std::min(h->cus * 2, (int)((h->n + 3) / 4)));
  KMC_CK(kmc::launch_refine(rp, g2, gA, h->stream));
  return KMC_OK;
}

kmc_status enqueue_update(kmc_handle* h, int mode, int exclude_empty) {
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
  KMC_CK(kmc::launch_sort(sp, h->stream));

  kmc::UpdateParams up{};
  up.X = h->X;
  up.C = h->C;
  up.Cnext = h->Cnext;
  up.members = h->members;
  up.start = h->start;
  up.count = h->count;
  up.ipart = h->ipart;
  up.done = done_ptr(h);
  up.K = h->K;
  up.D = h->D;
  up.DP = h->DP;
  KMC_CK(kmc::launch_update(up, h->stream));

  kmc::FinishParams fp{};
  fp.C = h->C;
  fp.Cnext = h->Cnext;
  fp.conv = h->conv;
  fp.shift2 = h->shift2;
  fp.count = h->count;
  fp.alive = h->alive;
  fp.done = done_ptr(h);
  fp.tol = h->cfg.tol;
  fp.K = h->K;
  fp.D = h->D;
  fp.DP = h->DP;
  fp.exclude_empty = exclude_empty;
  kmc::StatsParams st{};
  st.ipart = h->ipart;
  st.nparts = h->nparts;
  st.shift2 = h->shift2;
  st.conv = h->conv;
  st.qcnt = h->qcnt;
  st.state = h->state;
  st.done = done_ptr(h);
  st.tol = h->cfg.tol;
  st.K = h->K;
  st.mode = mode;
  KMC_CK(kmc::launch_finish(fp, st, h->stream));
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
  (void)h;
}

kmc_status load_points(kmc_handle* h, const float* dsrc, size_t ld) {
  double* xn64 = nullptr;
  KMC_CK(dalloc(&xn64, h->n));
  kmc::LoadParams lp{};
  lp.src = dsrc;
  lp.ld = ld;
  lp.n = (long long)h->n;
  lp.D = h->D;
  lp.DP = h->DP;
  lp.X = h->X;
  lp.Xhi = h->Xhi;
  lp.Xlo = h->Xlo;
  lp.xn2 = h->xn2;
  lp.xn2_64 = xn64;
  hipError_t e = kmc::launch_load_points(lp, h->stream);
  if (e == hipSuccess) e = kmc::launch_sum_f64(xn64, (long long)h->n, &h->state->sum_x2, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(xn64);
  KMC_CK(e);
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
  if (cfg->precision < 0 || cfg->precision > 2) return KMC_ERR_ARG;
  if (cfg->update != KMC_UPDATE_ORDERED) return KMC_ERR_ARG;
  if (!(cfg->tol >= 0.0f)) return KMC_ERR_ARG;
  const int DP = pick_dp((int)cfg->d);
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
  h->Kp = (h->K + 31) / 32 * 32;
  h->prec = cfg->precision;
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
    h->nparts = h->K * ((h->D + 63) / 64);

    // assign launch shape: all centroids resident in LDS when they fit
    const size_t kResidentMax = 144 * 1024, kChunkMax = 64 * 1024;
    h->kchunk = h->Kp;
    h->assign_lds = kmc::assign_lds(DP, h->prec, h->kchunk);
    if (h->assign_lds > kResidentMax) {
      int kc = 32;
      while (kc + 32 <= h->Kp && kmc::assign_lds(DP, h->prec, kc + 32) <= kChunkMax) kc += 32;
      h->kchunk = kc;
      h->assign_lds = kmc::assign_lds(DP, h->prec, kc);
    }
    const void* fn = kmc::assign_kernel_ptr(DP, h->prec);
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)h->assign_lds) != hipSuccess) {
      st = fail(h, KMC_ERR_HIP, "hipFuncSetAttribute(assign)");
      break;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, h->assign_lds) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    const long long tiles = (long long)((h->n + 127) / 128);
    h->assign_grid = (int)std::min<long long>(tiles, (long long)per_cu * h->cus);
    band_constants(h->prec, h->D, &h->alpha, &h->beta, &h->gref);
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
  hipError_t e = hipMemcpyAsync(tmp, X, bytes, hipMemcpyHostToDevice, h->stream);
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
  return load_points(h, dX, ld);
}

kmc_status kmc_set_centroids(kmc_handle* h, const float* C0) {
  if (!h || !C0) return fail(h, KMC_ERR_ARG, "kmc_set_centroids: bad argument");
  KMC_CK(hipSetDevice(h->dev));
  KMC_CK(hipMemcpy2DAsync(h->C, h->DP * sizeof(float), C0, h->D * sizeof(float),
                          h->D * sizeof(float), h->K, hipMemcpyHostToDevice, h->stream));
  KMC_CK(kmc::launch_fill_i32(h->alive, (long long)h->K, 1, h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
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
  const int mode = !check ? kmc::CONV_NONE : (score_mode ? kmc::CONV_SCORE : kmc::CONV_SHIFT);
  if (!exclude) KMC_CK(kmc::launch_fill_i32(h->alive, (long long)h->K, 1, h->stream));
  const uint64_t kMaxTimed = 4096;
  const uint64_t ntimed = timing ? std::min<uint64_t>(maxiter, kMaxTimed) : 0;
  while (h->events.size() < 2 * ntimed) {
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
  while (launched < maxiter) {
    const uint64_t m = check ? std::min<uint64_t>(chunk, maxiter - launched) : maxiter - launched;
    for (uint64_t it = 0; it < m; ++it) {
      const uint64_t gi = launched + it;
      hipEvent_t e0 = gi < ntimed ? h->events[2 * gi] : nullptr;
      hipEvent_t e1 = gi < ntimed ? h->events[2 * gi + 1] : nullptr;
      st = enqueue_assign(h, e0, e1);
      if (st == KMC_OK) st = enqueue_update(h, mode, exclude ? 1 : 0);
      if (st != KMC_OK) break;
    }
    if (st != KMC_OK) break;
    launched += m;
    if (check) {
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
    if (st == KMC_OK) {
      fill_result(h, s, res);
      if (res) {
        res->ms_total = ms;
        const uint64_t ran = std::min<uint64_t>(s.iter, ntimed);
        double acc = 0.0;
        for (uint64_t i = 0; i < ran; ++i) {
          float a = 0.f;
          if (hipEventElapsedTime(&a, h->events[2 * i], h->events[2 * i + 1]) == hipSuccess)
            acc += a;
        }
        res->ms_assign = acc;
        res->assign_launches = ran;
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
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
  KMC_CK(hipMemcpy2DAsync(C, h->D * sizeof(float), h->C, h->DP * sizeof(float),
                          h->D * sizeof(float), h->K, hipMemcpyDeviceToHost, h->stream));
  KMC_CK(hipStreamSynchronize(h->stream));
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
  unsigned q[2] = {0, 0};
  KMC_CK(hipMemcpyAsync(q, h->qcnt, sizeof(q), hipMemcpyDeviceToHost, h->stream));
  st = kmc_get_labels(h, labels);
  if (st != KMC_OK) return st;
  if (res) {
    res->refined_two = q[0];
    res->refined_full = q[1];
  }
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
