// kmc_kk_api.hip -- C ABI of kernel k-means (kmc_kk_* in include/kmc.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/kmc.h"
#include "kmc_copy.h"
#include "kmc_kk.h"

using kmc::KKState;

struct kmc_kk_handle {
  kmc_kk_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  std::string err;
  kmc::KKParams P{};
  float* dX = nullptr;
  double sum_diag = 0.0;
  bool have_points = false, have_labels = false;
  float ms_build = 0.f;
  // per timed iteration 4 events: start, after the V rebuild (sort or
  // one-hot), after the K product, end
  std::vector<hipEvent_t> events;
  uint64_t timed_iters = 0;
  kmc_setup_times setup{};
};

namespace {

kmc_status kfail(kmc_kk_handle* h, kmc_status s, const std::string& msg) {
  if (h) h->err = msg;
  return s;
}

#define KK_CK(expr)                                                               \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return kfail(h, e_ == hipErrorOutOfMemory ? KMC_ERR_OOM : KMC_ERR_HIP,      \
                   std::string(#expr) + ": " + hipGetErrorString(e_));            \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

void free_kk(kmc_kk_handle* h) {
  kmc::KKParams& P = h->P;
  void* ptrs[] = {h->dX, P.Kmat, P.xn2, P.kdiag, P.HT, P.labels, P.count, P.S,
                  P.gpart, P.g, P.opart, P.cpart, P.state, P.bcount, P.start,
                  P.members, P.tpart, P.gtile};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

kmc_status reset_kk_state(kmc_kk_handle* h, uint64_t maxiter) {
  KKState s{};
  s.maxiter = maxiter;
  s.converged_at = maxiter;
  s.sum_diag = h->sum_diag;
  // s lives on this stack frame: the copy must be done before it returns
  KK_CK(hipMemcpyAsync(h->P.state, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
  KK_CK(hipStreamSynchronize(h->stream));
  return KMC_OK;
}

}  // namespace

extern "C" {

kmc_status kmc_kk_create(const kmc_kk_config* cfg, kmc_kk_handle** out) {
  if (!cfg || !out) return KMC_ERR_ARG;
  *out = nullptr;
  if (cfg->n == 0 || cfg->d == 0 || cfg->k == 0 || cfg->k > cfg->n) return KMC_ERR_ARG;
  if (cfg->product < KMC_KK_AUTO || cfg->product > KMC_KK_SORTED) return KMC_ERR_ARG;
  // the cluster-sorted product streams K at a higher fraction of HBM
  // bandwidth than the dense one-hot product (no operand traffic besides K)
  const int product = cfg->product == KMC_KK_DENSE ? kmc::KK_DENSE : kmc::KK_SORTED;
  if (cfg->k > (uint32_t)(product == kmc::KK_DENSE ? kmc::KK_DENSE_MAX_K : kmc::KK_MAX_K))
    return KMC_ERR_UNSUPPORTED;
  if (cfg->kernel < KMC_KERNEL_LINEAR || cfg->kernel > KMC_KERNEL_RBF) return KMC_ERR_ARG;
  if (cfg->store != KMC_STORE_FP32 && cfg->store != KMC_STORE_BF16) return KMC_ERR_ARG;
  if (cfg->build < KMC_KBUILD_FP32 || cfg->build > KMC_KBUILD_BF16) return KMC_ERR_ARG;
  if (cfg->n > (size_t)0x7fffffff) return KMC_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return KMC_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev) return KMC_ERR_ARG;

  kmc_kk_handle* h = new (std::nothrow) kmc_kk_handle();
  if (!h) return KMC_ERR_OOM;
  h->cfg = *cfg;
  h->dev = cfg->device;
  kmc::KKParams& P = h->P;
  P.n = (long long)cfg->n;
  P.d = (int)cfg->d;
  P.ldx = cfg->d;
  P.Np = (P.n + 63) / 64 * 64;
  P.store = cfg->store == KMC_STORE_FP32 ? kmc::KS_FP32 : kmc::KS_BF16;
  P.kind = cfg->kernel;
  P.build = cfg->build == KMC_KBUILD_BF16X3 ? kmc::KB_BF16X3
            : cfg->build == KMC_KBUILD_BF16 ? kmc::KB_BF16
                                            : kmc::KB_FP32;
  P.gamma = cfg->gamma > 0.0f ? cfg->gamma : 1.0f / (float)cfg->d;
  P.K = (int)cfg->k;
  P.product = product;
  if (product == kmc::KK_DENSE) {
    P.Kp = P.K <= 16 ? 16 : (P.K <= 32 ? 32 : 64);
    P.sS_i = P.Kp;
    P.sS_c = 1;
  } else {
    P.Kp = (P.K + 15) / 16 * 16;
    P.sS_i = 1;
    P.sS_c = P.Np;
    // sort blocking as the Lloyd path: ~4 points per cluster counter per block
    const int ch = ((P.K + 255) / 256) * 256 * 4;
    P.CH = std::max(1024, std::min(16384, ch));
    P.nblk = (int)((P.n + P.CH - 1) / P.CH);
    P.lbits = 0;
    while ((1 << P.lbits) < P.K) ++P.lbits;
    P.nq = (int)((P.n + kmc::kKKRows - 1) / kmc::kKKRows);
    P.ntile = (int)((P.Np + 1023) / 1024);
  }
  P.nassign = (int)((P.n + 255) / 256);
  P.tol = cfg->tol;
  kmc_status st = KMC_OK;
  auto run = [&]() -> kmc_status {
    KK_CK(hipSetDevice(h->dev));
    KK_CK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    const size_t es = P.store == kmc::KS_FP32 ? 4 : 2;
    void* kbuf = nullptr;
    KK_CK(hipMalloc(&kbuf, (size_t)P.Np * (size_t)P.Np * es));
    P.Kmat = kbuf;
    KK_CK(dalloc(&h->dX, (size_t)P.n * P.d));
    KK_CK(dalloc(&P.xn2, (size_t)P.n));
    KK_CK(dalloc(&P.kdiag, (size_t)P.n));
    if (P.product == kmc::KK_DENSE) {
      void* hb = nullptr;
      KK_CK(hipMalloc(&hb, (size_t)P.Kp * (size_t)P.Np * es));
      P.HT = hb;
      KK_CK(dalloc(&P.S, (size_t)P.n * P.Kp));
      KK_CK(dalloc(&P.gpart, (size_t)((P.n + 15) / 16) * P.Kp));
    } else {
      KK_CK(dalloc(&P.S, (size_t)P.Kp * (size_t)P.Np));
      KK_CK(dalloc(&P.bcount, (size_t)P.K * (size_t)P.nblk));
      KK_CK(dalloc(&P.start, (size_t)P.K));
      KK_CK(dalloc(&P.members, (size_t)P.n));
      KK_CK(dalloc(&P.tpart, ((size_t)P.nq + (size_t)P.K) * (size_t)P.Np));
      KK_CK(dalloc(&P.gtile, (size_t)P.K * (size_t)P.ntile));
    }
    KK_CK(dalloc(&P.labels, (size_t)P.n));
    KK_CK(dalloc(&P.count, (size_t)P.Kp));
    KK_CK(dalloc(&P.g, (size_t)P.Kp));
    KK_CK(dalloc(&P.opart, (size_t)P.nassign));
    KK_CK(dalloc(&P.cpart, (size_t)P.nassign));
    KK_CK(dalloc(&P.state, 1));
    KK_CK(hipMemsetAsync(P.state, 0, sizeof(KKState), h->stream));
    KK_CK(hipStreamSynchronize(h->stream));
    P.X = h->dX;
    P.done = &P.state->done;
    return KMC_OK;
  };
  st = run();
  if (st != KMC_OK) {
    free_kk(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return st;
  }
  *out = h;
  return KMC_OK;
}

kmc_status kmc_kk_destroy(kmc_kk_handle* h) {
  if (!h) return KMC_ERR_ARG;
  (void)hipSetDevice(h->dev);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->events) (void)hipEventDestroy(e);
  free_kk(h);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return KMC_OK;
}

const char* kmc_kk_last_error(const kmc_kk_handle* h) { return h ? h->err.c_str() : ""; }

kmc_status kmc_kk_set_points(kmc_kk_handle* h, const float* X, size_t ld) {
  if (!h || !X || ld < (size_t)h->P.d) return kfail(h, KMC_ERR_ARG, "kmc_kk_set_points: bad argument");
  kmc::KKParams& P = h->P;
  KK_CK(hipSetDevice(h->dev));
  hipEvent_t e0, e1;
  KK_CK(hipEventCreate(&e0));
  KK_CK(hipEventCreate(&e1));
  KK_CK(hipEventRecord(e0, h->stream));
  KK_CK(kmc::rows_to_device(h->dX, (size_t)P.d, X, ld, (size_t)P.d, (size_t)P.n, h->stream));
  KK_CK(hipEventRecord(e1, h->stream));
  KK_CK(hipEventSynchronize(e1));
  float mc = 0.f;
  KK_CK(hipEventElapsedTime(&mc, e0, e1));
  h->setup.memcpy_ms = mc;
  // the bf16 builds read X as bf16 planes, split once here
  uint16_t* planes = nullptr;
  P.dpad = (P.d + 31) / 32 * 32;
  if (P.build != kmc::KB_FP32) {
    const size_t np = P.build == kmc::KB_BF16X3 ? 2 : 1;
    KK_CK(dalloc(&planes, np * (size_t)P.n * (size_t)P.dpad));
    P.Xh = planes;
    P.Xl = np == 2 ? planes + (size_t)P.n * (size_t)P.dpad : nullptr;
  }
  KK_CK(hipEventRecord(e0, h->stream));
  hipError_t be = kmc::kk_launch_setup(P, h->stream);
  if (be == hipSuccess) be = hipEventRecord(e1, h->stream);
  if (be == hipSuccess) be = hipEventSynchronize(e1);
  if (planes) (void)hipFree(planes);
  P.Xh = P.Xl = nullptr;
  KK_CK(be);
  KK_CK(hipEventElapsedTime(&h->ms_build, e0, e1));
  h->setup.operands_ms = h->ms_build;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<float> diag((size_t)P.n);
  KK_CK(hipMemcpy(diag.data(), P.kdiag, diag.size() * sizeof(float), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (float v : diag) s += v;
  h->sum_diag = s;
  h->have_points = true;
  return KMC_OK;
}

kmc_status kmc_kk_set_labels(kmc_kk_handle* h, const int32_t* labels0) {
  if (!h || !labels0) return kfail(h, KMC_ERR_ARG, "kmc_kk_set_labels: bad argument");
  for (long long i = 0; i < h->P.n; ++i)
    if (labels0[i] < 0 || labels0[i] >= h->P.K)
      return kfail(h, KMC_ERR_OUT_OF_BOUNDS, "kmc_kk_set_labels: label out of range");
  KK_CK(hipSetDevice(h->dev));
  hipEvent_t e0, e1;
  KK_CK(hipEventCreate(&e0));
  KK_CK(hipEventCreate(&e1));
  KK_CK(hipEventRecord(e0, h->stream));
  KK_CK(hipMemcpyAsync(h->P.labels, labels0, (size_t)h->P.n * sizeof(int32_t),
                       hipMemcpyHostToDevice, h->stream));
  KK_CK(hipEventRecord(e1, h->stream));
  KK_CK(hipEventSynchronize(e1));
  float ms = 0.f;
  KK_CK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->setup.init_ms = ms;
  h->have_labels = true;
  return KMC_OK;
}

kmc_status kmc_kk_run(kmc_kk_handle* h, uint64_t maxiter, uint32_t flags, kmc_kk_result* res) {
  if (!h) return KMC_ERR_ARG;
  if (!h->have_points || !h->have_labels)
    return kfail(h, KMC_ERR_STATE, "kmc_kk_run: points and initial labels must be set");
  if (res) std::memset(res, 0, sizeof(*res));
  KK_CK(hipSetDevice(h->dev));
  kmc::KKParams P = h->P;
  const bool check = (flags & KMC_RUN_CHECK_CONVERGED) != 0;
  P.mode = !check ? kmc::CONV_NONE
                  : ((flags & KMC_RUN_STOP_STABLE) ? kmc::CONV_STABLE : kmc::CONV_SCORE);
  kmc_status st = reset_kk_state(h, maxiter);
  if (st != KMC_OK) return st;
  const bool timing = (flags & KMC_RUN_TIME_KERNELS) != 0;
  const uint64_t ntimed = timing ? std::min<uint64_t>(maxiter, 4096) : 0;
  while (h->events.size() < 4 * ntimed) {
    hipEvent_t e;
    KK_CK(hipEventCreate(&e));
    h->events.push_back(e);
  }
  hipEvent_t t0, t1;
  KK_CK(hipEventCreate(&t0));
  KK_CK(hipEventCreate(&t1));
  KK_CK(hipEventRecord(t0, h->stream));
  uint64_t launched = 0, chunk = 1;
  KKState s{};
  while (launched < maxiter) {
    const uint64_t m = check ? std::min<uint64_t>(chunk, maxiter - launched) : maxiter - launched;
    for (uint64_t it = 0; it < m; ++it) {
      const uint64_t gi = launched + it;
      const bool tm = gi < ntimed;
      if (tm) KK_CK(hipEventRecord(h->events[4 * gi], h->stream));
      KK_CK(kmc::kk_launch_iter(P, tm ? h->events[4 * gi + 1] : nullptr,
                                tm ? h->events[4 * gi + 2] : nullptr, h->stream));
      if (tm) KK_CK(hipEventRecord(h->events[4 * gi + 3], h->stream));
    }
    launched += m;
    if (check) {
      KK_CK(hipMemcpyAsync(&s, P.state, sizeof(s), hipMemcpyDeviceToHost, h->stream));
      KK_CK(hipStreamSynchronize(h->stream));
      if (s.done) break;
      chunk = std::min<uint64_t>(chunk * 2, 16);
    }
  }
  KK_CK(hipEventRecord(t1, h->stream));
  KK_CK(hipEventSynchronize(t1));
  float ms = 0.f;
  KK_CK(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  KK_CK(hipMemcpy(&s, P.state, sizeof(s), hipMemcpyDeviceToHost));
  h->timed_iters = std::min<uint64_t>(s.iter, ntimed);
  if (res) {
    res->iterations = s.iter;
    res->converged_at = s.converged_at;
    res->objective = s.objective;
    res->score = s.objective - s.sum_diag;
    res->changed = s.changed;
    res->ms_total = ms;
    res->ms_build = h->ms_build;
    const uint64_t ran = std::min<uint64_t>(s.iter, ntimed);
    double acc = 0.0;
    for (uint64_t i = 0; i < ran; ++i) {
      float a = 0.f;
      if (hipEventElapsedTime(&a, h->events[4 * i + 1], h->events[4 * i + 2]) == hipSuccess)
        acc += a;
    }
    res->ms_spmm = acc;
    res->spmm_launches = ran;
  }
  return KMC_OK;
}

kmc_status kmc_kk_get_setup_times(const kmc_kk_handle* h, kmc_setup_times* out) {
  if (!h || !out) return KMC_ERR_ARG;
  *out = h->setup;
  return KMC_OK;
}

kmc_status kmc_kk_get_stage_times(kmc_kk_handle* h, kmc_stage_times* out, uint64_t cap,
                                  uint64_t* count) {
  if (!h || (!out && cap > 0)) return kfail(h, KMC_ERR_ARG, "kmc_kk_get_stage_times: bad argument");
  KK_CK(hipSetDevice(h->dev));
  const uint64_t m = std::min<uint64_t>(cap, h->timed_iters);
  for (uint64_t i = 0; i < m; ++i) {
    const hipEvent_t* ev = &h->events[4 * i];
    float c = 0.f, a = 0.f, r = 0.f;
    KK_CK(hipEventElapsedTime(&c, ev[0], ev[1]));
    KK_CK(hipEventElapsedTime(&a, ev[1], ev[2]));
    KK_CK(hipEventElapsedTime(&r, ev[2], ev[3]));
    out[i].distances_ms = a;
    out[i].argmin_ms = r;
    out[i].centroids_ms = c;
  }
  if (count) *count = h->timed_iters;
  return KMC_OK;
}

kmc_status kmc_kk_get_labels(kmc_kk_handle* h, int32_t* labels) {
  if (!h || !labels) return kfail(h, KMC_ERR_ARG, "kmc_kk_get_labels: bad argument");
  KK_CK(hipSetDevice(h->dev));
  KK_CK(hipMemcpy(labels, h->P.labels, (size_t)h->P.n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return KMC_OK;
}

kmc_status kmc_kk_distances(kmc_kk_handle* h, const int32_t* labels, float* D) {
  if (!h || !labels || !D) return kfail(h, KMC_ERR_ARG, "kmc_kk_distances: bad argument");
  if (!h->have_points) return kfail(h, KMC_ERR_STATE, "kmc_kk_distances: points not set");
  kmc_status st = kmc_kk_set_labels(h, labels);
  if (st != KMC_OK) return st;
  st = reset_kk_state(h, 1);
  if (st != KMC_OK) return st;
  kmc::KKParams P = h->P;
  float* dD = nullptr;
  KK_CK(dalloc(&dD, (size_t)P.n * P.K));
  P.Dout = dD;
  hipError_t e = kmc::kk_launch_distances(P, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(D, dD, (size_t)P.n * P.K * sizeof(float), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(dD);
  KK_CK(e);
  return KMC_OK;
}

kmc_status kmc_kk_get_kernel_rows(kmc_kk_handle* h, size_t row0, size_t nrows, float* out) {
  if (!h || !out || row0 + nrows > (size_t)h->P.n)
    return kfail(h, KMC_ERR_ARG, "kmc_kk_get_kernel_rows: bad argument");
  if (!h->have_points) return kfail(h, KMC_ERR_STATE, "kmc_kk_get_kernel_rows: points not set");
  const kmc::KKParams& P = h->P;
  KK_CK(hipSetDevice(h->dev));
  if (P.store == kmc::KS_FP32) {
    KK_CK(kmc::rows_to_host(out, (size_t)P.n,
                            reinterpret_cast<const float*>(P.Kmat) + row0 * (size_t)P.Np,
                            (size_t)P.Np, (size_t)P.n, nrows, h->stream));
  } else {
    std::vector<uint16_t> tmp(nrows * (size_t)P.n);
    KK_CK(kmc::rows_to_host(tmp.data(), (size_t)P.n,
                            reinterpret_cast<const uint16_t*>(P.Kmat) + row0 * (size_t)P.Np,
                            (size_t)P.Np, (size_t)P.n, nrows, h->stream));
    for (size_t i = 0; i < tmp.size(); ++i) {
      const uint32_t u = ((uint32_t)tmp[i]) << 16;
      float f;
      std::memcpy(&f, &u, 4);
      out[i] = f;
    }
  }
  return KMC_OK;
}

}  // extern "C"
