// kmc_static.hip -- kmc_assign_static: the assignment stage of kmc_assign
// with no device allocation at all, for callers that cannot allocate (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// Every buffer the stage needs besides its inputs and outputs (point norms,
// the prepared centroid operands, the refinement queues, the stop flag) lives
// in one file-scope device array, `scratch`, of kScratchBytes. Each GPU the
// library runs on holds its own copy of the array until the process ends.
// Host code never uses the array's name as an address: the first call on a
// device (that device current) takes the device address with
// hipGetSymbolAddress and keeps it per device; kernels get that address.
//
// Every call on a GPU shares that GPU's array and leaves its contents there:
// calls that overlap on one GPU (two streams, two threads) race. What must
// start at a known value (stop flag, liveness flags) is filled by a kernel at
// every call, so a captured graph replays correctly.
#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/kmc.h"
#include "kmc_kernels.h"

namespace {

constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Limits of the allocation-free path.
constexpr size_t kStaticMaxN = size_t(1) << 19;     // points
constexpr int kStaticMaxK = 1024;                   // clusters
constexpr int kStaticMaxKp = kStaticMaxK;           // operand rows (K rounded to 64)
constexpr size_t kStaticQueue = kStaticMaxN + 65536;  // refinement queue entries

// Byte layout of the scratch array (each part 256-byte aligned).
constexpr size_t kOffDone = 0;
constexpr size_t kOffAlive = kOffDone + kAlign;
constexpr size_t kOffXn2 = kOffAlive + align_up(sizeof(int) * kStaticMaxK);
constexpr size_t kOffCw = kOffXn2 + align_up(sizeof(float) * kStaticMaxN);
constexpr size_t kOffCn = kOffCw + align_up(sizeof(float) * kStaticMaxKp * kmc::kMaxDP);
constexpr size_t kOffCT = kOffCn + align_up(sizeof(float) * kStaticMaxKp);
constexpr size_t kOffPhi = kOffCT + align_up(sizeof(float) * kStaticMaxKp * kmc::kMaxDP);
constexpr size_t kOffPlo = kOffPhi + align_up(sizeof(uint16_t) * kStaticMaxKp * kmc::kMaxDP);
constexpr size_t kOffQ2cnt = kOffPlo + align_up(sizeof(uint16_t) * kStaticMaxKp * kmc::kMaxDP);
constexpr size_t kOffQacnt = kOffQ2cnt + align_up(sizeof(unsigned) * kmc::kMaxRegions);
constexpr size_t kOffQ2 = kOffQacnt + align_up(sizeof(unsigned) * kmc::kMaxRegions);
constexpr size_t kOffQa = kOffQ2 + align_up(sizeof(int) * 4 * kStaticQueue);
constexpr size_t kScratchBytes = kOffQa + align_up(sizeof(int) * kStaticQueue);
static_assert(kScratchBytes < (size_t(20) << 20), "scratch stays under 20 MiB");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
int g_cus[kMaxDevices] = {};

// The device address of this device's scratch array (taken once per device,
// with that device current).
hipError_t scratch_of(int dev, unsigned char** out, int* cus) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    int n = 0;
    e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    g_ptr[dev] = static_cast<unsigned char*>(p);
    g_cus[dev] = std::max(1, n);
  }
  *out = g_ptr[dev];
  *cus = g_cus[dev];
  return hipSuccess;
}

int pick_dp(uint32_t d) {
  for (int o : {16, 32, 64, 128, 256})
    if (d <= (uint32_t)o) return o;
  return 0;
}

}  // namespace

extern "C" kmc_status kmc_assign_static(const float* X, size_t n, uint32_t d, size_t ld,
                                        const float* C, uint32_t k, int32_t* labels,
                                        void* stream) {
  if (!X || !C || !labels || n == 0 || d == 0 || k == 0 || k > n) return KMC_ERR_ARG;
  const int DP = pick_dp(d);
  if (DP == 0 || ld != (size_t)DP) return KMC_ERR_ARG;
  if (n > kStaticMaxN || k > (uint32_t)kStaticMaxK) return KMC_ERR_UNSUPPORTED;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return KMC_ERR_NO_DEVICE;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;  // no device made current: the default
  unsigned char* base = nullptr;
  int cus = 0;
  if (scratch_of(dev, &base, &cus) != hipSuccess) return KMC_ERR_HIP;

  const int K = (int)k, Kp = (K + 63) / 64 * 64;
  int* done = reinterpret_cast<int*>(base + kOffDone);
  int* alive = reinterpret_cast<int*>(base + kOffAlive);
  float* xn2 = reinterpret_cast<float*>(base + kOffXn2);
  float* Cw = reinterpret_cast<float*>(base + kOffCw);
  float* cn = reinterpret_cast<float*>(base + kOffCn);
  float* CT = reinterpret_cast<float*>(base + kOffCT);
  uint16_t* Phi = reinterpret_cast<uint16_t*>(base + kOffPhi);
  uint16_t* Plo = reinterpret_cast<uint16_t*>(base + kOffPlo);
  unsigned* q2cnt = reinterpret_cast<unsigned*>(base + kOffQ2cnt);
  unsigned* qacnt = reinterpret_cast<unsigned*>(base + kOffQacnt);
  int* q2 = reinterpret_cast<int*>(base + kOffQ2);
  int* qa = reinterpret_cast<int*>(base + kOffQa);

  // Launch shape of the distance kernel: the first launch variant built for
  // this width (the one a new handle starts with), centroids resident in LDS
  // when they fit, else in chunks -- as kmc_create chooses.
  int variant = -1;
  const void* fn = nullptr;
  for (int v = 0; v < kmc::assign_variants() && !fn; ++v) {
    fn = kmc::assign_kernel_ptr(DP, kmc::PREC_FP32, v);
    variant = v;
  }
  if (!fn) return KMC_ERR_UNSUPPORTED;
  const size_t kResidentMax = 144 * 1024, kChunkMax = 64 * 1024;
  int kchunk = Kp;
  size_t lds = kmc::assign_lds(DP, kmc::PREC_FP32, kchunk);
  if (lds > kResidentMax) {
    int kc = 64;
    while (kc + 64 <= Kp && kmc::assign_lds(DP, kmc::PREC_FP32, kc + 64) <= kChunkMax) kc += 64;
    kchunk = kc;
    lds = kmc::assign_lds(DP, kmc::PREC_FP32, kc);
  }
  if (kmc::raise_lds_limit(fn, lds) != hipSuccess) return KMC_ERR_HIP;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kmc::assign_threads(variant),
                                                   lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const long long tile = kmc::assign_tile(variant);
  const long long tiles = ((long long)n + tile - 1) / tile;
  const long long g0 = std::min<long long>({tiles, (long long)per_cu * cus,
                                            (long long)kmc::kMaxRegions});
  // each workgroup's queue region holds all its tiles' points; the grid is
  // trimmed so that no region is left without a tile
  const long long per = (tiles + g0 - 1) / g0;
  const int grid = (int)((tiles + per - 1) / per);
  const long long qstride = per * tile;
  if ((size_t)grid * (size_t)qstride > kStaticQueue) return KMC_ERR_UNSUPPORTED;

  int idx_bits = 0;
  while ((1 << idx_bits) < Kp) ++idx_bits;
  float alpha = 0, beta = 0, gref = 0, xpk = 0, hop = 0;
  kmc::band_constants(kmc::PREC_FP32, (int)d, idx_bits, &alpha, &beta, &gref, &xpk, true, &hop);
  float alpha3 = 0, beta3 = 0, gref3 = 0, xpk3 = 0;
  kmc::band_constants(kmc::PREC_BF16X3, (int)d, idx_bits, &alpha3, &beta3, &gref3, &xpk3);

  if (kmc::launch_fill_i32(done, 1, 0, s) != hipSuccess) return KMC_ERR_HIP;
  if (kmc::launch_fill_i32(alive, K, 1, s) != hipSuccess) return KMC_ERR_HIP;

  // ||x||^2 of the caller's (already padded) rows, as kmc_set_points forms it
  kmc::LoadParams lp{};
  lp.src = X;
  lp.ld = ld;
  lp.n = (long long)n;
  lp.D = (int)d;
  lp.DP = DP;
  lp.xn2 = xn2;
  if (kmc::launch_load_points(lp, s) != hipSuccess) return KMC_ERR_HIP;

  kmc::PrepParams pp{};
  pp.C = C;
  pp.Cw = Cw;
  pp.cn = cn;
  pp.CT = CT;
  pp.Phi = Phi;
  pp.Plo = Plo;
  pp.alive = alive;
  pp.done = done;
  pp.K = K;
  pp.Kp = Kp;
  pp.DP = DP;
  pp.prec = kmc::PREC_FP32;
  if (kmc::launch_prep(pp, s) != hipSuccess) return KMC_ERR_HIP;

  kmc::AssignParams ap{};
  ap.X = X;
  ap.xn2 = xn2;
  ap.n = (long long)n;
  ap.Cw = Cw;
  ap.cn = cn;
  ap.hop = hop;
  ap.K = K;
  ap.Kp = Kp;
  ap.kchunk = kchunk;
  ap.idx_mask = (1u << idx_bits) - 1u;
  ap.alpha = alpha;
  ap.beta = beta;
  ap.gref = gref;
  ap.xpk = xpk;
  ap.labels = labels;
  ap.q2 = q2;
  ap.qa = qa;
  ap.q2cnt = q2cnt;
  ap.qacnt = qacnt;
  ap.qstride = qstride;
  ap.done = done;
  if (kmc::launch_assign(DP, kmc::PREC_FP32, variant, ap, grid, lds, s) != hipSuccess)
    return KMC_ERR_HIP;

  kmc::RefineParams rp{};
  rp.X = X;
  rp.C = C;
  rp.CT = CT;
  rp.q2 = q2;
  rp.qa = qa;
  rp.q2cnt = q2cnt;
  rp.qacnt = qacnt;
  rp.qstride = qstride;
  rp.nregions = grid;
  rp.alive = alive;
  rp.labels = labels;
  rp.done = done;
  rp.K = K;
  rp.Kp = Kp;
  rp.D = (int)d;
  rp.DP = DP;
  rp.cn = cn;
  rp.xn2 = xn2;
  rp.Phi = Phi;
  rp.Plo = Plo;
  rp.alpha3 = alpha3;
  rp.beta3 = beta3;
  rp.gref3 = gref3;
  rp.pq = 1;
  rp.prune_min = kmc::pq_shape(K, DP) ? 0 : kmc::kPruneMinPoints;
  if (kmc::launch_refine(rp, s) != hipSuccess) return KMC_ERR_HIP;
  return KMC_OK;
}
