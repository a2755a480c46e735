// kmc_probe.hip -- kmc_probe_mfma: one 32 x 32 block of the distance GEMM's
// matrix-core products on caller operands, for checking the error model the
// margin band is built on (the band constants assume the operand rounding of
// the stored forms and an accumulation error bound per form).
//
// The operands are rounded exactly as the engine stores them (fp32: as is;
// bf16: round to nearest even; fp8: one power-of-two scale per row so that
// its largest entry lands in [224, 448], then e4m3 round to nearest even),
// then one wave runs the engine's MFMA chain for one block (the same
// instructions, k-step order and lane layout as the distance kernel) from a
// zero accumulator.
#include <cstring>
#include <vector>

#include "../../include/kmc.h"
#include "kmc_device.h"
#include "kmc_kernels.h"

namespace {

using kmc::bf16_rne;
using kmc::bf16_to_f32;
using kmc::fp8_e4m3_rne;
using kmc::fp8_e4m3_value;
using kmc::fp8_scale_exp;

// A, B: 32 x D fp32 rows. out: 32 x 32, out[i * 32 + j] = A_i . B_j as the
// matrix core forms it. Aq, Bq: the stored operand values (decoded, with the
// row scale applied), 32 x D.
template <int D, int PREC>
__global__ __launch_bounds__(64) void k_probe_mfma(const float* A, const float* B, float* out,
                                                  float* Aq, float* Bq) {
  __shared__ __attribute__((aligned(16))) uint16_t sh[2][32][D];  // bf16 rows
  __shared__ __attribute__((aligned(16))) uint8_t sq[2][32][D];   // e4m3 rows
  __shared__ int ssc[2][32];                                      // E8M0 scales
  const int lane = threadIdx.x;
  const int p = lane & 31, h = lane >> 5;
  // lanes 0..31 store row p of A, lanes 32..63 row p of B
  {
    const float* src = (h == 0 ? A : B) + (size_t)p * D;
    float* dq = (h == 0 ? Aq : Bq) + (size_t)p * D;
    if (PREC == kmc::PREC_FP8) {
      float amax = 0.0f;
      for (int j = 0; j < D; ++j) amax = fmaxf(amax, fabsf(src[j]));
      const int e = fp8_scale_exp(amax);
      for (int j = 0; j < D; ++j) {
        const uint32_t code = fp8_e4m3_rne(ldexpf(src[j], -e));
        sq[h][p][j] = (uint8_t)code;
        dq[j] = ldexpf(fp8_e4m3_value(code), e);
      }
      ssc[h][p] = e + 127;
    } else if (PREC == kmc::PREC_BF16) {
      for (int j = 0; j < D; ++j) {
        const uint16_t b = bf16_rne(src[j]);
        sh[h][p][j] = b;
        dq[j] = bf16_to_f32(b);
      }
    } else {
      for (int j = 0; j < D; ++j) dq[j] = src[j];
    }
  }
  __syncthreads();
  kmc_f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  if constexpr (PREC == kmc::PREC_FP8) {
    // step s: bytes 64 s + 32 h + [0, 32) of each row
#pragma unroll
    for (int s = 0; s < D / 64; ++s) {
      const uint4 a0 = *reinterpret_cast<const uint4*>(&sq[0][p][64 * s + 32 * h]);
      const uint4 a1 = *reinterpret_cast<const uint4*>(&sq[0][p][64 * s + 32 * h + 16]);
      const uint4 b0 = *reinterpret_cast<const uint4*>(&sq[1][p][64 * s + 32 * h]);
      const uint4 b1 = *reinterpret_cast<const uint4*>(&sq[1][p][64 * s + 32 * h + 16]);
      const kmc_i32x8 av = {(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                            (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
      const kmc_i32x8 bv = {(int)b0.x, (int)b0.y, (int)b0.z, (int)b0.w,
                            (int)b1.x, (int)b1.y, (int)b1.z, (int)b1.w};
      acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, ssc[0][p], 0,
                                                            ssc[1][p]);
    }
  } else if (PREC == kmc::PREC_BF16) {
    // step s: dimensions 16 s + 8 h + [0, 8)
#pragma unroll
    for (int s = 0; s < D / 16; ++s) {
      const uint4 a = *reinterpret_cast<const uint4*>(&sh[0][p][16 * s + 8 * h]);
      const uint4 b = *reinterpret_cast<const uint4*>(&sh[1][p][16 * s + 8 * h]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(kmc_bf16x8, a),
                                                    __builtin_bit_cast(kmc_bf16x8, b), acc, 0,
                                                    0, 0);
    }
  } else {
    // fp32 step (q, c): dimension 8 q + 4 h + c
    const float* ar = A + (size_t)p * D;
    const float* br = B + (size_t)p * D;
#pragma unroll
    for (int q = 0; q < D / 8; ++q) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[8 * q + 4 * h + c], br[8 * q + 4 * h + c],
                                                   acc, 0, 0, 0);
    }
  }
  // accumulator r of lane (p, h): A row 8 (r >> 2) + 4 h + (r & 3), B row p
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r >> 2) + 4 * h + (r & 3);
    out[i * 32 + p] = acc[r];
  }
}

template <int PREC>
const void* probe_kernel(uint32_t d) {
  switch (d) {
    case 16: return PREC == kmc::PREC_FP8 ? nullptr
                                          : reinterpret_cast<const void*>(k_probe_mfma<16, PREC>);
    case 32: return PREC == kmc::PREC_FP8 ? nullptr
                                          : reinterpret_cast<const void*>(k_probe_mfma<32, PREC>);
    case 64: return reinterpret_cast<const void*>(k_probe_mfma<64, PREC>);
    case 128: return reinterpret_cast<const void*>(k_probe_mfma<128, PREC>);
    case 256: return reinterpret_cast<const void*>(k_probe_mfma<256, PREC>);
    default: return nullptr;
  }
}

}  // namespace

extern "C" kmc_status kmc_probe_mfma(int32_t precision, uint32_t d, const float* A,
                                     const float* B, float* out, float* Aq, float* Bq) {
  if (!A || !B || !out || !Aq || !Bq) return KMC_ERR_ARG;
  const void* fn = precision == KMC_PREC_FP32   ? probe_kernel<kmc::PREC_FP32>(d)
                   : precision == KMC_PREC_BF16 ? probe_kernel<kmc::PREC_BF16>(d)
                   : precision == KMC_PREC_FP8  ? probe_kernel<kmc::PREC_FP8>(d)
                                                : nullptr;
  if (!fn) return KMC_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return KMC_ERR_NO_DEVICE;
  const size_t nab = 32 * (size_t)d * sizeof(float), no = 32 * 32 * sizeof(float);
  float *dA = nullptr, *dB = nullptr, *dO = nullptr, *dAq = nullptr, *dBq = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&dA), nab);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dB), nab);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dO), no);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dAq), nab);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&dBq), nab);
  if (e == hipSuccess) e = hipMemcpy(dA, A, nab, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, nab, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    void* args[] = {&dA, &dB, &dO, &dAq, &dBq};
    e = hipLaunchKernel(fn, dim3(1), dim3(64), args, 0, nullptr);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dO, no, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(Aq, dAq, nab, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(Bq, dBq, nab, hipMemcpyDeviceToHost);
  for (float* p : {dA, dB, dO, dAq, dBq})
    if (p) (void)hipFree(p);
  return e == hipSuccess ? KMC_OK : KMC_ERR_HIP;
}
