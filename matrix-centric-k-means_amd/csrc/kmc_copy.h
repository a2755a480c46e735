// kmc_copy.h -- host <-> device transfers of row blocks whose host and device
// row pitches differ. Each is one contiguous transfer of the device block
// plus packing or unpacking on the host, and returns once the transfer is
// done. Pitched hipMemcpy2D transfers are avoided: from pageable memory they
// go row by row, and with a host process that loads a second HIP runtime
// before this library (a framework bundling its own) device-to-host pitched
// copies have been seen returning wrong bytes under a profiler.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

namespace kmc {

// rows x width elements: host src (pitch hld) -> device dst (pitch dld);
// the dld - width padding elements of each device row are written as zero
template <typename T>
hipError_t rows_to_device(T* dst, size_t dld, const T* src, size_t hld, size_t width,
                          size_t rows, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipError_t e;
  if (dld == width && hld == width) {
    e = hipMemcpyAsync(dst, src, rows * width * sizeof(T), hipMemcpyHostToDevice, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  std::vector<T> tmp(rows * dld, T(0));
  for (size_t r = 0; r < rows; ++r) std::memcpy(&tmp[r * dld], src + r * hld, width * sizeof(T));
  e = hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice, s);
  return e == hipSuccess ? hipStreamSynchronize(s) : e;  // tmp dies here
}

// rows x width elements: device src (pitch dld) -> host dst (pitch hld)
template <typename T>
hipError_t rows_to_host(T* dst, size_t hld, const T* src, size_t dld, size_t width, size_t rows,
                        hipStream_t s) {
  if (rows == 0) return hipSuccess;
  hipError_t e;
  if (dld == width && hld == width) {
    e = hipMemcpyAsync(dst, src, rows * width * sizeof(T), hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  std::vector<T> tmp((rows - 1) * dld + width);
  e = hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(T), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  for (size_t r = 0; r < rows; ++r) std::memcpy(dst + r * hld, &tmp[r * dld], width * sizeof(T));
  return hipSuccess;
}

}  // namespace kmc
