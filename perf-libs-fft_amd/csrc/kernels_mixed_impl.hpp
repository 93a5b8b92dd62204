// kernels_mixed_impl.hpp -- launch and occupancy wrappers of k_mixed<S>,
// included by one translation unit per storage precision.
#pragma once

#include "kernels_mixed.hpp"

namespace plf {

template <typename S>
hipError_t launch_mixed_t(const MixedArgs &a, int grid, size_t lds, hipStream_t s) {
  if (lds > 65536) {
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_mixed<S>), hipFuncAttributeMaxDynamicSharedMemorySize,
        160 * 1024);
    if (attr != hipSuccess) return attr;
  }
  hipLaunchKernelGGL(k_mixed<S>, dim3((unsigned)grid), dim3(256), lds, s, a);
  return hipGetLastError();
}

template <typename S>
int mixed_blocks_per_cu_t(size_t lds) {
  if (lds > 65536) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mixed<S>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024) != hipSuccess)
      return 0;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_mixed<S>, 256, lds) != hipSuccess)
    return 0;
  return nb;
}

}  // namespace plf
