// kernels_mixed_impl.hpp -- launch and occupancy wrappers of k_mixed<S, L>,
// included by one translation unit per storage precision.
#pragma once

#include "kernels_mixed.hpp"

namespace plf {

template <typename S, int L>
hipError_t launch_mixed_level(const MixedArgs &a, int grid, size_t lds, hipStream_t s) {
  if (lds > 65536) {
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_mixed<S, L>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (attr != hipSuccess) return attr;
  }
  hipLaunchKernelGGL((k_mixed<S, L>), dim3((unsigned)grid), dim3(256), lds, s, a);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_mixed_t(int level, const MixedArgs &a, int grid, size_t lds, hipStream_t s) {
  switch (level) {
    case 0: return launch_mixed_level<S, 0>(a, grid, lds, s);
    case 1: return launch_mixed_level<S, 1>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename S, int L>
int mixed_blocks_level(size_t lds) {
  if (lds > 65536) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mixed<S, L>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024) != hipSuccess)
      return 0;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_mixed<S, L>, 256, lds) != hipSuccess)
    return 0;
  return nb;
}

template <typename S>
int mixed_blocks_per_cu_t(int level, size_t lds) {
  switch (level) {
    case 0: return mixed_blocks_level<S, 0>(lds);
    case 1: return mixed_blocks_level<S, 1>(lds);
    default: return 0;
  }
}

}  // namespace plf
