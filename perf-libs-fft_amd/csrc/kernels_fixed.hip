// kernels_fixed.hip -- launchers of the fixed-point kernels (Q0.15, Q0.7).
#include "kernels_fixed.hpp"
#include "registry.hpp"

namespace plf {

namespace {

template <void (*K)(KArgs), int BLOCK>
hipError_t launch_q(const KArgs &a, int64_t grid, size_t lds, hipStream_t s) {
  if (lds > 65536) {
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(K), hipFuncAttributeMaxDynamicSharedMemorySize,
        160 * 1024);
    if (attr != hipSuccess) return attr;
  }
  hipLaunchKernelGGL(K, dim3((unsigned)grid), dim3(BLOCK), lds, s, a);
  return hipGetLastError();
}

template <int BITS, bool BWD>
hipError_t launch_dir(int tier, const KArgs &a, int64_t grid, size_t lds, hipStream_t s) {
  switch (tier) {
    case 0: return launch_q<k_fixed<BITS, 16, 16, 16, BWD>, 256>(a, grid, lds, s);
    case 1: return launch_q<k_fixed<BITS, 64, 4, 16, BWD>, 256>(a, grid, lds, s);
    case 2: return launch_q<k_fixed<BITS, 256, 1, 16, BWD>, 256>(a, grid, lds, s);
    case 3: return launch_q<k_fixed<BITS, 1024, 1, 16, BWD>, 1024>(a, grid, lds, s);
    case 4: return launch_q<k_fixed<BITS, 1024, 1, 32, BWD>, 1024>(a, grid, lds, s);
    // n = 4096 on the compile-time chain 16 x 16 x 16 (unit strides)
    case kFixedTier4096: return launch_q<k_fixed_static<BITS, 256, 1, 16, BWD, 16, 16, 16>, 256>(a, grid, lds, s);
    default: return hipErrorInvalidValue;
  }
}

// flags bit0 selects the backward instances
template <int BITS>
hipError_t launch_tier(int tier, const KArgs &a, int64_t grid, size_t lds, hipStream_t s) {
  return (a.flags & 1) ? launch_dir<BITS, true>(tier, a, grid, lds, s)
                       : launch_dir<BITS, false>(tier, a, grid, lds, s);
}

template <int BITS>
hipError_t launch_elem(int which, const KArgs &a, int64_t total, hipStream_t s) {
  const int64_t grid = (total + 255) / 256;
  switch (which) {
    case 0: return launch_q<k_fixed_widen<BITS>, 256>(a, grid, 0, s);
    case 1: return launch_q<k_fixed_half<BITS>, 256>(a, grid, 0, s);
    case 2: return launch_q<k_fixed_hermitian<BITS>, 256>(a, grid, 0, s);
    case 3: return launch_q<k_fixed_real<BITS>, 256>(a, grid, 0, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

const FixedTier kFixedTiers[kNumFixedTiers] = {
    {16, 16, 16, 256}, {64, 4, 16, 1024}, {256, 1, 16, 4096},
    {1024, 1, 16, 16384}, {1024, 1, 32, 32768}};

hipError_t launch_fixed(int bits, int tier, const KArgs &a, int64_t grid, size_t lds,
                        hipStream_t s) {
  return bits == 15 ? launch_tier<15>(tier, a, grid, lds, s) : launch_tier<7>(tier, a, grid, lds, s);
}

hipError_t launch_fixed_elem(int bits, int which, const KArgs &a, int64_t total, hipStream_t s) {
  return bits == 15 ? launch_elem<15>(which, a, total, s) : launch_elem<7>(which, a, total, s);
}

}  // namespace plf
