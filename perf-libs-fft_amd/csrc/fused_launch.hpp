// fused_launch.hpp -- launch wrappers and registration for fused kernels.
#pragma once

#include <initializer_list>

#include "fused_kernel.hpp"
#include "registry.hpp"

namespace plf {

template <typename K>
hipError_t launch_fused(const FusedArgs &a, int64_t grid, hipStream_t s) {
  constexpr size_t lds = K::F::lds_bytes;
  if constexpr (lds > 65536) {
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&fused_c2c_kernel<K>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (attr != hipSuccess) return attr;
  }
  hipLaunchKernelGGL(fused_c2c_kernel<K>, dim3((unsigned)grid), dim3(K::BLOCK),
                     lds, s, a);
  return hipGetLastError();
}

template <typename KF, typename KB>
void add_fused(std::vector<FusedEntry> &r, int prec, std::initializer_list<int> chain) {
  FusedEntry e;
  e.n = KF::F::N;
  e.prec = prec;
  e.tpt = KF::TPT_;
  e.tpb = KF::TPB_;
  e.padshift = KF::PAD_;
  e.chain.assign(chain.begin(), chain.end());
  e.lds_bytes = KF::F::lds_bytes;
  e.block = KF::BLOCK;
  e.bwd = false;
  e.launch = &launch_fused<KF>;
  r.push_back(e);
  e.bwd = true;
  e.launch = &launch_fused<KB>;
  r.push_back(e);
}

}  // namespace plf

// Registers forward and backward variants of one configuration.
#define PLF_FUSED(S, PREC, TPT, TPB, PAD, ...)                                \
  ::plf::add_fused<::plf::FusedCfg<S, false, TPT, TPB, PAD, __VA_ARGS__>,     \
                   ::plf::FusedCfg<S, true, TPT, TPB, PAD, __VA_ARGS__>>(     \
      r, PREC, {__VA_ARGS__})
