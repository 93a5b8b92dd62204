// launch.hpp -- launch wrappers and registration helpers for kernels.hpp.
#pragma once

#include <algorithm>
#include <initializer_list>

#include "kernels.hpp"
#include "registry.hpp"

namespace plf {

template <void (*K)(KArgs), int BLOCK>
hipError_t launch_k(const KArgs &a, int64_t grid, size_t lds, hipStream_t s) {
  if (lds > 65536) {
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(K), hipFuncAttributeMaxDynamicSharedMemorySize,
        160 * 1024);
    if (attr != hipSuccess) return attr;
  }
  hipLaunchKernelGGL(K, dim3((unsigned)grid), dim3(BLOCK), lds, s, a);
  return hipGetLastError();
}

// Persistent launch: at most `per_cu` blocks per compute unit.
template <void (*K)(KArgs), int BLOCK, int PER_CU>
hipError_t launch_pipe(const KArgs &a, int64_t grid, size_t lds, hipStream_t s) {
  static const int cus = [] {
    int d = 0, c = 0;
    if (hipGetDevice(&d) != hipSuccess ||
        hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess)
      return 256;
    return c > 0 ? c : 256;
  }();
  return launch_k<K, BLOCK>(a, std::min<int64_t>(grid, (int64_t)cus * PER_CU), lds, s);
}

inline void add_entry1(std::vector<KEntry> &r, int kind, int prec, int n,
                       std::initializer_list<int> chain, int tpt, int tpb, int pts,
                       bool uses_lds, KLaunch fwd) {
  KEntry e;
  e.kind = kind;
  e.prec = prec;
  e.n = n;
  e.chain.assign(chain.begin(), chain.end());
  e.tpt = tpt;
  e.tpb = tpb;
  e.pts = pts;
  e.uses_lds = uses_lds;
  e.bwd = false;
  e.launch = fwd;
  r.push_back(e);
}

inline void add_entry(std::vector<KEntry> &r, int kind, int prec, int n,
                      std::initializer_list<int> chain, int tpt, int tpb, int pts,
                      bool uses_lds, KLaunch fwd, KLaunch bwd) {
  KEntry e;
  e.kind = kind;
  e.prec = prec;
  e.n = n;
  e.chain.assign(chain.begin(), chain.end());
  e.tpt = tpt;
  e.tpb = tpb;
  e.pts = pts;
  e.uses_lds = uses_lds;
  e.bwd = false;
  e.launch = fwd;
  r.push_back(e);
  e.bwd = true;
  e.launch = bwd;
  r.push_back(e);
}

}  // namespace plf

// Static chain: KERNEL in {k_c2c, k_col, k_row}, KK kind, S storage.
#define PLF_STATIC(KERNEL, KK, S, PREC, TPT, TPB, ...)                                  \
  ::plf::add_entry(                                                                    \
      r, KK, PREC, ::plf::Chain<__VA_ARGS__>::N, {__VA_ARGS__}, TPT, TPB, 0,           \
      (KK != ::plf::KK_C2C) || (::plf::Chain<__VA_ARGS__>::count > 1),                 \
      &::plf::launch_k<                                                                \
          &::plf::KERNEL<S, false, TPB,                                                \
                         ::plf::SEng<typename ::plf::StorageT<S>::C, TPT, __VA_ARGS__>>, \
          TPT * TPB>,                                                                  \
      &::plf::launch_k<                                                                \
          &::plf::KERNEL<S, true, TPB,                                                 \
                         ::plf::SEng<typename ::plf::StorageT<S>::C, TPT, __VA_ARGS__>>, \
          TPT * TPB>)

// Runtime chain (dynamic engine).
#define PLF_DYNAMIC(KERNEL, KK, S, PREC, TPT, TPB, PTS)                                     \
  ::plf::add_entry(                                                                        \
      r, KK, PREC, 0, {}, TPT, TPB, PTS, true,                                             \
      &::plf::launch_k<                                                                    \
          &::plf::KERNEL<S, false, TPB, ::plf::DEng<typename ::plf::StorageT<S>::C, TPT, PTS, TPB>>, \
          TPT * TPB>,                                                                      \
      &::plf::launch_k<                                                                    \
          &::plf::KERNEL<S, true, TPB, ::plf::DEng<typename ::plf::StorageT<S>::C, TPT, PTS, TPB>>, \
          TPT * TPB>)

// Single-direction kinds (real transforms): one entry, bwd == false.
#define PLF_STATIC1(KERNEL, KK, S, PREC, TPT, TPB, ...)                                 \
  ::plf::add_entry1(                                                                   \
      r, KK, PREC, ::plf::Chain<__VA_ARGS__>::N, {__VA_ARGS__}, TPT, TPB, 0, true,     \
      &::plf::launch_k<                                                                \
          &::plf::KERNEL<S, false, TPB,                                                \
                         ::plf::SEng<typename ::plf::StorageT<S>::C, TPT, __VA_ARGS__>>, \
          TPT * TPB>)

#define PLF_DYNAMIC1(KERNEL, KK, S, PREC, TPT, TPB, PTS)                                    \
  ::plf::add_entry1(                                                                       \
      r, KK, PREC, 0, {}, TPT, TPB, PTS, true,                                             \
      &::plf::launch_k<                                                                    \
          &::plf::KERNEL<S, false, TPB, ::plf::DEng<typename ::plf::StorageT<S>::C, TPT, PTS, TPB>>, \
          TPT * TPB>)

// Split-exchange unit-stride C2C with at least WAVES waves per SIMD.
#define PLF_UNIT_SPLIT_OCC(S, PREC, TPT, TPB, PAD, WAVES, ...)                                \
  do {                                                                                    \
    using EngT = ::plf::SEngS<typename ::plf::StorageT<S>::C, TPT, PAD, __VA_ARGS__>;      \
    ::plf::add_entry(                                                                     \
        r, ::plf::KK_C2C_UNIT, PREC, ::plf::Chain<__VA_ARGS__>::N, {__VA_ARGS__}, TPT, TPB, 0, \
        ::plf::Chain<__VA_ARGS__>::count > 1,                                             \
        &::plf::launch_k<&::plf::k_c2c_unit_occ<S, false, TPB, EngT, WAVES>, TPT * TPB>,  \
        &::plf::launch_k<&::plf::k_c2c_unit_occ<S, true, TPB, EngT, WAVES>, TPT * TPB>);  \
    for (int i = 1; i <= 2; ++i) {                                                        \
      r[r.size() - i].pad = PAD;                                                          \
      r[r.size() - i].lds_scalars = 1;                                                    \
      r[r.size() - i].waves = WAVES;                                                      \
    }                                                                                     \
  } while (0)

// Single-direction kind on the split (real, then imaginary) LDS exchange.
#define PLF_STATIC1_SPLIT(KERNEL, KK, S, PREC, TPT, TPB, PAD, ...)                         \
  do {                                                                                    \
    using EngT = ::plf::SEngS<typename ::plf::StorageT<S>::C, TPT, PAD, __VA_ARGS__>;      \
    ::plf::add_entry1(r, KK, PREC, ::plf::Chain<__VA_ARGS__>::N, {__VA_ARGS__}, TPT, TPB, 0, \
                      true, &::plf::launch_k<&::plf::KERNEL<S, false, TPB, EngT>, TPT * TPB>); \
    r.back().pad = PAD;                                                                   \
    r.back().lds_scalars = 1;                                                             \
  } while (0)

// Unit-stride streaming C2C with an explicit LDS padding shift and
// butterfly-pairing choice.
#define PLF_UNIT(S, PREC, TPT, TPB, PAD, PAIR, ...)                                       \
  do {                                                                                    \
    using EngT = ::plf::SEngX<typename ::plf::StorageT<S>::C, TPT, PAD, PAIR, __VA_ARGS__>; \
    ::plf::add_entry(r, ::plf::KK_C2C_UNIT, PREC, ::plf::Chain<__VA_ARGS__>::N,              \
                     {__VA_ARGS__}, TPT, TPB, PAIR ? 1 : 0,                               \
                     ::plf::Chain<__VA_ARGS__>::count > 1,                                \
                     &::plf::launch_k<&::plf::k_c2c_unit<S, false, TPB, EngT>, TPT * TPB>, \
                     &::plf::launch_k<&::plf::k_c2c_unit<S, true, TPB, EngT>, TPT * TPB>); \
    r[r.size() - 1].pad = PAD;                                                            \
    r[r.size() - 2].pad = PAD;                                                            \
  } while (0)

// Cross-lane C2C of n = 16, 32, 64 (k_c2c_xlane): 256-thread blocks of four
// waves, each wave 1024 / n packed members; the chain names the radix-2
// stages.
#define PLF_XLANE(S, PREC, N, ...)                                                         \
  do {                                                                                     \
    ::plf::add_entry(r, ::plf::KK_C2C_UNIT, PREC, N, {__VA_ARGS__}, N / 16, 4096 / N, 0,  \
                     false, &::plf::launch_k<&::plf::k_c2c_xlane<S, false, N>, 256>,       \
                     &::plf::launch_k<&::plf::k_c2c_xlane<S, true, N>, 256>);             \
    r[r.size() - 1].xlane = true;                                                          \
    r[r.size() - 2].xlane = true;                                                          \
  } while (0)

// Packed C2C of a small odd prime N staged through LDS (k_c2c_staged):
// 256 members of one thread each per block.
#define PLF_STAGED(S, PREC, N)                                                             \
  do {                                                                                     \
    ::plf::add_entry(r, ::plf::KK_C2C_UNIT, PREC, N, {N}, 1, 256, 0, true,                 \
                     &::plf::launch_k<&::plf::k_c2c_staged<S, false, N>, 256>,            \
                     &::plf::launch_k<&::plf::k_c2c_staged<S, true, N>, 256>);            \
    for (int i_ = 1; i_ <= 2; ++i_) {                                                      \
      r[r.size() - i_].packed = true;                                                      \
      r[r.size() - i_].staged = true;                                                      \
    }                                                                                      \
  } while (0)

// Unit-stride software-pipelined C2C (k_c2c_pipe), PER_CU blocks per CU.
#define PLF_PIPE(S, PREC, TPT, PAD, PER_CU, ...)                                           \
  do {                                                                                    \
    using EngT = ::plf::SEngX<typename ::plf::StorageT<S>::C, TPT, PAD, false, __VA_ARGS__>; \
    ::plf::add_entry(r, ::plf::KK_C2C_UNIT, PREC, ::plf::Chain<__VA_ARGS__>::N,              \
                     {__VA_ARGS__}, TPT, 1, 0, true,                                      \
                     &::plf::launch_pipe<&::plf::k_c2c_pipe<S, false, 1, EngT>, TPT, PER_CU>, \
                     &::plf::launch_pipe<&::plf::k_c2c_pipe<S, true, 1, EngT>, TPT, PER_CU>); \
    for (int i = 1; i <= 2; ++i) {                                                        \
      r[r.size() - i].pad = PAD;                                                          \
      r[r.size() - i].persist = PER_CU;                                                   \
    }                                                                                     \
  } while (0)

// Unit-stride C2C on the split-exchange engine (half the LDS).
#define PLF_UNIT_SPLIT(S, PREC, TPT, TPB, PAD, ...)                                       \
  do {                                                                                    \
    using EngT = ::plf::SEngS<typename ::plf::StorageT<S>::C, TPT, PAD, __VA_ARGS__>;      \
    ::plf::add_entry(r, ::plf::KK_C2C_UNIT, PREC, ::plf::Chain<__VA_ARGS__>::N,              \
                     {__VA_ARGS__}, TPT, TPB, 0, ::plf::Chain<__VA_ARGS__>::count > 1,     \
                     &::plf::launch_k<&::plf::k_c2c_unit<S, false, TPB, EngT>, TPT * TPB>, \
                     &::plf::launch_k<&::plf::k_c2c_unit<S, true, TPB, EngT>, TPT * TPB>); \
    for (int i = 1; i <= 2; ++i) {                                                        \
      r[r.size() - i].pad = PAD;                                                          \
      r[r.size() - i].lds_scalars = 1;                                                    \
    }                                                                                     \
  } while (0)
