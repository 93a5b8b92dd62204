// kernels_xlane.hpp -- batched C2C of n = 16, 32 or 64 points with the
// butterflies across lanes (DPP and ds_swizzle) instead of an LDS exchange.
//
// Layout: the members are contiguous (unit stride, distance n). Lane l of a
// wave holds, in register i (0..15), element l + 64 i of the wave's 1024
// consecutive complex elements: point j = l mod n of member 64 i / n + l / n.
// Every load and store instruction of the wave covers 512 contiguous bytes.
// (Holding both halves of a member in one lane, so that the first stage runs
// in registers, measured slower: 0.47 instead of 0.36 ms for 2^27 points at
// n = 32, from the shorter contiguous runs.)
//
// Transform: radix-2 decimation in frequency over the lanes of a member,
// stages h = n/2, n/4, ..., 1. Lane j reads its partner j ^ h and keeps
//   v' = (v + p)                         if bit h of j is clear,
//   v' = (p - v) W_{2h}^{j mod h}        if it is set,
// as v' = (s v + p) w with the lane constants s = +-1 and w (w = 1 on the
// clear lanes and on the last stage). The result is the DFT in bit-reversed
// order, so lane j stores bin bitrev(j) of its member: a permutation inside
// the member's block, which keeps the stores contiguous.
//
// Partners: DPP quad_perm for h = 1, 2; DPP row_ror:8 for h = 8 (within a
// row of 16 lanes, i +- 8 mod 16 = i ^ 8); ds_swizzle in bit-mask mode for
// h = 4 and 16; v_permlane32_swap for h = 32. None of them touches LDS
// memory or needs a barrier.
#pragma once

#include "fft_device.hpp"
#include "kernel_args.hpp"
#include "kernels.hpp"

namespace plf {

template <int H>
__device__ __forceinline__ float xlane_partner(float v) {
  const int x = __builtin_bit_cast(int, v);
  int r;
  if constexpr (H == 1) {
    r = __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  } else if constexpr (H == 2) {
    r = __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  } else if constexpr (H == 8) {
    r = __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, true);  // row_ror:8
  } else if constexpr (H == 32) {
    // swapping the upper half of a copy with the lower half of another
    // leaves [lo | lo] in the first and [hi | hi] in the second
    const auto p = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false);
    r = (int)(__lane_id() < 32 ? p[1] : p[0]);
  } else {
    static_assert(H == 4 || H == 16, "partner distance");
    // bit-mask mode: and_mask 0x1f, or_mask 0, xor_mask H
    r = __builtin_amdgcn_ds_swizzle(x, 0x1F | (H << 10));
  }
  return __builtin_bit_cast(float, r);
}

template <int N>
struct XLane {
  static_assert(N == 16 || N == 32 || N == 64, "lengths 16, 32, 64");
  static constexpr int LG = N == 16 ? 4 : (N == 32 ? 5 : 6);
  static constexpr int MEMBERS_PER_WAVE = 1024 / N;

  static __device__ __forceinline__ int bitrev(int j) {
    return (int)(__builtin_bitreverse32((unsigned)j) >> (32 - LG));
  }

  // one DIF stage of half-size H on the 16 values of this lane
  template <int H>
  static __device__ __forceinline__ void stage(cpx<float> (&v)[16], int j,
                                               const cpx<float> *__restrict__ tw) {
    const float s = (j & H) ? -1.0f : 1.0f;
    cpx<float> w = {1.0f, 0.0f};
    if constexpr (H > 1) w = tw[(LG - 1 - __builtin_ctz(H)) * N + j];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float pr = xlane_partner<H>(v[i].x);
      const float pi = xlane_partner<H>(v[i].y);
      const cpx<float> d = {fmaf(s, v[i].x, pr), fmaf(s, v[i].y, pi)};
      v[i] = H > 1 ? cmul(d, w) : d;
    }
  }
};

// KArgs: in, out (packed members, distance n), tw = per-stage lane twiddles
// (stage of half-size h at row log2(n) - 1 - log2(h), n entries each), count
// (members), n. 256 threads per block = four waves of 1024 elements.
template <typename S, bool BWD, int N>
__global__ void __launch_bounds__(256) k_c2c_xlane(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  const int lane = __lane_id();
  const int wave = threadIdx.x / 64;
  const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 1024;
  const int64_t total = a.count * N;
  const E *in = static_cast<const E *>(a.in);
  E *out = static_cast<E *>(a.out);
  const cpx<float> *tw = static_cast<const cpx<float> *>(a.tw);
  const int j = lane % N;
  cpx<float> v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int64_t e = base + lane + 64 * i;
    cpx<float> x = {0.0f, 0.0f};
    if (e < total) x = ST::ldnt(in + e);
    v[i] = BWD ? cswap(x) : x;
  }
  if constexpr (N >= 64) XLane<N>::template stage<32>(v, j, tw);
  if constexpr (N >= 32) XLane<N>::template stage<16>(v, j, tw);
  XLane<N>::template stage<8>(v, j, tw);
  XLane<N>::template stage<4>(v, j, tw);
  XLane<N>::template stage<2>(v, j, tw);
  XLane<N>::template stage<1>(v, j, tw);
  const int k = XLane<N>::bitrev(j);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int64_t e = base + lane + 64 * i;
    if (e < total) ST::stnt(out + (e - j + k), BWD ? cswap(v[i]) : v[i]);
  }
}

}  // namespace plf
