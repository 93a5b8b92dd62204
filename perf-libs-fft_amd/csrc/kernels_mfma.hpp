// kernels_mfma.hpp -- 4096-point C2C with 16-bit storage whose first
// radix-16 stage is a matrix product on the matrix cores.
//
// Stage 0 of the 16 x 16 x 16 Stockham chain computes, for each of the 256
// columns t, Y_t[k] = sum_m W_16^{k m} x[t + 256 m]. In real form that is
// one 32 x 32 matrix times a 32 x 256 matrix:
//   [Yr; Yi] = [[Wr, -Wi], [Wi, Wr]] [Xr; Xi],
// computed by v_mfma_f32_32x32x16_f16 on 32-column blocks (K = 16 real
// parts, then 16 imaginary parts). The inputs are the stored f16 values
// themselves (exact); the DFT matrix is split into f16 high and low parts
// (w = hi + lo, |w - hi - lo| <= 2^-22), so each block takes four products
// accumulated in f32 and the stage is as accurate as the f32 butterflies.
// Stages 1 and 2 run on the split (real, then imaginary) LDS exchange with
// f32 butterflies, as in the plain kernel.
//
// Operand layouts of v_mfma_f32_32x32x16_f16 (wave of 64 lanes, lane l,
// element i): A[row = l % 32][k = 8 (l / 32) + i] (8 halves), B[k = 8 (l / 32)
// + i][col = l % 32] (8 halves), D[row = 8 (i / 4) + 4 (l / 32) + i % 4]
// [col = l % 32] (16 floats).
#pragma once

#include "engine.hpp"
#include "kernels.hpp"

namespace plf {

typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef float mf_f16 __attribute__((ext_vector_type(16)));
typedef unsigned mf_u4 __attribute__((ext_vector_type(4)));

// cos(2 pi q / 16), q = 0..15
__device__ __forceinline__ float mf_cos16(int q) {
  constexpr float c1 = 0.923879532511286756f, c2 = 0.707106781186547524f,
                  c3 = 0.382683432365089772f;
  switch (q & 15) {
    case 0: return 1.0f;
    case 1: case 15: return c1;
    case 2: case 14: return c2;
    case 3: case 13: return c3;
    case 4: case 12: return 0.0f;
    case 5: case 11: return -c3;
    case 6: case 10: return -c2;
    case 7: case 9: return -c1;
    default: return -1.0f;
  }
}

// A operand of this lane for K half KH (0: inputs' real parts, 1: imaginary
// parts), as high and low f16 parts. W_16^{km} = cos(a) - i sin(a), a = 2 pi k m / 16.
__device__ __forceinline__ void mf_dft16_a(int lane, int kh, mf_h8 &hi, mf_h8 &lo) {
  const int row = lane % 32, h = lane / 32;
  const int k = row % 16;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int m = 8 * h + i;
    const int q = (k * m) & 15;
    const float c = mf_cos16(q), s = mf_cos16(q + 12);  // s = sin(a)
    // rows 0..15: [Wr, -Wi] = [c, s]; rows 16..31: [Wi, Wr] = [-s, c]
    const float w = row < 16 ? (kh == 0 ? c : s) : (kh == 0 ? -s : c);
    const _Float16 wh = (_Float16)w;
    hi[i] = wh;
    lo[i] = (_Float16)(w - (float)wh);
  }
}

// KArgs: in, out (half_pair), tw = Stockham twiddles of the chain 16,16,16,
// count, idist, odist (unit element stride). 256 threads, one member per
// block.
template <bool BWD>
__global__ void __launch_bounds__(256) k_c2c_mfma16(KArgs a) {
  using ST = StorageT<half_pair>;
  using E = StaticEngineSplit<float, 256, 5, Chain<16, 16, 16>>;
  const int t = threadIdx.x;
  const int lane = t % 64, wave = t / 64, h = lane / 32, c = lane % 32;
  const int64_t member = blockIdx.x;
  if (member >= a.count) return;  // whole block: the grid is count blocks
  const uint32_t *in = reinterpret_cast<const uint32_t *>(a.in) + member * a.idist;
  half_pair *out = static_cast<half_pair *>(a.out) + member * a.odist;
  const cpx<float> *tw = static_cast<const cpx<float> *>(a.tw);
  float *lds = reinterpret_cast<float *>(dyn_lds<float>());

  mf_h8 a0h, a0l, a1h, a1l;
  mf_dft16_a(lane, 0, a0h, a0l);
  mf_dft16_a(lane, 1, a1h, a1l);

  mf_f16 acc[2];
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int col = 64 * wave + 32 * cb + c;
    mf_h8 bre, bim;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t u = __builtin_nontemporal_load(in + col + 256 * (8 * h + i));
      const uint16_t lo16 = (uint16_t)(u & 0xffffu), hi16 = (uint16_t)(u >> 16);
      // backward = swap o forward o swap
      bre[i] = __builtin_bit_cast(_Float16, BWD ? hi16 : lo16);
      bim[i] = __builtin_bit_cast(_Float16, BWD ? lo16 : hi16);
    }
    mf_f16 d = {};
    d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0h, bre, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0l, bre, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1h, bim, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1l, bim, d, 0, 0, 0);
    acc[cb] = d;
  }

  // stage 0's outputs: bin k of column t at t * 16 + k (Stockham, L = 1);
  // element i of the result holds bin 8 (i / 4) + 4 h + i % 4 (real part for
  // i < 8, imaginary part at i + 8)
  cpx<float> w[1][16];
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int col = 64 * wave + 32 * cb + c;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = 8 * (i / 4) + 4 * h + (i % 4);
        lds[lds_pad<5>(16 * col + k)] = acc[cb][i + 8 * comp];
      }
    }
    __syncthreads();
    if (comp == 0) E::template get<1, 0>(lds, w, t);
    else E::template get<1, 1>(lds, w, t);
    __syncthreads();
  }
  auto store = [&](int e, cpx<float> v) { ST::stnt(out + e, BWD ? cswap(v) : v); };
  E::template from<1>(w, lds, tw, t, true, store);
}

// Batched 16-point C2C with 16-bit storage as one matrix product per 32
// packed members (distance 16, unit stride): lane l (member c = l % 32 of
// the group, half h = l / 32) loads points 8h..8h+7 of its member (32
// contiguous bytes), which are exactly its B operand (real parts, then
// imaginary parts); four products (DFT matrix high and low parts, real and
// imaginary K halves) give D, whose element i < 8 is bin
// 8 (i / 4) + 4 h + i % 4 (real part; imaginary part at i + 8), so each lane
// stores two runs of four bins. No LDS, no exchange. Persistent launch:
// each wave walks groups of 32 members, two at a time (both loaded before
// either product).
template <bool BWD>
__device__ __forceinline__ mf_f16 mf_dft16_group(const mf_u4 (&u)[2], const mf_h8 &a0h,
                                                 const mf_h8 &a0l, const mf_h8 &a1h,
                                                 const mf_h8 &a1l) {
  mf_h8 bre, bim;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t w = u[i / 4][i % 4];
    const uint16_t lo16 = (uint16_t)(w & 0xffffu), hi16 = (uint16_t)(w >> 16);
    bre[i] = __builtin_bit_cast(_Float16, BWD ? hi16 : lo16);
    bim[i] = __builtin_bit_cast(_Float16, BWD ? lo16 : hi16);
  }
  mf_f16 d = {};
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0h, bre, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0l, bre, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1h, bim, d, 0, 0, 0);
  d = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1l, bim, d, 0, 0, 0);
  return d;
}

template <bool BWD>
__device__ __forceinline__ void mf_store16(uint32_t *dst, const mf_f16 &d) {
#pragma unroll
  for (int r = 0; r < 2; ++r) {  // bins 8r + 4h .. 8r + 4h + 3
    mf_u4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint16_t re = f2h(d[4 * r + j]), im = f2h(d[4 * r + j + 8]);
      v[j] = BWD ? (uint32_t(im) | (uint32_t(re) << 16)) : (uint32_t(re) | (uint32_t(im) << 16));
    }
    __builtin_nontemporal_store(v, reinterpret_cast<mf_u4 *>(dst + 8 * r));
  }
}

// KArgs: in, out (half_pair, packed members), tw = the operand table
// (4 x 64 lanes x 8 f16: real-half high, real-half low, imaginary-half high,
// imaginary-half low), count.
template <bool BWD>
__global__ void __launch_bounds__(256) k_c2c_mfma16_small(KArgs a) {
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int h = lane / 32, c = lane % 32;
  const mf_h8 *A = static_cast<const mf_h8 *>(a.tw);
  const mf_h8 a0h = A[lane], a0l = A[64 + lane], a1h = A[128 + lane], a1l = A[192 + lane];
  const uint32_t *in = static_cast<const uint32_t *>(a.in);
  uint32_t *out = static_cast<uint32_t *>(a.out);
  const int64_t groups = (a.count + 31) / 32;
  const int64_t stride = (int64_t)gridDim.x * 4;
  // two groups in flight per wave: both loads are issued before either
  // group's products
  for (int64_t g = (int64_t)blockIdx.x * 4 + wave; g < groups; g += 2 * stride) {
    const int64_t m0 = g * 32 + c, m1 = (g + stride) * 32 + c;
    const bool v0 = m0 < a.count, v1 = m1 < a.count;
    mf_u4 u0[2], u1[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      u0[r] = v0 ? __builtin_nontemporal_load(reinterpret_cast<const mf_u4 *>(in + m0 * 16 + 8 * h) + r)
                 : mf_u4{0u, 0u, 0u, 0u};
      u1[r] = v1 ? __builtin_nontemporal_load(reinterpret_cast<const mf_u4 *>(in + m1 * 16 + 8 * h) + r)
                 : mf_u4{0u, 0u, 0u, 0u};
    }
    const mf_f16 d0 = mf_dft16_group<BWD>(u0, a0h, a0l, a1h, a1l);
    const mf_f16 d1 = mf_dft16_group<BWD>(u1, a0h, a0l, a1h, a1l);
    if (v0) mf_store16<BWD>(out + m0 * 16 + 4 * h, d0);
    if (v1) mf_store16<BWD>(out + m1 * 16 + 4 * h, d1);
  }
}

}  // namespace plf
