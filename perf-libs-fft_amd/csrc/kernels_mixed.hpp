// kernels_mixed.hpp -- one persistent launch for a batch of mixed lengths.
//
// Every member is transformed by exactly one block of 256 threads, from its
// first read to its last write; blocks never wait for each other, so the
// launch needs no atomics, counters or inter-block synchronization. The host
// sorts the members by predicted cost (mixed_plan.cpp) and block b takes the
// members at positions b, 2G-1-b, 2G+b, 4G-1-b, ... of that order (G blocks),
// which balances the blocks to within about one member.
//
// A member runs one of four schedules:
//   direct      the whole member in LDS (runtime-chain engine, 256 threads)
//   Bluestein   chirp, DFT_M, pointwise B, DFT_M, chirp, all in LDS
//   four-step   column tiles (16 or 8 columns per pass through LDS, twiddled
//               into the workspace), then row tiles with a transposed store
//   chirp-z     Bluestein whose two DFT_M are four-step passes through the
//               workspace (chirp-in columns, pointwise rows into a natural-
//               order staging half, columns, chirp-out rows)
// The passes of one member are separated by block barriers, which also make
// the block's workspace writes visible to its own later reads.
#pragma once

#include "engine.hpp"
#include "mixed_args.hpp"

namespace plf {

template <typename S>
struct MixedExec {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename ST::C;
  using EngD = DynamicEngine<C, 256, 16, true>;   // one member per block
  using Eng16 = DynamicEngine<C, 16, 16, true>;   // 16 lanes x 16 threads
  using Eng8 = DynamicEngine<C, 32, 16, true>;    // 8 lanes x 32 threads
  static constexpr int PAD = 5;

  // pass kinds of the tile functions
  enum Pass : int { COL = 0, ROW = 1, CZ_COL1 = 2, CZ_ROW1 = 3, CZ_COL2 = 4, CZ_ROW2 = 5 };

  // backward = swap o forward o swap (swap exchanges re and im)
  static __device__ __forceinline__ cpx<C> io(const MixedArgs &a, cpx<C> v) {
    return a.bwd ? cswap(v) : v;
  }

  static __device__ __forceinline__ cpx<C> split_tw(const cpx<C> *aux, const MixedXf &x,
                                                    int64_t e) {
    const cpx<C> *lo = aux + x.tw4_off;
    const cpx<C> *hi = lo + (int64_t(1) << x.shift);
    return cmul(hi[e >> x.shift], lo[e & ((int64_t(1) << x.shift) - 1)]);
  }

  // direct (kind 0) or Bluestein (kind 2) member in one pass through LDS
  static __device__ __forceinline__ void one_block(const MixedArgs &a, const MixedXf &x,
                                                   cpx<C> *lds) {
    const E *in = static_cast<const E *>(a.in) + x.in_off;
    E *out = static_cast<E *>(a.out) + x.out_off;
    const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw) + a.desc_tw[x.d0];
    const GenericDesc *d = a.descs + x.d0;
    const int t = threadIdx.x;
    if (x.kind == 0) {
      EngD::template run<false>(
          d, lds, tw, t, true, [&](int e) { return io(a, ST::ld(in[e])); },
          [&](int e, cpx<C> v) { out[e] = ST::st(io(a, v)); });
      return;
    }
    const cpx<C> *aux = static_cast<const cpx<C> *>(a.aux);
    const cpx<C> *chirp = aux + x.chirp_off;
    const cpx<C> *B = aux + x.B_off;
    const int n = x.n, M = x.len;
    EngD::template run<true>(
        d, lds, tw, t, true,
        [&](int e) {
          if (e >= n) return cpx<C>{C(0), C(0)};
          return cmul(io(a, ST::ld(in[e])), chirp[e]);
        },
        [&](int, cpx<C>) {});
    for (int e = t; e < M; e += 256) {
      const int p = lds_pad<PAD>(e);
      lds[p] = cswap(cmul(lds[p], B[e]));
    }
    __syncthreads();
    EngD::template run<false, true>(
        d, lds, tw, t, true, [&](int) { return cpx<C>{C(0), C(0)}; },
        [&](int e, cpx<C> v) {
          if (e < n) out[e] = ST::st(io(a, cmul(cswap(v), chirp[e])));
        });
  }

  // column pass over columns [first, first + LANES) of the n1 x n2 matrix
  template <class EngT, int LANES>
  static __device__ __forceinline__ void col_tile(const MixedArgs &a, int pass, int first,
                                                  const MixedXf &x, cpx<C> *lds_all) {
    const int g = threadIdx.x % LANES;
    const int t = threadIdx.x / LANES;
    const int n1 = x.n1, n2 = x.n2;
    const int col = first + g;
    const bool valid = col < n2;
    cpx<C> *lds = lds_all + g * lds_padded<PAD>(n1);
    const GenericDesc *d = a.descs + x.d0;
    const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw) + a.desc_tw[x.d0];
    const cpx<C> *aux = static_cast<const cpx<C> *>(a.aux);
    cpx<C> *ws = static_cast<cpx<C> *>(a.ws) + x.ws_off;
    const E *in = static_cast<const E *>(a.in) + x.in_off;
    const cpx<C> *chirp = aux + x.chirp_off;
    const int64_t M = (int64_t)n1 * n2;
    EngT::template run<false>(
        d, lds, tw, t, valid,
        [&](int e) {
          const int64_t j = col + (int64_t)e * n2;
          if (pass == COL) return io(a, ST::ld(in[j]));
          if (pass == CZ_COL1) {
            if (j >= x.n) return cpx<C>{C(0), C(0)};
            return cmul(io(a, ST::ld(in[j])), chirp[j]);
          }
          return ws[M + j];  // CZ_COL2: staged natural-order data
        },
        [&](int k1, cpx<C> v) {
          ws[(int64_t)k1 * n2 + col] = cmul(v, split_tw(aux, x, (int64_t)col * k1));
        });
  }

  // row pass over rows [first, first + LANES), transposed store through LDS
  template <class EngT, int LANES>
  static __device__ __forceinline__ void row_tile(const MixedArgs &a, int pass, int first,
                                                  const MixedXf &x, cpx<C> *lds_all) {
    constexpr int TPT = 256 / LANES;
    const int g = threadIdx.x / TPT;
    const int t = threadIdx.x % TPT;
    const int n1 = x.n1, n2 = x.n2;
    const int row = first + g;
    const bool valid = row < n1;
    const int lper = lds_padded<PAD>(n2);
    cpx<C> *lds = lds_all + g * lper;
    const GenericDesc *d = a.descs + x.d1;
    const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw) + a.desc_tw[x.d1];
    const cpx<C> *aux = static_cast<const cpx<C> *>(a.aux);
    cpx<C> *ws = static_cast<cpx<C> *>(a.ws) + x.ws_off;
    const cpx<C> *src = ws + (int64_t)(valid ? row : 0) * n2;
    EngT::template run<true>(
        d, lds, tw, t, valid, [&](int e) { return src[e]; }, [&](int, cpx<C>) {});
    const int c = threadIdx.x % LANES;
    const int rc = first + c;
    if (rc >= n1) return;
    const cpx<C> *l = lds_all + c * lper;
    const int64_t M = (int64_t)n1 * n2;
    E *out = static_cast<E *>(a.out) + x.out_off;
    const cpx<C> *B = aux + x.B_off;
    const cpx<C> *chirp = aux + x.chirp_off;
    for (int k2 = threadIdx.x / LANES; k2 < n2; k2 += TPT) {
      const cpx<C> v = l[lds_pad<PAD>(k2)];
      const int64_t k = rc + (int64_t)n1 * k2;
      if (pass == ROW) {
        out[k] = ST::st(io(a, v));
      } else if (pass == CZ_ROW1) {
        ws[M + k] = cswap(cmul(v, B[k]));
      } else if (k < x.n) {  // CZ_ROW2
        out[k] = ST::st(io(a, cmul(cswap(v), chirp[k])));
      }
    }
  }

  // All column tiles (or all row tiles) of one member. Each tile ends with a
  // barrier: the next tile rewrites the LDS this one read, and a row pass
  // reads what the column pass before it wrote.
  static __device__ __forceinline__ void cols(const MixedArgs &a, int pass, const MixedXf &x,
                                              cpx<C> *lds) {
    for (int c = 0; c < x.n2; c += x.lanes0) {
      if (x.lanes0 == 16) col_tile<Eng16, 16>(a, pass, c, x, lds);
      else col_tile<Eng8, 8>(a, pass, c, x, lds);
      __syncthreads();
    }
  }
  static __device__ __forceinline__ void rows(const MixedArgs &a, int pass, const MixedXf &x,
                                              cpx<C> *lds) {
    for (int r = 0; r < x.n1; r += x.lanes1) {
      if (x.lanes1 == 16) row_tile<Eng16, 16>(a, pass, r, x, lds);
      else row_tile<Eng8, 8>(a, pass, r, x, lds);
      __syncthreads();
    }
  }

  static __device__ __forceinline__ void run(const MixedArgs &a, cpx<C> *lds) {
    const int G = gridDim.x, b = blockIdx.x;
    for (int r = 0;; ++r) {
      const int64_t p = (int64_t)r * G + ((r & 1) ? G - 1 - b : b);
      if (p >= a.count) break;
      const MixedXf x = a.xf[a.order[p]];
      switch (x.kind) {
        case 0:
        case 2:
          one_block(a, x, lds);
          __syncthreads();
          break;
        case 1:
          cols(a, COL, x, lds);
          rows(a, ROW, x, lds);
          break;
        default:
          cols(a, CZ_COL1, x, lds);
          rows(a, CZ_ROW1, x, lds);
          cols(a, CZ_COL2, x, lds);
          rows(a, CZ_ROW2, x, lds);
          break;
      }
    }
  }
};

// At most 128 VGPRs (4 waves per SIMD): the blocks hide HBM latency by
// occupancy, and LDS (about 34 KB per block for FP32) allows 4 per CU.
template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_mixed(MixedArgs a) {
  extern __shared__ unsigned char plf_dyn_lds[];
  MixedExec<S>::run(a, reinterpret_cast<cpx<typename StorageT<S>::C> *>(plf_dyn_lds));
}

}  // namespace plf
