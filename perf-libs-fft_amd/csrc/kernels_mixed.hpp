// kernels_mixed.hpp -- the tile kernels of a mixed-length batch.
//
// Members are split into tasks of one 256-thread block each:
//   ONE       a whole member in LDS: direct (runtime-chain engine) or
//             Bluestein (chirp, DFT_M, pointwise B, DFT_M, chirp)
//   COL/ROW   four-step tiles: 16 (or 8; 64 or 256 for short columns)
//             columns per pass through LDS,
//             twiddled into the member's workspace; then 16 (or 8) rows
//             with a transposed store
// (the rows of a four-step member whose n2 has a large prime factor run on
// the compiled Bluestein kernel, kernels_conv.hpp, after these columns)
// (chirp-z members run on the three-pass kernels of kernels_cz.hpp, in
// groups). A launch runs one dependency level (ONE + COL, then ROW); stream
// order separates the levels, so blocks never wait for each other and need
// no atomics. Within a launch the host orders the tasks by predicted cost
// and block b takes positions b, 2G-1-b, 2G+b, ... (snake order over G
// blocks).
#pragma once

#include "engine.hpp"
#include "mixed_args.hpp"

namespace plf {

template <typename S>
struct MixedExec {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename ST::C;
  // 7-smooth chains only (radices 2, 3, 4, 5, 7, 8, 16)
  using EngD = DynamicEngine<C, 256, 16, true, true>;   // one member per block
  using Eng16 = DynamicEngine<C, 16, 16, true, true>;   // 16 lanes x 16 threads
  using Eng8 = DynamicEngine<C, 32, 16, true, true>;    // 8 lanes x 32 threads
  using Eng64 = DynamicEngine<C, 4, 16, true, true>;    // 64 lanes x 4 threads
  using Eng256 = DynamicEngine<C, 1, 16, true, true>;   // 256 lanes x 1 thread
  static constexpr int PAD = 5;

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

  // column pass over columns [first, end) (at most LANES) of the n1 x n2
  // matrix; the host starts tiles where the member's rows are line-aligned
  template <class EngT, int LANES>
  static __device__ __forceinline__ void col_tile(const MixedArgs &a, int first, int end,
                                                  const MixedXf &x, cpx<C> *lds_all) {
    const int g = threadIdx.x % LANES;
    const int t = threadIdx.x / LANES;
    const int n1 = x.n1, n2 = x.n2;
    const int col = first + g;
    const bool valid = col < end;
    cpx<C> *lds = lds_all + g * lds_padded<PAD>(n1);
    const GenericDesc *d = a.descs + x.d0;
    const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw) + a.desc_tw[x.d0];
    const cpx<C> *aux = static_cast<const cpx<C> *>(a.aux);
    cpx<C> *ws = static_cast<cpx<C> *>(a.ws) + x.ws_off;
    const E *in = static_cast<const E *>(a.in) + x.in_off;
    EngT::template run<false>(
        d, lds, tw, t, valid, [&](int e) { return io(a, ST::ld(in[col + (int64_t)e * n2])); },
        [&](int k1, cpx<C> v) {
          // kind 4: the rows run on the Bluestein kernel, which swaps its
          // input again for a backward transform
          const cpx<C> w = cmul(v, split_tw(aux, x, (int64_t)col * k1));
          ws[(int64_t)k1 * n2 + col] = x.kind == 4 ? io(a, w) : w;
        });
  }

  // row pass over rows [first, first + LANES), transposed store through LDS
  template <class EngT, int LANES>
  static __device__ __forceinline__ void row_tile(const MixedArgs &a, int first,
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
    const cpx<C> *ws = static_cast<const cpx<C> *>(a.ws) + x.ws_off;
    const cpx<C> *src = ws + (int64_t)(valid ? row : 0) * n2;
    EngT::template run<true>(
        d, lds, tw, t, valid, [&](int e) { return src[e]; }, [&](int, cpx<C>) {});
    const int c = threadIdx.x % LANES;
    const int rc = first + c;
    if (rc >= n1) return;
    const cpx<C> *l = lds_all + c * lper;
    E *out = static_cast<E *>(a.out) + x.out_off;
    for (int k2 = threadIdx.x / LANES; k2 < n2; k2 += TPT)
      out[rc + (int64_t)n1 * k2] = ST::st(io(a, l[lds_pad<PAD>(k2)]));
  }

  // The task types of one dependency level: 0 = ONE, COL; 1 = ROW. One
  // kernel per level keeps each kernel's code to the engine instances that
  // level runs.
  template <int LEVEL>
  static __device__ __forceinline__ void run(const MixedArgs &a, cpx<C> *lds) {
    const int G = gridDim.x, b = blockIdx.x;
    for (int r = 0;; ++r) {
      const int64_t p = (int64_t)r * G + ((r & 1) ? G - 1 - b : b);
      if (p >= a.ntasks) break;
      const MixedTask tk = a.tasks[p];
      const MixedXf x = a.xf[tk.xf];
      if constexpr (LEVEL == 0) {
        if (tk.type == MT_ONE) {
          one_block(a, x, lds);
        } else {
          if (x.lanes0 == 16) col_tile<Eng16, 16>(a, tk.first, tk.end, x, lds);
          else if (x.lanes0 == 8) col_tile<Eng8, 8>(a, tk.first, tk.end, x, lds);
          else if (x.lanes0 == 64) col_tile<Eng64, 64>(a, tk.first, tk.end, x, lds);
          else col_tile<Eng256, 256>(a, tk.first, tk.end, x, lds);
        }
      } else {  // MT_ROW
        if (x.lanes1 == 16) row_tile<Eng16, 16>(a, tk.first, x, lds);
        else row_tile<Eng8, 8>(a, tk.first, x, lds);
      }
      __syncthreads();  // the next task rewrites the LDS this one read
    }
  }
};

// At most 128 VGPRs (4 waves per SIMD): the blocks hide HBM latency by
// occupancy, and LDS (about 34 KB per block for FP32) allows 4 per CU.
template <typename S, int LEVEL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_mixed(MixedArgs a) {
  extern __shared__ unsigned char plf_dyn_lds[];
  MixedExec<S>::template run<LEVEL>(
      a, reinterpret_cast<cpx<typename StorageT<S>::C> *>(plf_dyn_lds));
}

}  // namespace plf
