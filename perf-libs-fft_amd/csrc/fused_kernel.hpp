// fused_kernel.hpp -- the LDS-resident Stockham kernel for batched C2C.
//
// One kernel computes whole length-N transforms (N = product of a
// compile-time radix chain). A block holds TPB transforms, each handled by
// TPT threads. Stage s of radix R with L = product of earlier radices is the
// self-sorting (Stockham) DIT step
//   v[m] = data[b + m*N/R] * W_{L*R}^{(b mod L)*m}
//   V    = DFT_R(v)
//   data'[(b div L)*L*R + (b mod L) + m*L] = V[m]
// for butterfly b in [0, N/R). Stage 0 reads straight from global memory and
// the last stage writes straight to global memory; both patterns are
// b + m*N/R, i.e. consecutive threads touch consecutive elements, so global
// traffic is coalesced and each element crosses HBM exactly once. The
// stages in between exchange through LDS (one buffer, two barriers per
// stage).
//
// This replaces the reference's recursive cooley_tukey_ab / ac_twid stages
// and their temp-buffer ping-pong (src/planner/strategies/cooley_tukey.hpp:
// 36-143, 329-400, 582-748) and the loop strategy over the batch
// (src/planner/strategies/loop.hpp:21-82) with one launch over the batch.
#pragma once

#include "fft_device.hpp"
#include "kernel_args.hpp"

namespace plf {

template <typename S, bool BWD, int TPT, int TPB, int PADSHIFT, typename Ch>
struct FusedC2C {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename ST::C;
  static constexpr int N = Ch::N;
  static constexpr int NS = Ch::count;
  static constexpr int LDS_PER = lds_padded<PADSHIFT>(N);
  static constexpr size_t lds_bytes = (NS > 1) ? size_t(TPB) * LDS_PER * sizeof(cpx<C>) : 0;

  template <int SI>
  static __device__ __forceinline__ void stage(const FusedArgs &a, cpx<C> *lds,
                                               const E *in, E *out, int t,
                                               bool valid,
                                               const cpx<C> *__restrict__ tw) {
    constexpr int R = Ch::radix(SI);
    constexpr int L = Ch::L(SI);
    constexpr int NB = N / R;
    constexpr int KB = (NB + TPT - 1) / TPT;
    constexpr bool FIRST = (SI == 0);
    constexpr bool LAST = (SI == NS - 1);
    cpx<C> v[KB][R];

    // ---- gather inputs
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = t + k * TPT;
      const bool act = (NB % TPT == 0) || (b < NB);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        if constexpr (FIRST) {
          cpx<C> e = {C(0), C(0)};
          if (act && valid) e = ST::ld(in[(int64_t)(b + m * NB) * a.istride]);
          v[k][m] = BWD ? cswap(e) : e;
        } else {
          v[k][m] = act ? lds[lds_pad<PADSHIFT>(b + m * NB)] : cpx<C>{C(0), C(0)};
        }
      }
    }
    // ---- twiddle + butterfly
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = t + k * TPT;
      if constexpr (!FIRST) {
        const int j = b % L;
        const bool act = (NB % TPT == 0) || (b < NB);
        if (act) {
#pragma unroll
          for (int m = 1; m < R; ++m)
            v[k][m] = cmul(v[k][m], tw[Ch::tw_offset(SI) + (m - 1) * L + j]);
        }
      }
      Dft<R, C>::run(v[k]);
    }
    if constexpr (!FIRST && !LAST) __syncthreads();  // everyone has read
    // ---- scatter outputs
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = t + k * TPT;
      const bool act = (NB % TPT == 0) || (b < NB);
      if constexpr (LAST) {
        if (act && valid) {
#pragma unroll
          for (int m = 0; m < R; ++m) {
            cpx<C> o = BWD ? cswap(v[k][m]) : v[k][m];
            out[(int64_t)(b + m * NB) * a.ostride] = ST::st(o);
          }
        }
      } else {
        if (act) {
          const int j = b % L;
          const int base = (b / L) * L * R + j;
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad<PADSHIFT>(base + m * L)] = v[k][m];
        }
      }
    }
    if constexpr (!LAST) {
      __syncthreads();
      stage<SI + 1>(a, lds, in, out, t, valid, tw);
    }
  }

  static __device__ __forceinline__ void run(const FusedArgs &a, cpx<C> *lds_all) {
    const int g = threadIdx.x / TPT;
    const int t = threadIdx.x % TPT;
    const int64_t tr = (int64_t)blockIdx.x * TPB + g;
    const bool valid = tr < a.batch;
    const E *in = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
    E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
    cpx<C> *lds = lds_all + g * LDS_PER;
    stage<0>(a, lds, in, out, t, valid, static_cast<const cpx<C> *>(a.tw));
  }
};

template <typename K>
__global__ void __launch_bounds__(K::BLOCK) fused_c2c_kernel(FusedArgs a) {
  extern __shared__ unsigned char plf_dyn_lds[];
  K::F::run(a, reinterpret_cast<cpx<typename K::F::C> *>(plf_dyn_lds));
}

// Bundles the template parameters so one kernel template serves every
// configuration.
template <typename S, bool BWD, int TPT, int TPB, int PADSHIFT, int... Rs>
struct FusedCfg {
  using F = FusedC2C<S, BWD, TPT, TPB, PADSHIFT, Chain<Rs...>>;
  static constexpr int BLOCK = TPT * TPB;
  static constexpr int TPT_ = TPT, TPB_ = TPB, PAD_ = PADSHIFT;
};

}  // namespace plf
