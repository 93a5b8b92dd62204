// engine.hpp -- the LDS Stockham engine shared by every FFT kernel.
//
// An engine computes one length-N forward DFT cooperatively with TPT
// threads (thread index t). Data enter through a load callback (stage 0
// reads element e straight from wherever the caller says: global memory,
// a chirp-multiplied input, a permuted gather, ...) and leave through a
// store callback (the last stage writes output element e anywhere: global
// memory, a twiddled workspace, ...), or through LDS when the caller needs a
// cross-element epilogue (real-to-complex post-processing, transposed
// stores). The stages in between exchange through LDS.
//
// Stage s of radix R with L = product of earlier radices is the
// self-sorting (Stockham) DIT step
//   v[m] = data[b + m*N/R] * W_{L*R}^{(b mod L)*m}
//   V    = DFT_R(v)
//   data'[(b div L)*L*R + (b mod L) + m*L] = V[m]
// for butterfly b in [0, N/R). Stage 0 reads, and the last stage writes,
// elements b + m*N/R: consecutive threads touch consecutive elements.
//
// All threads of the block must call run() (it contains barriers); threads
// without a valid transform pass valid=false and their callbacks are not
// invoked.
#pragma once

#include "fft_device.hpp"
#include "kernel_args.hpp"

namespace plf {

// ---------------------------------------------------------------------------
// compile-time radix chain
// ---------------------------------------------------------------------------
// PAIR: thread t owns the KB consecutive butterflies t*KB .. t*KB+KB-1 of a
// stage (instead of t, t+TPT, ...), so that with unit stride its global
// loads and stores of neighbouring elements merge into wider accesses.
template <typename C, int TPT, int PAD, typename Ch, bool PAIR = false>
struct StaticEngine {
  static constexpr int N = Ch::N;
  static constexpr int NS = Ch::count;
  static constexpr int LDS_PER = lds_padded<PAD>(N);
  static constexpr bool USES_LDS = NS > 1;

  template <int SI, bool TO_LDS, bool LDS_IN, class Load, class Store, bool REGS_IN = false,
            bool KEEP = false>
  static __device__ __forceinline__ void stage(cpx<C> *lds, const cpx<C> *__restrict__ tw,
                                               int t, bool valid, Load &load, Store &store,
                                               const cpx<C> *vin = nullptr,
                                               cpx<C> *keep = nullptr) {
    constexpr int R = Ch::radix(SI);
    constexpr int L = Ch::L(SI);
    constexpr int NB = N / R;
    constexpr int KB = (NB + TPT - 1) / TPT;
    constexpr bool FIRST = (SI == 0);
    constexpr bool LAST = (SI == NS - 1);
    cpx<C> v[KB][R];
    auto bidx = [&](int k) { return PAIR ? t * KB + k : t + k * TPT; };
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = bidx(k);
      const bool act = (NB % TPT == 0) || (b < NB);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        if constexpr (FIRST && REGS_IN) {
          v[k][m] = vin[k * R + m];
        } else if constexpr (FIRST && !LDS_IN) {
          v[k][m] = (act && valid) ? load(b + m * NB) : cpx<C>{C(0), C(0)};
        } else {
          v[k][m] = act ? lds[lds_pad_step<PAD, NB>(b, m)] : cpx<C>{C(0), C(0)};
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = bidx(k);
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
    if constexpr ((!FIRST || LDS_IN) && (!LAST || TO_LDS)) __syncthreads();  // all reads done
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = bidx(k);
      const bool act = (NB % TPT == 0) || (b < NB);
      if constexpr (LAST && KEEP) {
#pragma unroll
        for (int m = 0; m < R; ++m) keep[k * R + m] = v[k][m];
      } else if constexpr (LAST && !TO_LDS) {
        if (act && valid) {
#pragma unroll
          for (int m = 0; m < R; ++m) store(b + m * NB, v[k][m]);
        }
      } else if constexpr (LAST) {
        if (act) {
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad_step<PAD, NB>(b, m)] = v[k][m];
        }
      } else {
        if (act) {
          const int j = b % L;
          const int base = (b / L) * L * R + j;
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad_step<PAD, L>(base, m)] = v[k][m];
        }
      }
    }
    if constexpr (!LAST || TO_LDS) __syncthreads();
    if constexpr (!LAST)
      stage<SI + 1, TO_LDS, false, Load, Store, false, KEEP>(lds, tw, t, valid, load, store,
                                                             nullptr, keep);
  }

  // Results stay in LDS (natural order, lds[lds_pad(e)]) when TO_LDS. With
  // LDS_IN the input is taken from LDS (natural order) instead of `load`.
  template <bool TO_LDS, bool LDS_IN = false, class Load, class Store>
  static __device__ __forceinline__ void run(cpx<C> *lds, const cpx<C> *tw, int t, bool valid,
                                             Load &&load, Store &&store) {
    stage<0, TO_LDS, LDS_IN>(lds, tw, t, valid, load, store);
  }

  // Stage 0 takes its inputs from registers: vin[k * R0 + m] holds element
  // b + m * (N / R0) of butterfly b = t + k * TPT (zero where b >= N / R0).
  // Lets a caller issue the next member's loads before this member's stages,
  // or compute stage-0 inputs with knowledge of the butterfly they feed.
  // With TO_LDS the results stay in LDS as for run<true>.
  template <bool TO_LDS = false, class Store>
  static __device__ __forceinline__ void run_regs(const cpx<C> *vin, cpx<C> *lds,
                                                  const cpx<C> *tw, int t, Store &&store,
                                                  bool valid = true) {
    auto none = [](int) { return cpx<C>{C(0), C(0)}; };
    stage<0, TO_LDS, false, decltype(none), Store, true>(lds, tw, t, valid, none, store, vin);
  }

  // run_regs with the last stage's outputs left in registers:
  // keep[k * RL + m] = element b + m * (N / RL) of butterfly b = t + k * TPT
  // (RL the last radix). Other threads may still read LDS on return.
  static __device__ __forceinline__ void run_regs_keep(const cpx<C> *vin, cpx<C> *lds,
                                                       const cpx<C> *tw, int t, cpx<C> *keep) {
    auto none = [](int) { return cpx<C>{C(0), C(0)}; };
    auto drop = [](int, cpx<C>) {};
    stage<0, false, false, decltype(none), decltype(drop), true, true>(lds, tw, t, true, none,
                                                                       drop, vin, keep);
  }
};

// ---------------------------------------------------------------------------
// compile-time chain, split exchange: the same stages as StaticEngine, but
// each LDS exchange moves the real parts, then the imaginary parts, through
// one real array of N (padded) scalars. Half the LDS per transform lets two
// 16384-point blocks share a CU, so one block's global traffic overlaps the
// other's exchanges. Global input and output only (no TO_LDS / LDS_IN).
//
// MIRROR bit0 (first stage) / bit1 (last stage): the stage's butterflies are
// handed out in mirrored pairs, slot 2q of thread t holding butterfly
// lo = t + q*TPT and slot 2q+1 butterfly NB - lo (NB/2 when lo = 0), so that
// elements e and N - e (mod N) of that stage sit in one thread: the real
// transforms' Hermitian pre- and post-twiddles then need no exchange. Needs
// NB = KB * TPT with KB even.
// ---------------------------------------------------------------------------
template <typename C, int TPT, int PAD, typename Ch, int MIRROR = 0>
struct StaticEngineSplit {
  static constexpr int N = Ch::N;
  static constexpr int NS = Ch::count;
  static constexpr bool USES_LDS = NS > 1;
  template <int SI>
  struct St {
    static constexpr int R = Ch::radix(SI);
    static constexpr int L = Ch::L(SI);
    static constexpr int NB = N / R;
    static constexpr int KB = (NB + TPT - 1) / TPT;
    static __device__ __forceinline__ bool act(int b) { return (NB % TPT == 0) || (b < NB); }
  };

  // butterfly in slot k of thread t at stage SI
  template <int SI>
  static __device__ __forceinline__ int bfly(int t, int k) {
    using T = St<SI>;
    constexpr bool MIR = ((MIRROR & 1) && SI == 0) || ((MIRROR & 2) && SI == NS - 1);
    if constexpr (MIR) {
      static_assert(T::NB == T::KB * TPT && T::KB % 2 == 0,
                    "mirrored stage: NB = KB * TPT with KB even");
      const int lo = t + (k >> 1) * TPT;
      if (!(k & 1)) return lo;
      return lo == 0 ? T::NB / 2 : T::NB - lo;
    } else {
      return t + k * TPT;
    }
  }

  // twiddles (all but the first stage) and butterflies of stage SI
  template <int SI>
  static __device__ __forceinline__ void compute(cpx<C> (&v)[St<SI>::KB][St<SI>::R],
                                                 const cpx<C> *__restrict__ tw, int t) {
    using T = St<SI>;
#pragma unroll
    for (int k = 0; k < T::KB; ++k) {
      const int b = bfly<SI>(t, k);
      if constexpr (SI > 0) {
        const int j = b % T::L;
        if (T::act(b)) {
#pragma unroll
          for (int m = 1; m < T::R; ++m)
            v[k][m] = cmul(v[k][m], tw[Ch::tw_offset(SI) + (m - 1) * T::L + j]);
        }
      }
      Dft<T::R, C>::run(v[k]);
    }
  }

  // stage SI's outputs (component COMP) into LDS
  template <int SI, int COMP>
  static __device__ __forceinline__ void put(C *lds, const cpx<C> (&v)[St<SI>::KB][St<SI>::R],
                                             int t) {
    using T = St<SI>;
#pragma unroll
    for (int k = 0; k < T::KB; ++k) {
      const int b = bfly<SI>(t, k);
      if (T::act(b)) {
        const int base = (b / T::L) * T::L * T::R + b % T::L;
#pragma unroll
        for (int m = 0; m < T::R; ++m)
          lds[lds_pad_step<PAD, T::L>(base, m)] = COMP ? v[k][m].y : v[k][m].x;
      }
    }
  }

  // stage SI's inputs (component COMP) from LDS
  template <int SI, int COMP>
  static __device__ __forceinline__ void get(const C *lds, cpx<C> (&v)[St<SI>::KB][St<SI>::R],
                                             int t) {
    using T = St<SI>;
#pragma unroll
    for (int k = 0; k < T::KB; ++k) {
      const int b = bfly<SI>(t, k);
#pragma unroll
      for (int m = 0; m < T::R; ++m) {
        const C x = T::act(b) ? lds[lds_pad_step<PAD, T::NB>(b, m)] : C(0);
        if (COMP) v[k][m].y = x;
        else v[k][m].x = x;
      }
    }
  }

  // KEEP: the last stage's outputs go to keep[k * R + m] (registers)
  // instead of through `store`
  template <int SI, class Store, bool KEEP = false>
  static __device__ __forceinline__ void from(cpx<C> (&v)[St<SI>::KB][St<SI>::R], C *lds,
                                              const cpx<C> *__restrict__ tw, int t, bool valid,
                                              Store &store, cpx<C> *keep = nullptr) {
    using T = St<SI>;
    compute<SI>(v, tw, t);
    if constexpr (SI == NS - 1 && KEEP) {
#pragma unroll
      for (int k = 0; k < T::KB; ++k)
#pragma unroll
        for (int m = 0; m < T::R; ++m) keep[k * T::R + m] = v[k][m];
    } else if constexpr (SI == NS - 1) {
#pragma unroll
      for (int k = 0; k < T::KB; ++k) {
        const int b = bfly<SI>(t, k);
        if (T::act(b) && valid) {
#pragma unroll
          for (int m = 0; m < T::R; ++m) store(b + m * T::NB, v[k][m]);
        }
      }
    } else {
      cpx<C> w[St<SI + 1>::KB][St<SI + 1>::R];
      put<SI, 0>(lds, v, t);
      __syncthreads();
      get<SI + 1, 0>(lds, w, t);
      __syncthreads();
      put<SI, 1>(lds, v, t);
      __syncthreads();
      get<SI + 1, 1>(lds, w, t);
      if constexpr (SI + 1 < NS - 1) __syncthreads();  // the next exchange rewrites LDS
      from<SI + 1, Store, KEEP>(w, lds, tw, t, valid, store, keep);
    }
  }

  template <bool TO_LDS, bool LDS_IN = false, class Load, class Store>
  static __device__ __forceinline__ void run(cpx<C> *lds, const cpx<C> *tw, int t, bool valid,
                                             Load &&load, Store &&store) {
    static_assert(!TO_LDS && !LDS_IN, "split exchange: global input and output only");
    using T = St<0>;
    cpx<C> v[T::KB][T::R];
#pragma unroll
    for (int k = 0; k < T::KB; ++k) {
      const int b = bfly<0>(t, k);
#pragma unroll
      for (int m = 0; m < T::R; ++m)
        v[k][m] = (T::act(b) && valid) ? load(b + m * T::NB) : cpx<C>{C(0), C(0)};
    }
    from<0>(v, reinterpret_cast<C *>(lds), tw, t, valid, store);
  }

  // Every stage, with the last stage's outputs left in registers:
  // keep[k * RL + m] = element b + m * (N / RL) of butterfly b = bfly(t, k)
  // of the last stage (RL its radix; natural order). LDS may still be being
  // read by other threads on return: barrier before reusing it.
  template <class Load>
  static __device__ __forceinline__ void run_keep(cpx<C> *lds, const cpx<C> *tw, int t,
                                                  bool valid, Load &&load, cpx<C> *keep) {
    using T = St<0>;
    cpx<C> v[T::KB][T::R];
#pragma unroll
    for (int k = 0; k < T::KB; ++k) {
      const int b = bfly<0>(t, k);
#pragma unroll
      for (int m = 0; m < T::R; ++m)
        v[k][m] = (T::act(b) && valid) ? load(b + m * T::NB) : cpx<C>{C(0), C(0)};
    }
    auto none = [](int, cpx<C>) {};
    from<0, decltype(none), true>(v, reinterpret_cast<C *>(lds), tw, t, valid, none, keep);
  }

  // Stage 0 takes its inputs from registers, laid out as for
  // StaticEngine::run_regs, with butterfly bfly(t, k) in slot k.
  template <bool TO_LDS = false, class Store>
  static __device__ __forceinline__ void run_regs(const cpx<C> *vin, cpx<C> *lds,
                                                  const cpx<C> *tw, int t, Store &&store,
                                                  bool valid = true) {
    static_assert(!TO_LDS, "split exchange: global output only");
    using T = St<0>;
    cpx<C> v[T::KB][T::R];
#pragma unroll
    for (int k = 0; k < T::KB; ++k)
#pragma unroll
      for (int m = 0; m < T::R; ++m) v[k][m] = vin[k * T::R + m];
    from<0>(v, reinterpret_cast<C *>(lds), tw, t, valid, store);
  }
};

// ---------------------------------------------------------------------------
// runtime radix chain (descriptor in device memory). BIGP enables the
// radix 17/19/23 butterflies, which need more registers than a 1024-thread
// block can give each thread. SMALL compiles only radices 2, 3, 4, 5, 7, 8
// and 16 (7-smooth chains): a kernel that runs many engine instances keeps
// its code small.
// ---------------------------------------------------------------------------
template <typename C, int TPT, int PTS, bool BIGP, bool SMALL = false>
struct DynamicEngine {
  static constexpr int PAD = 5;

  template <int R, bool TO_LDS, class Load, class Store>
  static __device__ __forceinline__ void stage(const GenericDesc *__restrict__ d, int s,
                                               cpx<C> *lds, const cpx<C> *__restrict__ tw,
                                               int t, bool valid, bool lds_in, Load &load,
                                               Store &store) {
    constexpr int KMAX = (PTS + R - 1) / R;
    // Opaque copy of the thread index: without it the compiler hoists the
    // address arithmetic of every radix case out of the stage loop and keeps
    // it live across all stages (150 instead of 90 VGPRs at 16 points).
    asm volatile("" : "+v"(t));
    const int n = d->n;
    const int NB = n / R;
    const int L = d->L[s];
    const bool first = (s == 0);
    const bool from_global = first && !lds_in;
    const bool last = (s == d->ns - 1);
    const cpx<C> *tws = tw + d->twoff[s];
    cpx<C> v[KMAX][R];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      const bool act = b < NB;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        cpx<C> e = {C(0), C(0)};
        if (act) {
          if (from_global) {
            if (valid) e = load(b + m * NB);
          } else {
            e = lds[lds_pad<PAD>(b + m * NB)];
          }
        }
        v[k][m] = e;
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      if (!first && b < NB) {
        const int j = b % L;
#pragma unroll
        for (int m = 1; m < R; ++m) v[k][m] = cmul(v[k][m], tws[(m - 1) * L + j]);
      }
      Dft<R, C>::run(v[k]);
    }
    if (!from_global && (!last || TO_LDS)) __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      if (b < NB) {
        if (last && !TO_LDS) {
          if (valid) {
#pragma unroll
            for (int m = 0; m < R; ++m) store(b + m * NB, v[k][m]);
          }
        } else if (last) {
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad<PAD>(b + m * NB)] = v[k][m];
        } else {
          const int j = b % L;
          const int base = (b / L) * L * R + j;
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad<PAD>(base + m * L)] = v[k][m];
        }
      }
    }
    if (!last || TO_LDS) __syncthreads();
  }

  template <bool TO_LDS, bool LDS_IN = false, class Load, class Store>
  static __device__ __forceinline__ void run(const GenericDesc *__restrict__ d, cpx<C> *lds,
                                             const cpx<C> *tw, int t, bool valid, Load &&load,
                                             Store &&store) {
    const int ns = d->ns;
    if (ns == 0) {  // n == 1: identity
      if (TO_LDS) {
        if (t == 0 && valid && !LDS_IN) lds[0] = load(0);
        __syncthreads();
      } else if (t == 0 && valid) {
        store(0, LDS_IN ? lds[0] : load(0));
      }
      return;
    }
    for (int s = 0; s < ns; ++s) {
      const bool lds_in = LDS_IN && s == 0;
      switch (d->radix[s]) {
#define PLF_DYN_CASE(R)                                                      \
  case R:                                                                    \
    stage<R, TO_LDS>(d, s, lds, tw, t, valid, lds_in, load, store);          \
    break;
        PLF_DYN_CASE(2)
        PLF_DYN_CASE(3)
        PLF_DYN_CASE(4)
        PLF_DYN_CASE(5)
        PLF_DYN_CASE(7)
        PLF_DYN_CASE(8)
        PLF_DYN_CASE(16)
#undef PLF_DYN_CASE
        case 11:
          if constexpr (!SMALL) stage<11, TO_LDS>(d, s, lds, tw, t, valid, lds_in, load, store);
          break;
        case 13:
          if constexpr (!SMALL) stage<13, TO_LDS>(d, s, lds, tw, t, valid, lds_in, load, store);
          break;
        case 17:
          if constexpr (BIGP && !SMALL)
            stage<17, TO_LDS>(d, s, lds, tw, t, valid, lds_in, load, store);
          break;
        case 19:
          if constexpr (BIGP && !SMALL)
            stage<19, TO_LDS>(d, s, lds, tw, t, valid, lds_in, load, store);
          break;
        case 23:
          if constexpr (BIGP && !SMALL)
            stage<23, TO_LDS>(d, s, lds, tw, t, valid, lds_in, load, store);
          break;
        default:
          break;
      }
    }
  }
};

}  // namespace plf
