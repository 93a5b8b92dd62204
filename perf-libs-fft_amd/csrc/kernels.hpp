// kernels.hpp -- batched transform kernels built on the Stockham engines.
//
// Engine policies:
//   SEng<C, TPT, Rs...>  compile-time radix chain (fully specialised)
//   DEng<C, TPT, PTS, TPB> runtime chain from a GenericDesc in device memory
// Kernels (one argument block, KArgs):
//   k_c2c  batched C2C, member = TPT-thread group      (replaces the
//          reference's loop + cooley_tukey ab/ac plans,
//          src/planner/strategies/{loop,cooley_tukey}.hpp)
//   k_col  four-step pass A: length-n1 column DFTs over a tile of TPB
//          consecutive columns, times W_N^{col*k1}, into the workspace
//   k_row  four-step pass B: length-n2 row DFTs of the workspace with a
//          transposed (column-contiguous) store through LDS
// A backward transform swaps re/im at the global load and the final store.
#pragma once

#include "engine.hpp"

namespace plf {

template <typename C_, int TPT_, int PAD_, bool PAIR_, int... Rs>
struct SEngX {
  using C = C_;
  static constexpr int TPT = TPT_;
  static constexpr int PAD = PAD_;
  using Ch = Chain<Rs...>;
  using E = StaticEngine<C, TPT, PAD, Ch, PAIR_>;
  static constexpr int NFIX = Ch::N;
  // stage-0 inputs may be handed over in registers (E::run_regs)
  static constexpr bool REGS_IN = !PAIR_;
  static constexpr bool USES_LDS = E::USES_LDS;
  static __host__ __device__ constexpr int lds_per(int64_t) { return lds_padded<PAD>(Ch::N); }
  template <bool TO_LDS, bool LDS_IN = false, class Load, class Store>
  static __device__ __forceinline__ void run(const GenericDesc *, cpx<C> *lds, const cpx<C> *tw,
                                             int t, bool valid, Load &&load, Store &&store) {
    E::template run<TO_LDS, LDS_IN>(lds, tw, t, valid, load, store);
  }
};
template <typename C_, int TPT_, int... Rs>
using SEng = SEngX<C_, TPT_, 5, false, Rs...>;

// compile-time chain with the split (real, then imaginary) LDS exchange:
// half the LDS of SEngX; lds_per counts complex slots
template <typename C_, int TPT_, int PAD_, int... Rs>
struct SEngS {
  using C = C_;
  static constexpr int TPT = TPT_;
  static constexpr int PAD = PAD_;
  using Ch = Chain<Rs...>;
  using E = StaticEngineSplit<C, TPT, PAD, Ch>;
  static constexpr int NFIX = Ch::N;
  static constexpr bool REGS_IN = true;
  static constexpr bool USES_LDS = E::USES_LDS;
  static __host__ __device__ constexpr int lds_per(int64_t) {
    return (lds_padded<PAD>(Ch::N) + 1) / 2;
  }
  template <bool TO_LDS, bool LDS_IN = false, class Load, class Store>
  static __device__ __forceinline__ void run(const GenericDesc *, cpx<C> *lds, const cpx<C> *tw,
                                             int t, bool valid, Load &&load, Store &&store) {
    E::template run<TO_LDS, LDS_IN>(lds, tw, t, valid, load, store);
  }
};

template <typename C_, int TPT_, int PTS_, int TPB_>
struct DEng {
  using C = C_;
  static constexpr int TPT = TPT_;
  static constexpr int PTS = PTS_;
  static constexpr int PAD = 5;
  static constexpr int NFIX = 0;
  static constexpr bool REGS_IN = false;
  static constexpr bool USES_LDS = true;
  static constexpr bool BIGP = TPT_ <= 256 && TPB_ <= 16 && TPT_ * TPB_ <= 256;
  using E = DynamicEngine<C, TPT, PTS, BIGP>;
  static __host__ __device__ constexpr int lds_per(int64_t n) { return lds_padded<PAD>((int)n); }
  template <bool TO_LDS, bool LDS_IN = false, class Load, class Store>
  static __device__ __forceinline__ void run(const GenericDesc *d, cpx<C> *lds, const cpx<C> *tw,
                                             int t, bool valid, Load &&load, Store &&store) {
    E::template run<TO_LDS, LDS_IN>(d, lds, tw, t, valid, load, store);
  }
};

// Logical block index that keeps consecutive indices on one XCD. The
// dispatcher hands block b to XCD b mod 8; with this order neighbouring
// column tiles, which share the cache lines at their edges (and, for
// Rader's permuted columns, every line of a member's input), run in one L2.
// Block b is the (b >> 3)-th block of XCD x = b & 7, and XCD x owns the
// q + [x < r] logical indices after those of XCDs 0..x-1 (q = G >> 3,
// r = G & 7), so the map is a bijection of [0, G). Used by the chirp-z and
// Rader passes; the four-step passes k_col / k_row measured 3-7 % slower in
// this order at 65536 and 2^20 points (FP32) and keep the dispatch order.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t G) {
  const int64_t x = b & 7, q = G >> 3, r = G & 7;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + (b >> 3);
}

template <typename C>
__device__ __forceinline__ cpx<C> *dyn_lds() {
  extern __shared__ unsigned char plf_dyn_lds[];
  return reinterpret_cast<cpx<C> *>(plf_dyn_lds);
}

// ---------------------------------------------------------------------------
// batched C2C: in[b*idist + e*istride] -> out[b*odist + k*ostride]
// ---------------------------------------------------------------------------
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_c2c(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const E *in = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  const int64_t is = a.istride, os = a.ostride;
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        cpx<C> v = ST::ld(in[(int64_t)e * is]);
        return BWD ? cswap(v) : v;
      },
      [&](int e, cpx<C> v) { out[(int64_t)e * os] = ST::st(BWD ? cswap(v) : v); });
}

// Batched C2C of a small odd prime N (3..23) with packed members (unit
// stride, distance N): one member per thread, 256 members per block. The
// block's 256 N consecutive elements pass through LDS on the way in and
// out, so every global load and store instruction of a wave covers 512
// contiguous bytes instead of 64 lanes N elements apart; the member's
// DFT_N runs in registers (lds[t N + j] with N odd meets each bank about
// twice per wave). KArgs: in, out, count. In place is safe: a block reads
// all of its members before it writes any.
template <typename S, bool BWD, int N>
__global__ void __launch_bounds__(256) k_c2c_staged(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename ST::C;
  cpx<C> *lds = dyn_lds<C>();
  const int t = threadIdx.x;
  const int64_t m0 = (int64_t)blockIdx.x * 256;
  const int members = (int)(a.count - m0 < 256 ? a.count - m0 : 256);
  const int total = members * N;
  const E *in = static_cast<const E *>(a.in) + m0 * N;
  E *out = static_cast<E *>(a.out) + m0 * N;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int e = t + 256 * k;
    if (e < total) lds[e] = ST::ldnt(in + e);
  }
  __syncthreads();
  cpx<C> v[N];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const cpx<C> x = lds[t * N + j];
    v[j] = BWD ? cswap(x) : x;
  }
  Dft<N, C>::run(v);
#pragma unroll
  for (int j = 0; j < N; ++j) lds[t * N + j] = BWD ? cswap(v[j]) : v[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int e = t + 256 * k;
    if (e < total) ST::stnt(out + e, lds[e]);
  }
}

// batched C2C, unit element stride: streaming (non-temporal) global access;
// with a PAIR engine neighbouring elements of one thread merge into 16-byte
// accesses.
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_c2c_unit(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const E *in = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        cpx<C> v = ST::ldnt(in + e);
        return BWD ? cswap(v) : v;
      },
      [&](int e, cpx<C> v) { ST::stnt(out + e, BWD ? cswap(v) : v); });
}

// k_c2c_unit with at least WAVES waves per SIMD: the register budget is cut
// to 512 / WAVES so that more than one large block fits a CU (for 16384
// points on the split exchange: two 1024-thread blocks at 64 registers).
template <typename S, bool BWD, int TPB, class Eng, int WAVES>
__global__ void __launch_bounds__(Eng::TPT *TPB) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_c2c_unit_occ(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const E *in = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        cpx<C> v = ST::ldnt(in + e);
        return BWD ? cswap(v) : v;
      },
      [&](int e, cpx<C> v) { ST::stnt(out + e, BWD ? cswap(v) : v); });
}

// batched C2C, unit element stride, software-pipelined: a block walks the
// members blockIdx.x, +gridDim.x, ... and issues the stage-0 loads of its
// next member before running the LDS stages of the current one, so each
// block keeps global loads in flight through its own exchanges. Launched
// with a grid of at most (CUs x resident blocks per CU).
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT) k_c2c_pipe(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  using Ch = typename Eng::Ch;
  constexpr int TPT = Eng::TPT;
  constexpr int N = Ch::N;
  constexpr int R0 = Ch::radix(0);
  constexpr int NB0 = N / R0;
  constexpr int KB0 = (NB0 + TPT - 1) / TPT;
  const int t = threadIdx.x;
  cpx<C> *lds = dyn_lds<C>();
  const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw);
  cpx<C> nxt[KB0 * R0];
  auto fetch = [&](int64_t mem) {
    const E *in = static_cast<const E *>(a.in) + mem * a.idist;
#pragma unroll
    for (int k = 0; k < KB0; ++k) {
      const int b = t + k * TPT;
      const bool act = (NB0 % TPT == 0) || (b < NB0);
#pragma unroll
      for (int m = 0; m < R0; ++m) {
        cpx<C> v = act ? ST::ldnt(in + b + m * NB0) : cpx<C>{C(0), C(0)};
        nxt[k * R0 + m] = BWD ? cswap(v) : v;
      }
    }
  };
  int64_t mem = blockIdx.x;
  if (mem < a.count) fetch(mem);
  for (; mem < a.count; mem += gridDim.x) {
    cpx<C> cur[KB0 * R0];
#pragma unroll
    for (int i = 0; i < KB0 * R0; ++i) cur[i] = nxt[i];
    if (mem + gridDim.x < a.count) fetch(mem + gridDim.x);
    E *out = static_cast<E *>(a.out) + mem * a.odist;
    Eng::E::run_regs(cur, lds, tw, t,
                     [&](int e, cpx<C> v) { ST::stnt(out + e, BWD ? cswap(v) : v); });
    __syncthreads();  // the next member's stage 0 rewrites this member's LDS
  }
}

// W_N^e from the split table: hi[e >> shift] * lo[e & (2^shift - 1)].
template <typename C>
__device__ __forceinline__ cpx<C> split_twiddle(const cpx<C> *__restrict__ lo,
                                                const cpx<C> *__restrict__ hi, int64_t e,
                                                int shift) {
  const int64_t mask = (int64_t(1) << shift) - 1;
  return cmul(hi[e >> shift], lo[e & mask]);
}

// ---------------------------------------------------------------------------
// four-step pass A. Column c2 of big transform bi (col = bi*n2 + c2, count =
// chunk*n2 columns): x[bi*idist + (c2 + e*n2)*istride], e < n1. Output
// k1 goes to ws[bi*bigdist + k1*n2 + c2] times W_N^{c2*k1}. Threads are
// column-fastest so that a wave reads and writes TPB-wide contiguous runs.
// ---------------------------------------------------------------------------
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_col(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  const int g = threadIdx.x % TPB;
  const int t = threadIdx.x / TPB;
  const int64_t col = (int64_t)blockIdx.x * TPB + g;
  const bool valid = col < a.count;
  const int64_t bi = valid ? col / a.n2 : 0;
  const int64_t c2 = valid ? col % a.n2 : 0;
  const E *in = static_cast<const E *>(a.in) + bi * a.idist + c2 * a.istride;
  cpx<C> *ws = static_cast<cpx<C> *>(a.out) + bi * a.bigdist + c2;
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  const int64_t es = a.n2 * a.istride, n2 = a.n2;
  const cpx<C> *lo = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *hi = static_cast<const cpx<C> *>(a.aux1);
  const int shift = a.shift;
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        cpx<C> v = ST::ld(in[(int64_t)e * es]);
        return BWD ? cswap(v) : v;
      },
      [&](int k1, cpx<C> v) {
        ws[(int64_t)k1 * n2] = cmul(v, split_twiddle(lo, hi, c2 * (int64_t)k1, shift));
      });
}

// ---------------------------------------------------------------------------
// four-step pass B. Row k1 of big transform bi (row = bi*n1 + k1, count =
// chunk*n1 rows): ws[bi*bigdist + k1*n2 + e], e < n2. Output k2 goes to
// out[bi*odist + (k1 + n1*k2)*ostride]; the block's TPB consecutive rows are
// written column-contiguously through LDS (a block may straddle two big
// transforms).
// ---------------------------------------------------------------------------
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_row(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t row0 = (int64_t)blockIdx.x * TPB;
  const int64_t row = row0 + g;
  const bool valid = row < a.count;
  const int lper = Eng::lds_per(a.n);
  cpx<C> *lds_all = dyn_lds<C>();
  cpx<C> *lds = lds_all + g * lper;
  const int64_t n1 = a.n1, n2 = a.n2;
  const cpx<C> *ws = static_cast<const cpx<C> *>(a.in) +
                     (valid ? (row / n1) * a.bigdist + (row % n1) * n2 : 0);
  Eng::template run<true>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid, [&](int e) { return ws[e]; },
      [&](int, cpx<C>) {});
  // transposed store: thread -> column c = tid % TPB, k2 = tid / TPB + j*TPT
  const int c = threadIdx.x % TPB;
  const int64_t rc = row0 + c;
  if (rc >= a.count) return;
  const int64_t os = a.ostride;
  E *out = static_cast<E *>(a.out) + (rc / n1) * a.odist + (rc % n1) * os;
  const cpx<C> *src = lds_all + c * lper;
  const int64_t step = n1 * os;
  for (int k2 = threadIdx.x / TPB; k2 < n2; k2 += TPT) {
    cpx<C> v = src[lds_pad<Eng::PAD>(k2)];
    out[k2 * step] = ST::st(BWD ? cswap(v) : v);
  }
}

}  // namespace plf
