// kernels_conv.hpp -- prime lengths: fused Bluestein (chirp-z) and Rader
// kernels, one block per member with everything resident in LDS.
//
// Bluestein (src/planner/strategies/bluestein.hpp:66-146): with the chirp
// c_k = exp(-i pi (k^2 mod 2n) / n) (phase reduced in integers on the host),
//   X = c .* IDFT_M( DFT_M(x .* c, zero-padded to M) .* B ),
//   B = DFT_M(b) / M,  b_k = b_{M-k} = conj(c_k) for 0 <= k < n, else 0.
// Rader (src/planner/strategies/rader.hpp:36-124), n prime, m = n - 1,
// generator g, perm[k] = g^k mod n:
//   a_k = x[perm[k]],  A = DFT_m(a),  X[0] = x0 + A[0],
//   conv = IDFT_m(A .* B),  B = DFT_m(b) / m,  b_{(m-k) mod m} = W_n^{perm[k]},
//   X[perm[(m - e) mod m]] = x0 + conv[e].
// Both run the forward engine twice in one kernel: the inverse DFT is
// swap(DFT(swap(.))). A backward transform swaps at the first load and the
// final store, like the other kernels.
#pragma once

#include "kernels.hpp"

namespace plf {

// Between the two transforms of a static-chain engine: B is loaded before
// the first transform at the slots where the second one's stage 0 takes its
// inputs (butterfly b = t + k TPT, element b + m N / R0), and one LDS read
// turns the first transform's natural-order output into those inputs,
// times B and swapped; `first` sees element 0. The caller puts a barrier
// after it and runs the second transform from registers.
template <class Eng>
struct ConvMid {
  using C = typename Eng::C;
  using Ch = typename Eng::Ch;
  static constexpr int TPT = Eng::TPT;
  static constexpr int R0 = Ch::radix(0);
  static constexpr int NB0 = Ch::N / R0;
  static constexpr int KB0 = (NB0 + TPT - 1) / TPT;
  static constexpr int SLOTS = KB0 * R0;
  static __device__ __forceinline__ bool act(int t, int k) {
    return (NB0 % TPT == 0) || (t + k * TPT < NB0);
  }
  static __device__ __forceinline__ void load_B(const cpx<C> *B, int t, bool valid,
                                                cpx<C> (&bv)[SLOTS]) {
#pragma unroll
    for (int k = 0; k < KB0; ++k)
#pragma unroll
      for (int m = 0; m < R0; ++m)
        bv[k * R0 + m] = (act(t, k) && valid) ? B[t + k * TPT + m * NB0] : cpx<C>{C(0), C(0)};
  }
  template <class First>
  static __device__ __forceinline__ void gather(const cpx<C> *lds, int t,
                                                const cpx<C> (&bv)[SLOTS], cpx<C> (&vin)[SLOTS],
                                                First &&first) {
#pragma unroll
    for (int k = 0; k < KB0; ++k)
#pragma unroll
      for (int m = 0; m < R0; ++m) {
        const cpx<C> A = act(t, k) ? lds[lds_pad<Eng::PAD>(t + k * TPT + m * NB0)]
                                   : cpx<C>{C(0), C(0)};
        if (k == 0 && m == 0 && t == 0) first(A);
        vin[k * R0 + m] = cswap(cmul(A, bv[k * R0 + m]));
      }
  }
};

// aux0: B (M entries), aux1: chirp c (n entries); a.n = M, a.n1 = n.
// Gathered launch (aux2 = BlueMember records, one per transform; a mixed
// batch's members and rows of equal M): length, offsets, output stride and
// table offsets per record, unit input stride.
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_bluestein(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int M = (int)a.n;
  const BlueMember *mt = static_cast<const BlueMember *>(a.aux2);
  const BlueMember *rec = mt ? mt + (valid ? tr : 0) : nullptr;
  const int n = mt ? (int)rec->n : (int)a.n1;
  const E *in = static_cast<const E *>(a.in) + (mt ? rec->in_off : valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (mt ? rec->out_off : valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(M);
  const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw);
  const cpx<C> *B = static_cast<const cpx<C> *>(a.aux0) + (mt ? rec->B_off : 0);
  const cpx<C> *chirp = static_cast<const cpx<C> *>(a.aux1) + (mt ? rec->chirp_off : 0);
  const int64_t is = mt ? 1 : a.istride, os = mt ? rec->ostride : a.ostride;
  auto load = [&](int e) {
    if (e >= n) return cpx<C>{C(0), C(0)};
    cpx<C> v = ST::ld(in[(int64_t)e * is]);
    if (BWD) v = cswap(v);
    return cmul(v, chirp[e]);
  };
  auto store = [&](int e, cpx<C> v) {
    if (e < n) {
      cpx<C> r = cmul(cswap(v), chirp[e]);
      out[(int64_t)e * os] = ST::st(BWD ? cswap(r) : r);
    }
  };
  if constexpr (Eng::NFIX > 0) {
    using Mid = ConvMid<Eng>;
    cpx<C> bv[Mid::SLOTS], vin[Mid::SLOTS];
    Mid::load_B(B, t, valid, bv);
    Eng::template run<true>(a.desc, lds, tw, t, valid, load, [&](int, cpx<C>) {});
    Mid::gather(lds, t, bv, vin, [](cpx<C>) {});
    __syncthreads();  // the second transform's first exchange rewrites LDS
    Eng::E::run_regs(vin, lds, tw, t, store, valid);
  } else {
    Eng::template run<true>(a.desc, lds, tw, t, valid, load, [&](int, cpx<C>) {});
    for (int e = t; e < M; e += TPT) {
      const int p = lds_pad<Eng::PAD>(e);
      lds[p] = cswap(cmul(lds[p], B[e]));
    }
    __syncthreads();
    Eng::template run<false, true>(
        a.desc, lds, tw, t, valid, [&](int) { return cpx<C>{C(0), C(0)}; }, store);
  }
}

// aux0: B (m entries), aux1: perm (int32, m entries); a.n = m, a.n1 = n.
template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_rader(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int m = (int)a.n;
  const E *in = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(m);
  const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw);
  const cpx<C> *B = static_cast<const cpx<C> *>(a.aux0);
  const int *perm = static_cast<const int *>(a.aux1);
  const int64_t is = a.istride, os = a.ostride;
  cpx<C> x0 = {C(0), C(0)};
  if (valid) {
    x0 = ST::ld(in[0]);
    if (BWD) x0 = cswap(x0);
  }
  auto load = [&](int e) {
    cpx<C> v = ST::ld(in[(int64_t)perm[e] * is]);
    return BWD ? cswap(v) : v;
  };
  auto store = [&](int e, cpx<C> v) {
    const int k = perm[e == 0 ? 0 : m - e];
    const cpx<C> r = cadd(x0, cswap(v));
    out[(int64_t)k * os] = ST::st(BWD ? cswap(r) : r);
  };
  auto first = [&](cpx<C> A) {  // X_0 = x_0 + A_0
    if (valid) {
      const cpx<C> y0 = cadd(x0, A);
      out[0] = ST::st(BWD ? cswap(y0) : y0);
    }
  };
  // (the register-resident middle step of k_bluestein measured slower here:
  // the Rader chains 4.8.3 and 16.3.3.7 are already register-bound)
  Eng::template run<true>(a.desc, lds, tw, t, valid, load, [&](int, cpx<C>) {});
  for (int e = t; e < m; e += TPT) {
    const int p = lds_pad<Eng::PAD>(e);
    const cpx<C> A = lds[p];
    if (e == 0) first(A);
    lds[p] = cswap(cmul(A, B[e]));
  }
  __syncthreads();
  Eng::template run<false, true>(
      a.desc, lds, tw, t, valid, [&](int) { return cpx<C>{C(0), C(0)}; }, store);
}

}  // namespace plf
