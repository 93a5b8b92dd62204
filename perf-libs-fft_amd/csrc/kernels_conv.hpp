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
  Eng::template run<true>(
      a.desc, lds, tw, t, valid,
      [&](int e) {
        if (e >= n) return cpx<C>{C(0), C(0)};
        cpx<C> v = ST::ld(in[(int64_t)e * is]);
        if (BWD) v = cswap(v);
        return cmul(v, chirp[e]);
      },
      [&](int, cpx<C>) {});
  for (int e = t; e < M; e += TPT) {
    const int p = lds_pad<Eng::PAD>(e);
    lds[p] = cswap(cmul(lds[p], B[e]));
  }
  __syncthreads();
  Eng::template run<false, true>(
      a.desc, lds, tw, t, valid, [&](int) { return cpx<C>{C(0), C(0)}; },
      [&](int e, cpx<C> v) {
        if (e < n) {
          cpx<C> r = cmul(cswap(v), chirp[e]);
          out[(int64_t)e * os] = ST::st(BWD ? cswap(r) : r);
        }
      });
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
  Eng::template run<true>(
      a.desc, lds, tw, t, valid,
      [&](int e) {
        cpx<C> v = ST::ld(in[(int64_t)perm[e] * is]);
        return BWD ? cswap(v) : v;
      },
      [&](int, cpx<C>) {});
  for (int e = t; e < m; e += TPT) {
    const int p = lds_pad<Eng::PAD>(e);
    const cpx<C> A = lds[p];
    if (e == 0 && valid) {
      const cpx<C> y0 = cadd(x0, A);
      out[0] = ST::st(BWD ? cswap(y0) : y0);
    }
    lds[p] = cswap(cmul(A, B[e]));
  }
  __syncthreads();
  Eng::template run<false, true>(
      a.desc, lds, tw, t, valid, [&](int) { return cpx<C>{C(0), C(0)}; },
      [&](int e, cpx<C> v) {
        const int k = perm[e == 0 ? 0 : m - e];
        const cpx<C> r = cadd(x0, cswap(v));
        out[(int64_t)k * os] = ST::st(BWD ? cswap(r) : r);
      });
}

}  // namespace plf
