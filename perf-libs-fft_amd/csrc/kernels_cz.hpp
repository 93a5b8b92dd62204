// kernels_cz.hpp -- chirp-z (Bluestein whose DFT_M is a four-step) in three
// HBM passes on the compile-time column / row engines.
//
// With M = n1 * n2 >= 2n - 2, c_j the chirp and B = DFT_M(b) / M
// (bluestein.hpp:66-98), y = conj-chirp o IDFT_M(B o DFT_M(c o x)) is the
// DFT of x. The passes over a member's workspace (M complex, row-major
// n1 x n2):
//   1. column c2: a_j = x_j c_j (j = c2 + n2 e < n, else 0), DFT_n1 over e,
//      times W_M^{c2 k1}, into row k1, column c2;
//   2. row k1: DFT_n2 gives A[k1 + n1 k2]; times B (stored row-major by k1),
//      swapped; the inverse transform's four-step, taken with the sides
//      exchanged, starts with a DFT_n2 over the same row, so it runs on the
//      same LDS contents, is twiddled by W_M^{k1 m2} and written back in place;
//   3. column m2: DFT_n1 over k1 gives bin m = m2 + n2 m1; swapped back,
//      times c_m, stored for m < n.
// The backward transform is swap o forward o swap on the input and output.
//
// KArgs of all three: tw = the engine's Stockham twiddles, aux0 = W_M split
// table (lo: 2^shift entries, then hi), n = engine length, n1, n2 = the
// sides, bigdist = the transform length n, count = columns (members * n2) or
// rows (members * n1). Pass 1: in = input (storage type, istride, idist),
// out = workspace, aux1 = chirp. Pass 2: out = workspace, aux1 = B. Pass 3:
// in = workspace, out = output (ostride, odist), aux1 = chirp.
// Gathered launch (aux2 = CzMember records, one per member of the launch;
// the members of a mixed-length batch that share M): each member's length,
// unit-stride offsets, workspace and table offsets come from its record.
#pragma once

#include "kernels.hpp"

namespace plf {

template <typename C>
__device__ __forceinline__ cpx<C> cz_twiddle(const cpx<C> *__restrict__ split, int shift,
                                             int64_t e) {
  const cpx<C> *hi = split + (int64_t(1) << shift);
  return cmul(hi[e >> shift], split[e & ((int64_t(1) << shift) - 1)]);
}

template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_cz_col_in(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  const int g = threadIdx.x % TPB;
  const int t = threadIdx.x / TPB;
  const int64_t col = (int64_t)blockIdx.x * TPB + g;
  const bool valid = col < a.count;
  const int64_t n1 = a.n1, n2 = a.n2;
  const int64_t bi = valid ? col / n2 : 0, c2 = valid ? col % n2 : 0;
  const CzMember *mt = static_cast<const CzMember *>(a.aux2);
  const int64_t nn = mt ? mt[bi].n : a.bigdist;
  const E *in = static_cast<const E *>(a.in) + (mt ? mt[bi].in_off : bi * a.idist);
  cpx<C> *ws = static_cast<cpx<C> *>(a.out) + (mt ? mt[bi].ws_off : bi * n1 * n2) + c2;
  const cpx<C> *split = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *chirp = static_cast<const cpx<C> *>(a.aux1) + (mt ? mt[bi].chirp_off : 0);
  const int64_t is = mt ? 1 : a.istride;
  const int shift = a.shift;
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        const int64_t j = c2 + n2 * e;
        if (j >= nn) return cpx<C>{C(0), C(0)};  // zero padding
        cpx<C> v = ST::ld(in[j * is]);
        return cmul(BWD ? cswap(v) : v, chirp[j]);
      },
      [&](int k1, cpx<C> v) { ws[(int64_t)k1 * n2] = cmul(v, cz_twiddle(split, shift, c2 * k1)); });
}

template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_cz_row_mid(KArgs a) {
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t row = (int64_t)blockIdx.x * TPB + g;
  const bool valid = row < a.count;
  const int64_t n1 = a.n1, n2 = a.n2;
  const int64_t bi = valid ? row / n1 : 0, k1 = valid ? row % n1 : 0;
  const CzMember *mt = static_cast<const CzMember *>(a.aux2);
  cpx<C> *rowp = static_cast<cpx<C> *>(a.out) + (mt ? mt[bi].ws_off : bi * n1 * n2) + k1 * n2;
  const cpx<C> *split = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *Brow = static_cast<const cpx<C> *>(a.aux1) + (mt ? mt[bi].B_off : 0) + k1 * n2;
  const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw);
  const int shift = a.shift;
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  Eng::template run<true>(
      a.desc, lds, tw, t, valid, [&](int e) { return rowp[e]; }, [&](int, cpx<C>) {});
  if (valid)
    for (int k2 = t; k2 < n2; k2 += TPT) {
      const int p = lds_pad<Eng::PAD>(k2);
      lds[p] = cswap(cmul(lds[p], Brow[k2]));
    }
  __syncthreads();
  Eng::template run<false, true>(
      a.desc, lds, tw, t, valid, [&](int) { return cpx<C>{C(0), C(0)}; },
      [&](int m2, cpx<C> v) { rowp[m2] = cmul(v, cz_twiddle(split, shift, k1 * m2)); });
}

template <typename S, bool BWD, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_cz_col_out(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  const int g = threadIdx.x % TPB;
  const int t = threadIdx.x / TPB;
  const int64_t col = (int64_t)blockIdx.x * TPB + g;
  const bool valid = col < a.count;
  const int64_t n1 = a.n1, n2 = a.n2;
  const int64_t bi = valid ? col / n2 : 0, c2 = valid ? col % n2 : 0;
  const CzMember *mt = static_cast<const CzMember *>(a.aux2);
  const int64_t nn = mt ? mt[bi].n : a.bigdist;
  const cpx<C> *ws = static_cast<const cpx<C> *>(a.in) + (mt ? mt[bi].ws_off : bi * n1 * n2) + c2;
  E *out = static_cast<E *>(a.out) + (mt ? mt[bi].out_off : bi * a.odist);
  const cpx<C> *chirp = static_cast<const cpx<C> *>(a.aux1) + (mt ? mt[bi].chirp_off : 0);
  const int64_t os = mt ? 1 : a.ostride;
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return ws[(int64_t)e * n2]; },
      [&](int m1, cpx<C> v) {
        const int64_t m = c2 + n2 * m1;
        if (m < nn) {
          const cpx<C> y = cmul(cswap(v), chirp[m]);
          out[m * os] = ST::st(BWD ? cswap(y) : y);
        }
      });
}

}  // namespace plf
