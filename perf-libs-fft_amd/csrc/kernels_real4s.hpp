// kernels_real4s.hpp -- R2C / C2R of even length n = 2M when the half
// length M is too long for one LDS tile.
//
// The packed half-length transform of kernels_real.hpp (z[j] = x[2j] +
// i x[2j+1], one length-M complex DFT, a Hermitian post- or pre-twiddle) runs
// on the four-step passes of kernels.hpp with M = m1 * m2:
//   R2C  pass A (k_col_r2c): column c2 < m2 of member bi reads z[c2 + m2*e],
//        e < m1, straight from the reals, and writes its length-m1 DFT times
//        W_M^{c2*k1} to ws[bi*M + k1*m2 + c2] (the complex pass A, k_col).
//        Pass B (k_row_r2c): one block runs the length-m2 DFTs of rows k1 and
//        m1 - k1 together (slot s < m1/2 holds rows s and m1 - s; slot 0 holds
//        rows 0 and m1/2), so that Z[k] and Z[M - k] are both in its LDS:
//          k = k1 + m1*k2  ->  M - k = (m1 - k1) + m1*(m2 - 1 - k2)  (k1 > 0)
//                                      m1*((m2 - k2) mod m2)        (k1 = 0)
//        and the post-twiddle
//          X[k] = (Z[k] + conj Z[M-k])/2 - (i/2) W_n^k (Z[k] - conj Z[M-k]),
//          X[M] = Re Z[0] - Im Z[0]
//        is applied on the transposed store.
//   C2R  pass A (k_col_c2r): element j = c2 + m2*e of column c2 is formed on
//        the load from X[j] and X[M - j] (two global loads; the partner column
//        m2 - c2 is read by a block of the same member, launched alongside):
//          Z'[j] = (X[j] + conj X[M-j]) + i W_n^{-j} (X[j] - conj X[M-j]),
//        then the inverse DFT (forward DFT of the swapped values) as in k_col.
//        Pass B (k_row_c2r): the row DFTs of k_row, whose transposed store
//        writes z[j] = n (x[2j] + i x[2j+1]) as pairs of reals.
// Two HBM round trips on M complex values: the same passes as a complex
// four-step of length M, half the bytes of widening to a length-n complex
// transform. Replaces the reference's Hermitian-folded Cooley-Tukey stages
// (src/planner/strategies/cooley_tukey.hpp:145-323, 402-580) for long n.
//
// Twiddles: aux0 is the split table of W_M (lo: 2^shift entries, then hi),
// aux1 the split table of W_n (lo: 2^shift2 entries, then hi); forward sign.
// flags bit0: the real side may be accessed as pairs (unit real stride, even
// distance, pair-aligned base); bit1: the adjoint weighting (interior bins
// doubled for R2C, halved for C2R), as in kernels_real.hpp.
#pragma once

#include "kernels.hpp"
#include "kernels_real.hpp"

namespace plf {

// W^e from a packed split table: lo = tab[0 .. 2^shift), hi = tab[2^shift ..).
template <typename C>
__device__ __forceinline__ cpx<C> packed_twiddle(const cpx<C> *__restrict__ tab, int64_t e,
                                                 int shift) {
  return split_twiddle(tab, tab + (int64_t(1) << shift), e, shift);
}

// ---------------------------------------------------------------------------
// R2C pass A: count = chunk * m2 columns; idist counts reals.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_col_r2c(KArgs a) {
  using ST = StorageT<S>;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  const int g = threadIdx.x % TPB;
  const int t = threadIdx.x / TPB;
  const int64_t col = (int64_t)blockIdx.x * TPB + g;
  const bool valid = col < a.count;
  const int64_t bi = valid ? col / a.n2 : 0;
  const int64_t c2 = valid ? col % a.n2 : 0;
  const Rs *x = static_cast<const Rs *>(a.in) + bi * a.idist;
  cpx<C> *ws = static_cast<cpx<C> *>(a.out) + bi * a.bigdist + c2;
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  const int64_t n2 = a.n2, is = a.istride;
  const bool pair_ok = a.flags & 1;
  const cpx<C> *wm = static_cast<const cpx<C> *>(a.aux0);
  const int shift = a.shift;
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return load_pair<ST>(x, c2 + (int64_t)e * n2, is, pair_ok); },
      [&](int k1, cpx<C> v) {
        ws[(int64_t)k1 * n2] = cmul(v, packed_twiddle(wm, c2 * (int64_t)k1, shift));
      });
}

// ---------------------------------------------------------------------------
// R2C pass B: count = chunk members; a member's m1/2 slots are spread over
// ceil((m1/2) / (TPB/2)) consecutive blocks; LDS row g < TPB/2 holds the low
// row of slot s0 + g, row g + TPB/2 its partner. Output: odist / ostride in
// complex elements, bins 0..M.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_row_r2c(KArgs a) {
  static_assert(TPB % 2 == 0, "rows come in slot pairs");
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  constexpr int HP = TPB / 2;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t m1 = a.n1, m2 = a.n2, half = m1 / 2;
  const int64_t bpm = (half + HP - 1) / HP;  // blocks per member
  const int64_t bi = blockIdx.x / bpm;
  const int64_t s0 = (blockIdx.x % bpm) * HP;
  // row of LDS slot gg (-1: past the member's slots)
  auto row_of = [&](int gg) -> int64_t {
    const int64_t s = s0 + (gg % HP);
    if (s >= half) return -1;
    if (gg < HP) return s;
    return s == 0 ? half : m1 - s;
  };
  const bool member = bi < a.count;
  const int64_t row = row_of(g);
  const bool valid = member && row >= 0;
  const int lper = Eng::lds_per(a.n);
  cpx<C> *lds_all = dyn_lds<C>();
  const cpx<C> *ws = static_cast<const cpx<C> *>(a.in) + bi * a.bigdist + (valid ? row * m2 : 0);
  Eng::template run<true>(
      a.desc, lds_all + g * lper, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return ws[e]; }, [&](int, cpx<C>) {});
  if (!member) return;
  // transposed store: thread -> LDS row c, k2 = tid / TPB + j*TPT; the low
  // rows of a block write ascending runs of HP bins, the partner rows
  // descending runs
  const int c = threadIdx.x % TPB;
  const int64_t r = row_of(c);
  if (r < 0) return;
  const int cp = r == 0 || r == half ? c : (c < HP ? c + HP : c - HP);
  const cpx<C> *src = lds_all + c * lper;
  const cpx<C> *par = lds_all + cp * lper;
  const int64_t M = m1 * m2, os = a.ostride;
  E *out = static_cast<E *>(a.out) + bi * a.odist;
  const cpx<C> *wn = static_cast<const cpx<C> *>(a.aux1);
  const int shift2 = a.shift2;
  const C h = (a.flags & 2) ? C(1) : C(0.5);
  for (int k2 = threadIdx.x / TPB; k2 < m2; k2 += TPT) {
    const int64_t k = r + m1 * k2;
    const int k2p = r == 0 ? (k2 == 0 ? 0 : (int)(m2 - k2)) : (int)(m2 - 1 - k2);
    const cpx<C> A = src[lds_pad<Eng::PAD>(k2)];
    if (k == 0) {
      out[0] = ST::st({A.x + A.y, C(0)});
      out[M * os] = ST::st({A.x - A.y, C(0)});
      continue;
    }
    const cpx<C> B = par[lds_pad<Eng::PAD>(k2p)];
    const cpx<C> xe = {h * (A.x + B.x), h * (A.y - B.y)};
    const cpx<C> xo = {h * (A.y + B.y), h * (B.x - A.x)};  // -i (A - conj B) h
    out[k * os] = ST::st(cadd(xe, cmul(packed_twiddle(wn, k, shift2), xo)));
  }
}

// ---------------------------------------------------------------------------
// C2R pass A: count = chunk * m2 columns; input bins 0..M at istride /
// idist (complex elements); the imaginary parts of X[0] and X[M] are
// ignored.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_col_c2r(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename Eng::C;
  const int g = threadIdx.x % TPB;
  const int t = threadIdx.x / TPB;
  const int64_t col = (int64_t)blockIdx.x * TPB + g;
  const bool valid = col < a.count;
  const int64_t bi = valid ? col / a.n2 : 0;
  const int64_t c2 = valid ? col % a.n2 : 0;
  const E *X = static_cast<const E *>(a.in) + bi * a.idist;
  cpx<C> *ws = static_cast<cpx<C> *>(a.out) + bi * a.bigdist + c2;
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(a.n);
  const int64_t n2 = a.n2, is = a.istride, M = a.n1 * a.n2;
  const cpx<C> *wm = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *wn = static_cast<const cpx<C> *>(a.aux1);
  const int shift = a.shift, shift2 = a.shift2;
  const C hw = (a.flags & 2) ? C(0.5) : C(1);  // adjoint of R2C: interior bins halved
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        const int64_t j = c2 + (int64_t)e * n2;
        cpx<C> A = ST::ld(X[j * is]);
        cpx<C> B = ST::ld(X[(M - j) * is]);
        if (j == 0) {
          A.y = C(0);
          B.y = C(0);
        } else {
          A = {hw * A.x, hw * A.y};
          B = {hw * B.x, hw * B.y};
        }
        const cpx<C> s = {A.x + B.x, A.y - B.y};  // A + conj B
        const cpx<C> d = {A.x - B.x, A.y + B.y};  // A - conj B
        const cpx<C> wd = cconj_mul(d, packed_twiddle(wn, j, shift2));  // W_n^{-j} d
        return cswap(cpx<C>{s.x - wd.y, s.y + wd.x});  // s + i wd, swapped: inverse DFT
      },
      [&](int k1, cpx<C> v) {
        ws[(int64_t)k1 * n2] = cmul(v, packed_twiddle(wm, c2 * (int64_t)k1, shift));
      });
}

// ---------------------------------------------------------------------------
// C2R pass B: count = chunk * m1 rows (row = bi*m1 + k1); output reals at
// odist / ostride (real elements): z[j] -> x[2j], x[2j+1], j = k1 + m1*k2.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_row_c2r(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t row0 = (int64_t)blockIdx.x * TPB;
  const int64_t row = row0 + g;
  const bool valid = row < a.count;
  const int lper = Eng::lds_per(a.n);
  cpx<C> *lds_all = dyn_lds<C>();
  const int64_t m1 = a.n1, m2 = a.n2;
  const cpx<C> *ws = static_cast<const cpx<C> *>(a.in) +
                     (valid ? (row / m1) * a.bigdist + (row % m1) * m2 : 0);
  Eng::template run<true>(
      a.desc, lds_all + g * lper, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return ws[e]; }, [&](int, cpx<C>) {});
  const int c = threadIdx.x % TPB;
  const int64_t rc = row0 + c;
  if (rc >= a.count) return;
  const int64_t os = a.ostride;
  const int64_t k1 = rc % m1;
  Rs *y = static_cast<Rs *>(a.out) + (rc / m1) * a.odist;
  const cpx<C> *src = lds_all + c * lper;
  const bool pair_ok = a.flags & 1;
  for (int k2 = threadIdx.x / TPB; k2 < m2; k2 += TPT) {
    const int64_t j = k1 + m1 * k2;
    const cpx<C> v = cswap(src[lds_pad<Eng::PAD>(k2)]);
    if (pair_ok) {
      reinterpret_cast<E *>(y)[j] = ST::st(v);
    } else {
      y[(2 * j) * os] = ST::str(v.x);
      y[(2 * j + 1) * os] = ST::str(v.y);
    }
  }
}

}  // namespace plf
