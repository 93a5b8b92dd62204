// kernels_real.hpp -- real-data transforms on the Stockham engines.
//
// Even n: the n reals of a member are read as m = n/2 complex values
// z[j] = x[2j] + i x[2j+1] and one length-m complex DFT runs on them.
//   R2C post-twiddle (k = 0..m/2, pairs k and m-k in one thread):
//     Xe = (Z[k] + conj Z[m-k]) / 2,  Xo = -i (Z[k] - conj Z[m-k]) / 2
//     X[k] = Xe + W_n^k Xo,  X[m-k] = conj(Xe - W_n^k Xo)
//     X[0] = Re Z0 + Im Z0,  X[m] = Re Z0 - Im Z0
//   C2R pre-twiddle (k = 0..m-1, Im X[0] and Im X[m] ignored):
//     Z'[k] = (X[k] + conj X[m-k]) + i W_n^{-k} (X[k] - conj X[m-k])
//     z = unnormalized inverse DFT_m(Z') = n (x_even + i x_odd)
// Odd n: a length-n complex DFT of (x, 0), keeping bins 0..n/2 (R2C), or
// of the Hermitian extension X[n-k] = conj X[k], keeping the real part (C2R).
//
// This replaces the reference's Hermitian-folded codelet families
// (ab_nfoh / tfj / tfol for R2C, nbih / tbil / tbih for C2R;
// src/planner/strategies/cooley_tukey.hpp:145-323, 402-580).
// Real-side strides count reals, complex-side strides count complex values
// (docs/plfft.md:136-170).
#pragma once

#include "kernels.hpp"

namespace plf {

// Load two consecutive reals as one complex value when that is legal.
template <typename ST>
__device__ __forceinline__ cpx<typename ST::C> load_pair(const typename ST::Rs *x, int64_t j,
                                                         int64_t is, bool pair_ok) {
  using E = typename ST::E;
  if (pair_ok) return ST::ld(reinterpret_cast<const E *>(x)[j]);
  return {ST::ldr(x[(2 * j) * is]), ST::ldr(x[(2 * j + 1) * is])};
}

// ---------------------------------------------------------------------------
// R2C, even n = 2m. flags bit0: reals may be read in pairs (unit stride,
// even distance). aux0: W_n^k, k = 0..m/2 (forward sign).
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_r2c_even(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int m = (int)a.n;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(m);
  const int64_t is = a.istride, os = a.ostride;
  const bool pair_ok = a.flags & 1;
  Eng::template run<true>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return load_pair<ST>(x, e, is, pair_ok); }, [&](int, cpx<C>) {});
  if (!valid) return;
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  const C h = C(0.5);
  for (int k = t; 2 * k <= m; k += TPT) {
    if (k == 0) {
      const cpx<C> z0 = lds[lds_pad<Eng::PAD>(0)];
      out[0] = ST::st({z0.x + z0.y, C(0)});
      out[(int64_t)m * os] = ST::st({z0.x - z0.y, C(0)});
      continue;
    }
    const cpx<C> A = lds[lds_pad<Eng::PAD>(k)];
    const cpx<C> B = lds[lds_pad<Eng::PAD>(m - k)];
    const cpx<C> xe = {h * (A.x + B.x), h * (A.y - B.y)};
    const cpx<C> xo = {h * (A.y + B.y), h * (B.x - A.x)};  // -i (A - conj B) / 2
    const cpx<C> wx = cmul(W[k], xo);
    out[(int64_t)k * os] = ST::st(cadd(xe, wx));
    if (2 * k != m) out[(int64_t)(m - k) * os] = ST::st({xe.x - wx.x, wx.y - xe.y});
  }
}

// ---------------------------------------------------------------------------
// C2R, even n = 2m. flags bit0: reals may be written in pairs.
// aux0: W_n^k, k = 0..m-1 (forward sign; conjugated here).
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_c2r_even(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int m = (int)a.n;
  const E *X = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(m);
  const int64_t is = a.istride, os = a.ostride;
  const bool pair_ok = a.flags & 1;
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        cpx<C> A = ST::ld(X[(int64_t)e * is]);
        cpx<C> B = ST::ld(X[(int64_t)(m - e) * is]);
        if (e == 0) {
          A.y = C(0);
          B.y = C(0);
        }
        const cpx<C> s = {A.x + B.x, A.y - B.y};   // A + conj B
        const cpx<C> d = {A.x - B.x, A.y + B.y};   // A - conj B
        const cpx<C> wd = cconj_mul(d, W[e]);      // W^{-e} (A - conj B)
        const cpx<C> zp = {s.x - wd.y, s.y + wd.x};  // s + i wd
        return cswap(zp);                           // inverse DFT via swap
      },
      [&](int e, cpx<C> v) {
        v = cswap(v);
        if (pair_ok) {
          reinterpret_cast<E *>(y)[e] = ST::st(v);
        } else {
          y[(int64_t)(2 * e) * os] = ST::str(v.x);
          y[(int64_t)(2 * e + 1) * os] = ST::str(v.y);
        }
      });
}

// ---------------------------------------------------------------------------
// R2C, any n, full-length complex DFT of (x, 0); bins 0..n/2 are stored.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_r2c_full(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int n = (int)a.n;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(n);
  const int64_t is = a.istride, os = a.ostride;
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return cpx<C>{ST::ldr(x[(int64_t)e * is]), C(0)}; },
      [&](int e, cpx<C> v) {
        if (2 * e <= n) out[(int64_t)e * os] = ST::st(v);
      });
}

// ---------------------------------------------------------------------------
// C2R, any n, full-length inverse DFT of the Hermitian extension.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_c2r_full(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int n = (int)a.n;
  const E *X = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(n);
  const int64_t is = a.istride, os = a.ostride;
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) {
        cpx<C> v;
        if (2 * e <= n) {
          v = ST::ld(X[(int64_t)e * is]);
          if (e == 0 || 2 * e == n) v.y = C(0);
        } else {
          v = ST::ld(X[(int64_t)(n - e) * is]);
          v.y = -v.y;
        }
        return cswap(v);
      },
      [&](int e, cpx<C> v) { y[(int64_t)e * os] = ST::str(v.y); });  // Re of swap(v)
}

}  // namespace plf
