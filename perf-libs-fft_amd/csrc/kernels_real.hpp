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

#include <type_traits>

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
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  // flags bit1: interior bins doubled (the adjoint of C2R)
  const C h = (a.flags & 2) ? C(1) : C(0.5);
  // output bins k and m-k from Z[k], Z[m-k]; so = output stride (1 when
  // known at compile time)
  auto post = [&](int k, cpx<C> w, int64_t so) {
    if (k == 0) {
      const cpx<C> z0 = lds[lds_pad<Eng::PAD>(0)];
      out[0] = ST::st({z0.x + z0.y, C(0)});
      out[(int64_t)m * so] = ST::st({z0.x - z0.y, C(0)});
      return;
    }
    const cpx<C> A = lds[lds_pad<Eng::PAD>(k)];
    const cpx<C> B = lds[lds_pad<Eng::PAD>(m - k)];
    const cpx<C> xe = {h * (A.x + B.x), h * (A.y - B.y)};
    const cpx<C> xo = {h * (A.y + B.y), h * (B.x - A.x)};  // -i (A - conj B) / 2
    const cpx<C> wx = cmul(w, xo);
    out[(int64_t)k * so] = ST::st(cadd(xe, wx));
    if (2 * k != m) out[(int64_t)(m - k) * so] = ST::st({xe.x - wx.x, wx.y - xe.y});
  };
  constexpr int NF = Eng::NFIX;
  if constexpr (NF > 0) {
    // compile-time length: the post-twiddles are fetched before the FFT so
    // that their latency hides behind it, and the epilogue is unrolled.
    // Unit strides (paired real loads, unit output stride) take a copy
    // without stride arithmetic.
    constexpr int KP = (NF / 2 + TPT) / TPT;  // ceil((NF/2 + 1) / TPT)
    cpx<C> wk[KP];
#pragma unroll
    for (int i = 0; i < KP; ++i) {
      const int k = t + i * TPT;
      wk[i] = (2 * k <= NF) ? W[k] : cpx<C>{C(0), C(0)};
    }
    auto body = [&](auto unit) {
      constexpr bool U = decltype(unit)::value;
      Eng::template run<true>(
          a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
          [&](int e) {
            if constexpr (U) return ST::ld(reinterpret_cast<const E *>(x)[e]);
            else return load_pair<ST>(x, e, is, pair_ok);
          },
          [&](int, cpx<C>) {});
      if (!valid) return;
#pragma unroll
      for (int i = 0; i < KP; ++i) {
        const int k = t + i * TPT;
        if (2 * k <= NF) post(k, wk[i], U ? int64_t(1) : os);
      }
    };
    if (pair_ok && os == 1) body(std::true_type{});
    else body(std::false_type{});
  } else {
    Eng::template run<true>(
        a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
        [&](int e) { return load_pair<ST>(x, e, is, pair_ok); }, [&](int, cpx<C>) {});
    if (!valid) return;
    for (int k = t; 2 * k <= m; k += TPT) post(k, W[k], os);
  }
}

// ---------------------------------------------------------------------------
// R2C and C2R, even n = 2m, on the split (real, then imaginary) exchange
// engine with mirrored butterflies (StaticEngineSplit, MIRROR): the whole
// length-m complex DFT stays in registers and in half the LDS of the plain
// engine, so m = 32768 runs in one pass with 1024 threads (32 points each).
// The R2C's last stage and the C2R's first stage hand each thread the
// butterflies lo = t + q*TPT and NB - lo (NB/2 for lo = 0), so the
// Hermitian partners Z[e] and Z[m - e] of the post- and pre-twiddles are in
// the same thread:
//   e = lo + j*NB  <->  m - e = (NB - lo) + (R-1-j)*NB       (lo > 0)
//   e = j*NB       <->  m - e = ((R - j) mod R)*NB            (lo = 0)
//   e = NB/2 + j*NB <-> m - e = NB/2 + (R-1-j)*NB
// R2C computes X[e] and X[m-e] = conj(Xe - W Xo) from one pair; C2R loads
// each X[k] once and forms Z'[e] and Z'[m-e] from it. flags and aux0 as for
// k_r2c_even / k_c2r_even.
// ---------------------------------------------------------------------------
template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_r2c_even_split(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  using Ch = typename Eng::Ch;
  constexpr int TPT = Eng::TPT;
  using EM = StaticEngineSplit<C, TPT, Eng::PAD, Ch, 2>;  // last stage mirrored
  constexpr int M = Ch::N;
  constexpr int RL = Ch::radix(Ch::count - 1);
  constexpr int NB = M / RL;
  constexpr int KB = NB / TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(M);
  const int64_t is = a.istride, os = a.ostride;
  const bool pair_ok = a.flags & 1;
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  const C h = (a.flags & 2) ? C(1) : C(0.5);
  // post-twiddles of the low butterflies: W_n^{lo + j NB} = W_n^lo W_n^{j NB}
  // with W_n^{NB} = exp(-2 pi i / 2RL), so only W_n^lo is loaded
  cpx<C> wl[KB / 2];
#pragma unroll
  for (int q = 0; q < KB / 2; ++q) wl[q] = W[t + q * TPT];
  static_assert(2 * M == 2 * RL * NB, "n = 2 RL NB");
  cpx<C> z[KB * RL];
  EM::run_keep(lds, static_cast<const cpx<C> *>(a.tw), t, valid,
               [&](int e) { return load_pair<ST>(x, e, is, pair_ok); }, z);
  if (!valid) return;
  auto body = [&](auto unit) {
    constexpr bool U = decltype(unit)::value;  // unit output stride
    const int64_t so = U ? int64_t(1) : os;
    // X[e] = Xe + w Xo and X[m - e] = conj(Xe - w Xo) from A = Z[e], B = Z[m-e]
    auto pair = [&](int e, cpx<C> A, cpx<C> B, cpx<C> w, bool both) {
      const cpx<C> xe = {h * (A.x + B.x), h * (A.y - B.y)};
      const cpx<C> xo = {h * (A.y + B.y), h * (B.x - A.x)};  // -i (A - conj B) h
      const cpx<C> wx = cmul(w, xo);
      out[(int64_t)e * so] = ST::st(cadd(xe, wx));
      if (both) out[(int64_t)(M - e) * so] = ST::st({xe.x - wx.x, wx.y - xe.y});
    };
#pragma unroll
    for (int q = 0; q < KB / 2; ++q) {
      const int lo = t + q * TPT;
      const cpx<C> *A = z + (2 * q) * RL;      // butterfly lo (or 0)
      const cpx<C> *B = z + (2 * q + 1) * RL;  // butterfly NB - lo (or NB/2)
      if (lo != 0) {
        static_for<RL>([&](auto J) {
          constexpr int j = decltype(J)::value;
          pair(lo + j * NB, A[j], B[RL - 1 - j], mul_wc<2 * RL, j>(wl[q]), true);
        });
      } else {
        // butterfly 0: bins 0 and m from Z[0], pairs (j, R - j) for j <= R/2
        out[0] = ST::st({A[0].x + A[0].y, C(0)});
        out[(int64_t)M * so] = ST::st({A[0].x - A[0].y, C(0)});
        static_for<RL / 2 + 1>([&](auto J) {
          constexpr int j = decltype(J)::value;
          if constexpr (j >= 1)
            pair(j * NB, A[j], A[RL - j], root_c<2 * RL, j, C>(), 2 * j != RL);
        });
        // butterfly NB/2: pairs (j, R - 1 - j)
        static_for<RL / 2>([&](auto J) {
          constexpr int j = decltype(J)::value;
          pair(NB / 2 + j * NB, B[j], B[RL - 1 - j], root_c<4 * RL, 2 * j + 1, C>(), true);
        });
      }
    }
  };
  if (os == 1) body(std::true_type{});
  else body(std::false_type{});
}

template <typename S, bool UNUSED, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_c2r_even_split(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  using Ch = typename Eng::Ch;
  constexpr int TPT = Eng::TPT;
  using EM = StaticEngineSplit<C, TPT, Eng::PAD, Ch, 1>;  // first stage mirrored
  constexpr int M = Ch::N;
  constexpr int R0 = Ch::radix(0);
  constexpr int NB = M / R0;
  constexpr int KB = NB / TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const E *X = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(M);
  const int64_t is = a.istride, os = a.ostride;
  const bool pair_ok = a.flags & 1;
  const C hw = (a.flags & 2) ? C(0.5) : C(1);  // adjoint of R2C: interior bins halved
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  // Z'[e] from A = X[e], B = X[m-e] (both interior unless e == 0), w = W_n^e,
  // swapped for the inverse DFT
  auto pre = [&](cpx<C> A, cpx<C> B, cpx<C> w, bool e0) {
    if (e0) {
      A.y = C(0);
      B.y = C(0);
    } else {
      A = {hw * A.x, hw * A.y};
      B = {hw * B.x, hw * B.y};
    }
    const cpx<C> s = {A.x + B.x, A.y - B.y};     // A + conj B
    const cpx<C> d = {A.x - B.x, A.y + B.y};     // A - conj B
    const cpx<C> wd = cconj_mul(d, w);           // W^{-e} (A - conj B)
    return cswap(cpx<C>{s.x - wd.y, s.y + wd.x});  // s + i wd
  };
  cpx<C> vin[KB * R0];
  auto build = [&](auto unit) {
    constexpr bool U = decltype(unit)::value;  // unit input stride
    const int64_t si = U ? int64_t(1) : is;
    auto ld = [&](int64_t k) { return ST::ld(X[k * si]); };
#pragma unroll
    for (int q = 0; q < KB / 2; ++q) {
      const int lo = t + q * TPT;
      const int hi = lo == 0 ? NB / 2 : NB - lo;
      cpx<C> xl[R0], xh[R0];
      if (valid) {
#pragma unroll
        for (int j = 0; j < R0; ++j) {
          xl[j] = ld(lo + j * NB);
          xh[j] = ld(hi + j * NB);
        }
      } else {
#pragma unroll
        for (int j = 0; j < R0; ++j) xl[j] = xh[j] = cpx<C>{C(0), C(0)};
      }
      // each pair of loaded values feeds two outputs, so the inputs can be
      // overwritten as the outputs are formed
      // W_n^{b + j NB} = W_n^b exp(-2 pi i j / 2R0) (n = 2 R0 NB), and
      // W_n^{NB - lo} = exp(-2 pi i / 2R0) conj(W_n^lo): one table load per
      // butterfly pair
      cpx<C> *vl = vin + (2 * q) * R0, *vh = vin + (2 * q + 1) * R0;
      if (lo != 0) {
        const cpx<C> wlo = W[lo];
        const cpx<C> whi = mul_wc<2 * R0, 1>(cpx<C>{wlo.x, -wlo.y});
        static_for<R0>([&](auto J) {
          constexpr int j = decltype(J)::value;
          const cpx<C> A = xl[j], B = xh[R0 - 1 - j];
          vl[j] = pre(A, B, mul_wc<2 * R0, j>(wlo), false);
          vh[R0 - 1 - j] = pre(B, A, mul_wc<2 * R0, R0 - 1 - j>(whi), false);
        });
      } else {
        const cpx<C> xm = valid ? ld(M) : cpx<C>{C(0), C(0)};  // X[m] pairs with X[0]
        vl[0] = pre(xl[0], xm, cpx<C>{C(1), C(0)}, true);
        static_for<R0 / 2 + 1>([&](auto J) {
          constexpr int j = decltype(J)::value;
          if constexpr (j >= 1) {
            const cpx<C> A = xl[j], B = xl[R0 - j];
            vl[j] = pre(A, B, root_c<2 * R0, j, C>(), false);
            if constexpr (2 * j != R0) vl[R0 - j] = pre(B, A, root_c<2 * R0, R0 - j, C>(), false);
          }
        });
        static_for<R0 / 2>([&](auto J) {
          constexpr int j = decltype(J)::value;
          const cpx<C> A = xh[j], B = xh[R0 - 1 - j];
          vh[j] = pre(A, B, root_c<4 * R0, 2 * j + 1, C>(), false);
          vh[R0 - 1 - j] = pre(B, A, root_c<4 * R0, 2 * (R0 - 1 - j) + 1, C>(), false);
        });
      }
    }
  };
  if (is == 1) build(std::true_type{});
  else build(std::false_type{});
  EM::run_regs(
      vin, lds, static_cast<const cpx<C> *>(a.tw), t,
      [&](int e, cpx<C> v) {
        v = cswap(v);
        if (pair_ok) {
          ST::stnt(reinterpret_cast<E *>(y) + e, v);
        } else {
          y[(int64_t)(2 * e) * os] = ST::str(v.x);
          y[(int64_t)(2 * e + 1) * os] = ST::str(v.y);
        }
      },
      valid);
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
  const C hw = (a.flags & 2) ? C(0.5) : C(1);  // adjoint of R2C: interior bins halved
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  auto store = [&](int e, cpx<C> v) {
    v = cswap(v);
    if (pair_ok) {
      ST::stnt(reinterpret_cast<E *>(y) + e, v);  // streaming: written once, not re-read
    } else {
      y[(int64_t)(2 * e) * os] = ST::str(v.x);
      y[(int64_t)(2 * e + 1) * os] = ST::str(v.y);
    }
  };
  // Z'[e] for stage 0, w = W_n^e; si = input stride (1 when known at
  // compile time)
  auto pre = [&](int e, cpx<C> w, int64_t si) {
    cpx<C> A = ST::ld(X[(int64_t)e * si]);
    cpx<C> B = ST::ld(X[(int64_t)(m - e) * si]);
    if (e == 0) {
      A.y = C(0);
      B.y = C(0);
    } else {
      A = {hw * A.x, hw * A.y};
      B = {hw * B.x, hw * B.y};
    }
    const cpx<C> s = {A.x + B.x, A.y - B.y};     // A + conj B
    const cpx<C> d = {A.x - B.x, A.y + B.y};     // A - conj B
    const cpx<C> wd = cconj_mul(d, w);           // W^{-e} (A - conj B)
    const cpx<C> zp = {s.x - wd.y, s.y + wd.x};  // s + i wd
    return cswap(zp);                            // inverse DFT via swap
  };
  constexpr int NF = Eng::NFIX;
  if constexpr (NF > 0 && Eng::REGS_IN) {
    // compile-time length: thread t feeds the stage-0 butterflies
    // b = t + k TPT with elements e = b + j NB (NB = m / R0), and
    // W_n^e = W_n^b W_n^{j NB}: one per-thread twiddle load and R0 loads at
    // block-uniform addresses instead of one per-lane load per element
    using Ch = typename Eng::Ch;
    constexpr int R0 = Ch::radix(0);
    constexpr int NB = NF / R0;
    constexpr int KB = (NB + TPT - 1) / TPT;
    cpx<C> vin[KB * R0];
    auto build = [&](auto unit) {
      constexpr bool U = decltype(unit)::value;  // unit input stride
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int b = t + k * TPT;
        const bool act = valid && ((NB % TPT == 0) || (b < NB));
        const cpx<C> wb = act ? W[b] : cpx<C>{C(1), C(0)};
#pragma unroll
        for (int j = 0; j < R0; ++j)
          vin[k * R0 + j] = act ? pre(b + j * NB, j == 0 ? wb : cmul(wb, W[j * NB]),
                                      U ? int64_t(1) : is)
                                : cpx<C>{C(0), C(0)};
      }
    };
    if (is == 1) build(std::true_type{});
    else build(std::false_type{});
    Eng::E::run_regs(vin, lds, static_cast<const cpx<C> *>(a.tw), t, store, valid);
  } else {
    Eng::template run<false>(
        a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
        [&](int e) { return pre(e, W[e], is); }, store);
  }
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
        if (2 * e <= n) {
          if ((a.flags & 2) && e > 0 && 2 * e != n) v = {C(2) * v.x, C(2) * v.y};
          out[(int64_t)e * os] = ST::st(v);
        }
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
        } else {
          v = ST::ld(X[(int64_t)(n - e) * is]);
          v.y = -v.y;
        }
        if (e == 0 || 2 * e == n) v.y = C(0);
        else if (a.flags & 2) v = {C(0.5) * v.x, C(0.5) * v.y};
        return cswap(v);
      },
      [&](int e, cpx<C> v) { y[(int64_t)e * os] = ST::str(v.y); });  // Re of swap(v)
}

// ---------------------------------------------------------------------------
// DCT-II / DST-II (FFTW REDFT10 / RODFT10), following the reference's
// Makhoul steps (src/r2r_buffer_helper.hpp:318-335 and :457-470 for the input
// permutation, :688-701 and :830-857 for the post-processing):
//   v[i] = x[2i] (i < ceil(n/2)),  v[i] = x[2(n-i)-1] (otherwise),
//   V = R2C_n(v),  y_k = 2 Re(w_k V_k),  y_{n-k} = -2 Im(w_k V_k),
//   w_k = exp(-i pi k / 2n).
// DST-II is the same with the odd-sourced v[i] negated and the output
// reversed (y_{n-1-k}). flags bit1 scales the output of internal index 0 by
// 1/2 (the adjoint of DCT-III / DST-III).
// Even n uses the packed half-length engine (m = n/2): aux0 = W_n^k
// (k <= m/2), aux1 = w_k (k <= m). Odd n runs the full-length engine with
// aux1 = w_k (k <= n/2).
// ---------------------------------------------------------------------------
template <typename ST>
__device__ __forceinline__ typename ST::C dct2_src(const typename ST::Rs *x, int i, int n,
                                                  int64_t is, bool dst) {
  using C = typename ST::C;
  if (2 * i < n) return ST::ldr(x[(int64_t)(2 * i) * is]);
  const C v = ST::ldr(x[(int64_t)(2 * (n - i) - 1) * is]);
  return dst ? -v : v;
}

// Thread t's share of for (o = 0; o < len; ++o): first loads every value
// (ld(o)), then consumes them (use(o, v)). With a compile-time length
// (LEN > 0) the loop is unrolled, so all of the thread's loads are in flight
// together.
template <int TPT, int LEN, class V, class Ld, class Use>
__device__ __forceinline__ void gather_then(int t, int len, Ld &&ld, Use &&use) {
  if constexpr (LEN > 0) {
    constexpr int IT = (LEN + TPT - 1) / TPT;
    V v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int o = t + i * TPT;
      if (LEN % TPT == 0 || o < LEN) v[i] = ld(o);
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const int o = t + i * TPT;
      if (LEN % TPT == 0 || o < LEN) use(o, v[i]);
    }
  } else {
    for (int o = t; o < len; o += TPT) use(o, ld(o));
  }
}

template <typename S, bool DST, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_dct2_even(KArgs a) {
  using ST = StorageT<S>;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int m = (int)a.n;
  const int n = 2 * m;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(m);
  const int64_t is = a.istride, os = a.ostride;
  // Coalesced loads of the member's reals, scattered into LDS as the packed
  // sequence z_e = u_{2e} + i u_{2e+1}: y-index o holds u_{o/2} (o even) or
  // u_{n-(o+1)/2} (o odd, negated for DST).
  constexpr int NF = Eng::NFIX;
  const int nn = NF > 0 ? 2 * NF : n;
  if (valid) {
    C *lr = reinterpret_cast<C *>(lds);
    auto put = [&](int o, C v) {
      const int i = (o & 1) ? n - (o + 1) / 2 : o / 2;
      lr[2 * lds_pad<Eng::PAD>(i >> 1) + (i & 1)] = (DST && (o & 1)) ? -v : v;
    };
    if (is == 1) {
      // unit stride: every load of the thread in flight at once, addressed
      // from one base
      gather_then<TPT, 2 * NF, C>(t, nn, [&](int o) { return ST::ldr(x[o]); }, put);
    } else {
#pragma unroll 8
      for (int o = t; o < nn; o += TPT) put(o, ST::ldr(x[(int64_t)o * is]));
    }
  }
  __syncthreads();
  Eng::template run<true, true>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int) { return cpx<C>{C(0), C(0)}; }, [&](int, cpx<C>) {});
  if (!valid) return;
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *w = static_cast<const cpx<C> *>(a.aux1);
  const C h = C(0.5), two = C(2);
  const C s0 = (a.flags & 2) ? C(0.5) : C(1);
  auto put = [&](int o, C v) { y[(int64_t)(DST ? n - 1 - o : o) * os] = ST::str(v); };
  // y_k and y_{n-k} from V_k
  auto emit = [&](int k, cpx<C> V) {
    const cpx<C> r = cmul(w[k], V);
    put(k, two * r.x);
    put(n - k, -two * r.y);
  };
  for (int k = t; 2 * k <= m; k += TPT) {
    if (k == 0) {
      const cpx<C> z0 = lds[lds_pad<Eng::PAD>(0)];
      put(0, s0 * two * (z0.x + z0.y));
      const C vm = z0.x - z0.y;  // V_m is real
      put(m, two * vm * w[m].x);
      continue;
    }
    const cpx<C> A = lds[lds_pad<Eng::PAD>(k)];
    const cpx<C> B = lds[lds_pad<Eng::PAD>(m - k)];
    const cpx<C> xe = {h * (A.x + B.x), h * (A.y - B.y)};
    const cpx<C> xo = {h * (A.y + B.y), h * (B.x - A.x)};
    const cpx<C> wx = cmul(W[k], xo);
    emit(k, cadd(xe, wx));
    if (2 * k != m) emit(m - k, {xe.x - wx.x, wx.y - xe.y});
  }
}

template <typename S, bool DST, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_dct2_full(KArgs a) {
  using ST = StorageT<S>;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int n = (int)a.n;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(n);
  const int64_t is = a.istride, os = a.ostride;
  Eng::template run<true>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int e) { return cpx<C>{dct2_src<ST>(x, e, n, is, DST), C(0)}; },
      [&](int, cpx<C>) {});
  if (!valid) return;
  const cpx<C> *w = static_cast<const cpx<C> *>(a.aux1);
  const C two = C(2);
  const C s0 = (a.flags & 2) ? C(0.5) : C(1);
  auto put = [&](int o, C v) { y[(int64_t)(DST ? n - 1 - o : o) * os] = ST::str(v); };
  for (int k = t; 2 * k <= n; k += TPT) {
    const cpx<C> r = cmul(w[k], lds[lds_pad<Eng::PAD>(k)]);
    if (k == 0) {
      put(0, s0 * two * r.x);
    } else {
      put(k, two * r.x);
      if (2 * k != n) put(n - k, -two * r.y);
    }
  }
}

// ---------------------------------------------------------------------------
// DCT-III / DST-III (REDFT01 / RODFT01): the inverse Makhoul steps,
//   X_k = (u_k - i u_{n-k}) conj(w_k)  (u_n = 0),  v = C2R_n(X),
//   y[2i] = v_i (i < ceil(n/2)),  y[2(n-i)-1] = v_i (otherwise).
// DST-III(u)_j = (-1)^j DCT-III(reverse u)_j. flags bit1 doubles the input
// of internal index 0 (the adjoint of DCT-II / DST-II).
// Even n: packed half-length engine, aux0 = W_n^k (k < m), aux1 = w_k (k <= m).
// Odd n: full-length engine, aux1 = w_k (k < n).
// ---------------------------------------------------------------------------
template <typename ST, bool DST>
__device__ __forceinline__ typename ST::C dct3_in(const typename ST::Rs *x, int k, int n,
                                                 int64_t is, typename ST::C s0) {
  using C = typename ST::C;
  if (k >= n) return C(0);
  const int src = DST ? n - 1 - k : k;
  C v = ST::ldr(x[(int64_t)src * is]);
  return k == 0 ? s0 * v : v;
}

template <typename S, bool DST, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_dct3_even(KArgs a) {
  using ST = StorageT<S>;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int m = (int)a.n;
  const int n = 2 * m;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(m);
  const int64_t is = a.istride, os = a.ostride;
  const cpx<C> *W = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *w = static_cast<const cpx<C> *>(a.aux1);
  const C s0 = (a.flags & 2) ? C(2) : C(1);
  // stage-0 input e from X(e) and X(m-e), w = W_n^e
  auto z = [&](cpx<C> A, cpx<C> B, cpx<C> w) {
    const cpx<C> sm = {A.x + B.x, A.y - B.y};
    const cpx<C> d = {A.x - B.x, A.y + B.y};
    const cpx<C> wd = cconj_mul(d, w);
    return cswap(cpx<C>{sm.x - wd.y, sm.y + wd.x});
  };
  constexpr int NF = Eng::NFIX;
  if constexpr (NF > 0 && Eng::REGS_IN) {
    // compile-time length: the n reals u are staged in LDS with coalesced
    // loads; thread t then builds the inputs of its stage-0 butterflies
    // b = t + k TPT, elements e = b + j NB (NB = m / R0), with
    // W_n^e = W_n^b W_n^{j NB}, w_e = w_b w_{j NB} and w_{m-e} = w_m conj(w_e)
    C *lu = reinterpret_cast<C *>(lds);
    if (valid) {
      if (is == 1) {  // unit stride: loads addressed from one base
        gather_then<TPT, 2 * NF, C>(
            t, 2 * NF, [&](int k) { return ST::ldr(x[DST ? n - 1 - k : k]); },
            [&](int k, C v) { lu[k] = k == 0 ? s0 * v : v; });
      } else {
        gather_then<TPT, 2 * NF, C>(
            t, 2 * NF, [&](int k) { return dct3_in<ST, DST>(x, k, n, is, s0); },
            [&](int k, C v) { lu[k] = v; });
      }
    }
    __syncthreads();
    using Ch = typename Eng::Ch;
    constexpr int R0 = Ch::radix(0);
    constexpr int NB = NF / R0;
    constexpr int KB = (NB + TPT - 1) / TPT;
    const cpx<C> wm = w[NF];
    cpx<C> vin[KB * R0];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int b = t + k * TPT;
      const bool act = valid && ((NB % TPT == 0) || (b < NB));
      const cpx<C> Wb = act ? W[b] : cpx<C>{C(1), C(0)};
      const cpx<C> wb = act ? w[b] : cpx<C>{C(1), C(0)};
#pragma unroll
      for (int j = 0; j < R0; ++j) {
        const int e = b + j * NB;
        const cpx<C> We = j == 0 ? Wb : cmul(Wb, W[j * NB]);
        const cpx<C> we = j == 0 ? wb : cmul(wb, w[j * NB]);
        const cpx<C> wme = cconj_mul(wm, we);
        if (act) {
          const cpx<C> A = cconj_mul(cpx<C>{lu[e], e == 0 ? C(0) : -lu[n - e]}, we);
          const cpx<C> B = cconj_mul(cpx<C>{lu[NF - e], -lu[NF + e]}, wme);
          vin[k * R0 + j] = z(A, B, We);
        } else {
          vin[k * R0 + j] = cpx<C>{C(0), C(0)};
        }
      }
    }
    __syncthreads();  // u is read; stage 0 overwrites the LDS
    Eng::E::template run_regs<true>(vin, lds, static_cast<const cpx<C> *>(a.tw), t,
                                    [&](int, cpx<C>) {}, valid);
  } else {
    auto X = [&](int k) {  // (u_k - i u_{n-k}) * conj(w_k), k in [0, m]
      const C ur = dct3_in<ST, DST>(x, k, n, is, s0);
      const C ui = k == 0 ? C(0) : -dct3_in<ST, DST>(x, n - k, n, is, s0);
      return cconj_mul(cpx<C>{ur, ui}, w[k]);
    };
    Eng::template run<true>(
        a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
        [&](int e) { return z(X(e), X(m - e), W[e]); }, [&](int, cpx<C>) {});
  }
  if (!valid) return;
  // swap(lds[e]) = v_{2e} + i v_{2e+1}; coalesced stores y[o] = v_i with
  // i = o/2 (o even) or n - (o+1)/2 (o odd, negated for DST)
  const int nn = NF > 0 ? 2 * NF : n;
  const C *lr = reinterpret_cast<const C *>(lds);
  auto get = [&](int o) {
    const int i = (o & 1) ? n - (o + 1) / 2 : o / 2;
    return lr[2 * lds_pad<Eng::PAD>(i >> 1) + 1 - (i & 1)];
  };
  if (os == 1)
    gather_then<TPT, 2 * NF, C>(t, nn, get,
                                [&](int o, C v) { y[o] = ST::str((DST && (o & 1)) ? -v : v); });
  else
    gather_then<TPT, 2 * NF, C>(
        t, nn, get, [&](int o, C v) { y[(int64_t)o * os] = ST::str((DST && (o & 1)) ? -v : v); });
}

template <typename S, bool DST, int TPB, class Eng>
__global__ void __launch_bounds__(Eng::TPT *TPB) k_dct3_full(KArgs a) {
  using ST = StorageT<S>;
  using Rs = typename ST::Rs;
  using C = typename Eng::C;
  constexpr int TPT = Eng::TPT;
  const int g = threadIdx.x / TPT;
  const int t = threadIdx.x % TPT;
  const int64_t tr = (int64_t)blockIdx.x * TPB + g;
  const bool valid = tr < a.count;
  const int n = (int)a.n;
  const Rs *x = static_cast<const Rs *>(a.in) + (valid ? tr * a.idist : 0);
  Rs *y = static_cast<Rs *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(n);
  const int64_t is = a.istride, os = a.ostride;
  const cpx<C> *w = static_cast<const cpx<C> *>(a.aux1);
  const C s0 = (a.flags & 2) ? C(2) : C(1);
  Eng::template run<false>(
      a.desc, lds, static_cast<const cpx<C> *>(a.tw), t, valid,
      [&](int k) {
        const C ur = dct3_in<ST, DST>(x, k, n, is, s0);
        const C ui = k == 0 ? C(0) : -dct3_in<ST, DST>(x, n - k, n, is, s0);
        return cswap(cconj_mul(cpx<C>{ur, ui}, w[k]));
      },
      [&](int i, cpx<C> v) {
        const int o = 2 * i < n ? 2 * i : 2 * (n - i) - 1;
        const C r = v.y;  // Re of swap(v)
        y[(int64_t)o * os] = ST::str((DST && (o & 1)) ? -r : r);
      });
}

}  // namespace plf
