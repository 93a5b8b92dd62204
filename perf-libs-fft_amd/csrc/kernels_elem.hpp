// kernels_elem.hpp -- element-wise staging kernels around C2C plans:
// chirp-z pre/post for long prime lengths, real<->complex staging for real
// transforms without a real kernel, and Makhoul pre/post for DCT/DST
// lengths without a DCT kernel.
#pragma once

#include "fft_device.hpp"
#include "kernel_args.hpp"

namespace plf {

// ---------------------------------------------------------------------------
// element-wise helpers for chirp-z over lengths beyond one LDS tile
// ---------------------------------------------------------------------------
// ws[b*M + e] = e < n ? x[b*idist + e*istride] .* chirp[e] : 0 (swap if BWD)
template <typename S, bool BWD>
__global__ void __launch_bounds__(256) k_chirp_in(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename ST::C;
  const int64_t M = a.n, n = a.n1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * M) return;
  const int64_t b = i / M, e = i % M;
  cpx<C> v = {C(0), C(0)};
  if (e < n) {
    v = ST::ld(static_cast<const E *>(a.in)[b * a.idist + e * a.istride]);
    if (BWD) v = cswap(v);
    v = cmul(v, static_cast<const cpx<C> *>(a.aux1)[e]);
  }
  static_cast<cpx<C> *>(a.out)[i] = v;
}

// ws[b*M + e] = swap(ws[b*M + e] .* B[e])
template <typename C>
__global__ void __launch_bounds__(256) k_pointwise_swap(KArgs a) {
  const int64_t M = a.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * M) return;
  cpx<C> *w = static_cast<cpx<C> *>(a.out);
  w[i] = cswap(cmul(w[i], static_cast<const cpx<C> *>(a.aux0)[i % M]));
}

// y[b*odist + e*ostride] = swap(ws[b*M + e]) .* chirp[e], e < n (swap if BWD)
template <typename S, bool BWD>
__global__ void __launch_bounds__(256) k_chirp_out(KArgs a) {
  using ST = StorageT<S>;
  using E = typename ST::E;
  using C = typename ST::C;
  const int64_t M = a.n, n = a.n1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, e = i % n;
  cpx<C> v = cmul(cswap(static_cast<const cpx<C> *>(a.in)[b * M + e]),
                  static_cast<const cpx<C> *>(a.aux1)[e]);
  static_cast<E *>(a.out)[b * a.odist + e * a.ostride] = ST::st(BWD ? cswap(v) : v);
}

// ---------------------------------------------------------------------------
// real <-> complex staging for real transforms routed through a C2C plan
// ---------------------------------------------------------------------------
// ws[b*n + e] = (x[b*idist + e*istride], 0)
template <typename S>
__global__ void __launch_bounds__(256) k_widen(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  const int64_t n = a.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, e = i % n;
  const C v = ST::ldr(static_cast<const typename ST::Rs *>(a.in)[b * a.idist + e * a.istride]);
  static_cast<cpx<C> *>(a.out)[i] = {v, C(0)};
}

// out[b*odist + k*ostride] = ws[b*n + k] * w_k, k <= n/2; flags bit1 doubles
// the interior bins (adjoint of C2R)
template <typename S>
__global__ void __launch_bounds__(256) k_narrow(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  const int64_t n = a.n, h = n / 2 + 1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * h) return;
  const int64_t b = i / h, k = i % h;
  cpx<C> v = static_cast<const cpx<C> *>(a.in)[b * n + k];
  if ((a.flags & 2) && k > 0 && 2 * k != n) v = {C(2) * v.x, C(2) * v.y};
  static_cast<typename ST::E *>(a.out)[b * a.odist + k * a.ostride] = ST::st(v);
}

// ws[b*n + e] = swap(Hermitian extension of X), Im X[0] (and Im X[n/2]) dropped;
// flags bit1 halves the interior bins (adjoint of R2C)
template <typename S>
__global__ void __launch_bounds__(256) k_hermitian(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  const int64_t n = a.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, e = i % n;
  const typename ST::E *X = static_cast<const typename ST::E *>(a.in) + b * a.idist;
  const int64_t k = 2 * e <= n ? e : n - e;
  cpx<C> v = ST::ld(X[k * a.istride]);
  if (k == 0 || 2 * k == n) v.y = C(0);
  else if (a.flags & 2) v = {C(0.5) * v.x, C(0.5) * v.y};
  if (2 * e > n) v.y = -v.y;
  static_cast<cpx<C> *>(a.out)[i] = cswap(v);
}

// y[b*odist + e*ostride] = Re swap(ws[b*n + e])
template <typename S>
__global__ void __launch_bounds__(256) k_realpart(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  const int64_t n = a.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, e = i % n;
  static_cast<typename ST::Rs *>(a.out)[b * a.odist + e * a.ostride] =
      ST::str(static_cast<const cpx<C> *>(a.in)[i].y);
}

// ---------------------------------------------------------------------------
// DCT/DST II/III through a C2C plan of the same length (lengths without a
// DCT kernel). aux1 = w_k = exp(-i pi k / 2n), k < n. flags: bit1 adjoint
// weight, bit2 DST, bit3 type III.
// ---------------------------------------------------------------------------
// type II pre: ws[b*n + i] = (v_i, 0), v = Makhoul permutation (DST: odd
// sources negated); type III pre: ws = swap((u_k - i u_{n-k}) conj(w_k)),
// DST-III reads u reversed, adjoint weight doubles internal u_0
template <typename S>
__global__ void __launch_bounds__(256) k_dct_pre(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  using Rs = typename ST::Rs;
  const int64_t n = a.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, e = i % n;
  const Rs *x = static_cast<const Rs *>(a.in) + b * a.idist;
  const bool dst = a.flags & 4, three = a.flags & 8;
  cpx<C> out;
  if (!three) {
    C v;
    if (2 * e < n) v = ST::ldr(x[(2 * e) * a.istride]);
    else {
      v = ST::ldr(x[(2 * (n - e) - 1) * a.istride]);
      if (dst) v = -v;
    }
    out = {v, C(0)};
  } else {
    const C s0 = (a.flags & 2) ? C(2) : C(1);
    auto u = [&](int64_t k) -> C {
      if (k >= n) return C(0);
      C v = ST::ldr(x[(dst ? n - 1 - k : k) * a.istride]);
      return k == 0 ? s0 * v : v;
    };
    const C ur = u(e), ui = e == 0 ? C(0) : -u(n - e);
    out = cswap(cconj_mul(cpx<C>{ur, ui}, static_cast<const cpx<C> *>(a.aux1)[e]));
  }
  static_cast<cpx<C> *>(a.out)[i] = out;
}

// type II post: y_k = 2 Re(w_k V_k), y_{n-k} = -2 Im(w_k V_k) (DST: output
// reversed; adjoint weight halves internal y_0); type III post:
// y[perm(i)] = Re swap(v_i), DST-III negates odd outputs. One thread per
// (member, i).
template <typename S>
__global__ void __launch_bounds__(256) k_dct_post(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  using Rs = typename ST::Rs;
  const int64_t n = a.n;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, e = i % n;
  const cpx<C> *V = static_cast<const cpx<C> *>(a.in) + b * n;
  Rs *y = static_cast<Rs *>(a.out) + b * a.odist;
  const bool dst = a.flags & 4, three = a.flags & 8;
  if (!three) {
    // output index o = e; from V_k with k = e (real part) or k = n - e (imag)
    const cpx<C> *w = static_cast<const cpx<C> *>(a.aux1);
    C val;
    if (2 * e <= n) {
      val = C(2) * cmul(w[e], V[e]).x;
      if (e == 0 && (a.flags & 2)) val *= C(0.5);
    } else {
      val = -C(2) * cmul(w[n - e], V[n - e]).y;
    }
    y[(dst ? n - 1 - e : e) * a.ostride] = ST::str(val);
  } else {
    const int64_t o = 2 * e < n ? 2 * e : 2 * (n - e) - 1;
    const C r = V[e].y;
    y[o * a.ostride] = ST::str((dst && (o & 1)) ? -r : r);
  }
}

// ---------------------------------------------------------------------------
// Remaining R2R kinds (FFTW definitions, docs/plfft.md:186-230) through one
// complex DFT of length L per member: a pre kernel builds the complex input
// in a staging buffer (count x L, compute precision), a post kernel maps the
// transform to the real output. KArgs: n = logical length, n1 = L,
// flags = r2r kind | 256 (adjoint weights), aux0 / aux1 = pre / post
// rotation tables (DCT-IV / DST-IV), strides and distances of the real
// sides.
//   DCT-I  (0): L = 2(n-1), even extension;           y_k = Re V_k
//   DCT-IV (3): L = 2n, x_j e^{-i pi j/2n}, zero pad;  y_k = 2 Re(q_k V_k)
//   DST-I  (4): L = 2(n+1), odd extension;            y_k = -Im V_{k+1}
//   DST-IV (7): as DCT-IV;                             y_k = -2 Im(q_k V_k)
//   DHT    (8): L = n, (x, 0);                         y_k = Re V_k - Im V_k
//   R2HC   (9): L = n, (x, 0);   y_k = Re V_k (k <= n/2), Im V_{n-k} (k > n/2)
//   HC2R  (10): L = n, Hermitian unpack, backward DFT; y_j = Re v_j
// with q_k = e^{-i pi (2k+1)/4n}. Adjoint weights: DCT-I halves interior
// inputs and doubles interior outputs; HC2R (the adjoint of R2HC) halves
// interior inputs; R2HC (the adjoint of HC2R) doubles interior outputs.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_r2r_pre(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  using Rs = typename ST::Rs;
  const int64_t n = a.n, L = a.n1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * L) return;
  const int64_t b = i / L, j = i % L;
  const Rs *x = static_cast<const Rs *>(a.in) + b * a.idist;
  const int kind = a.flags & 255;
  const bool adj = a.flags & 256;
  auto ld = [&](int64_t k) { return ST::ldr(x[k * a.istride]); };
  cpx<C> v = {C(0), C(0)};
  switch (kind) {
    case 0: {  // DCT-I
      const int64_t k = j < n ? j : L - j;
      v.x = ld(k);
      if (adj && k != 0 && k != n - 1) v.x *= C(0.5);
      break;
    }
    case 3:
    case 7:  // DCT-IV / DST-IV
      if (j < n) {
        const C xv = ld(j);
        const cpx<C> p = static_cast<const cpx<C> *>(a.aux0)[j];
        v = {xv * p.x, xv * p.y};
      }
      break;
    case 4:  // DST-I
      if (j >= 1 && j <= n) v.x = ld(j - 1);
      else if (j >= n + 2) v.x = -ld(2 * n + 1 - j);
      break;
    case 10: {  // HC2R: X_k = (h_k, h_{n-k}), X_{n-k} = conj X_k
      const C w = (adj && j != 0 && 2 * j != n) ? C(0.5) : C(1);
      if (j == 0) v = {ld(0), C(0)};
      else if (2 * j == n) v = {w * ld(j), C(0)};
      else if (2 * j < n) v = {w * ld(j), w * ld(n - j)};
      else v = {w * ld(n - j), -w * ld(j)};
      v = cswap(v);  // backward DFT via swap
      break;
    }
    default:  // DHT, R2HC
      v.x = ld(j);
      break;
  }
  static_cast<cpx<C> *>(a.out)[i] = v;
}

template <typename S>
__global__ void __launch_bounds__(256) k_r2r_post(KArgs a) {
  using ST = StorageT<S>;
  using C = typename ST::C;
  using Rs = typename ST::Rs;
  const int64_t n = a.n, L = a.n1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * n) return;
  const int64_t b = i / n, k = i % n;
  const cpx<C> *V = static_cast<const cpx<C> *>(a.in) + b * L;
  const int kind = a.flags & 255;
  const bool adj = a.flags & 256;
  C y;
  switch (kind) {
    case 0:  // DCT-I
      y = V[k].x;
      if (adj && k != 0 && k != n - 1) y *= C(2);
      break;
    case 3:
    case 7: {  // DCT-IV / DST-IV
      const cpx<C> r = cmul(static_cast<const cpx<C> *>(a.aux1)[k], V[k]);
      y = kind == 3 ? C(2) * r.x : C(-2) * r.y;
      break;
    }
    case 4:  // DST-I
      y = -V[k + 1].y;
      break;
    case 8:  // DHT
      y = V[k].x - V[k].y;
      break;
    case 9:  // R2HC
      y = 2 * k <= n ? V[k].x : V[n - k].y;
      if (adj && k != 0 && 2 * k != n) y *= C(2);
      break;
    default:  // HC2R: real part of the swapped-back result
      y = V[k].y;
      break;
  }
  static_cast<Rs *>(a.out)[b * a.odist + k * a.ostride] = ST::str(y);
}

}  // namespace plf
