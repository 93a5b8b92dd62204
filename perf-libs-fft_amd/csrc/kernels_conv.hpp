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
  const int n = (int)a.n1;
  const E *in = static_cast<const E *>(a.in) + (valid ? tr * a.idist : 0);
  E *out = static_cast<E *>(a.out) + (valid ? tr * a.odist : 0);
  cpx<C> *lds = dyn_lds<C>() + g * Eng::lds_per(M);
  const cpx<C> *tw = static_cast<const cpx<C> *>(a.tw);
  const cpx<C> *B = static_cast<const cpx<C> *>(a.aux0);
  const cpx<C> *chirp = static_cast<const cpx<C> *>(a.aux1);
  const int64_t is = a.istride, os = a.ostride;
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

}  // namespace plf
