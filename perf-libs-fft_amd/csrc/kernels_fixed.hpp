// kernels_fixed.hpp -- Q0.15 / Q0.7 fixed-point C2C for power-of-two n.
//
// Semantics of the reference's fixed-point codelets (src/providers/jit/
// print_algo.cpp:1142-1239, 2110-2196; src/providers/jit/emit.hpp:150-430):
// every radix-R codelet first shifts its inputs right by log2(R) with
// rounding (srshr), multiplies by the stage twiddles with one rounding per
// component (widening multiply-accumulate, rounding saturating narrow),
// then runs its butterflies with saturating add/sub; the transform is X/n.
// Twiddles and internal constants are sat(trunc(float(x) * 2^bits)).
//
// The op graph fixed here: a Stockham DIT over the chain [2^(e mod 4) if
// nonzero, then 16, 16, ...], stage twiddle W_{LR}^{j m} skipped when
// j m = 0, radix 8 = two radix-4 + W_8 + radix-2, radix 16 = radix-4 over
// stride-4 groups + W_16^{j k} + radix-4. Backward = swap o forward o swap,
// exact in integers. All arithmetic fits int32: a twiddle product sees
// inputs already shifted right by at least one bit.
#pragma once

#include <stdint.h>

#include "fft_device.hpp"
#include "kernel_args.hpp"

namespace plf {

template <int BITS>
struct QPair;
template <>
struct QPair<15> {
  int16_t re, im;
};
template <>
struct QPair<7> {
  int8_t re, im;
};

struct qc {
  int re, im;
};

template <int BITS>
struct QOps {
  static constexpr int HI = (1 << BITS) - 1, LO = -(1 << BITS);
  using P = QPair<BITS>;
  static __device__ __forceinline__ int sat(int v) { return v > HI ? HI : (v < LO ? LO : v); }
  static __device__ __forceinline__ qc add(qc a, qc b) { return {sat(a.re + b.re), sat(a.im + b.im)}; }
  static __device__ __forceinline__ qc sub(qc a, qc b) { return {sat(a.re - b.re), sat(a.im - b.im)}; }
  // x * (-i) = (im, -re); the negation saturates (-LO > HI)
  static __device__ __forceinline__ qc negi(qc a) { return {a.im, sat(-a.re)}; }
  // srshr: (v + 2^(s-1)) >> s, arithmetic (floor)
  template <int S>
  static __device__ __forceinline__ int rshr(int v) {
    if constexpr (S == 0) return v;
    else return (v + (1 << (S - 1))) >> S;
  }
  // complex product, one rounding per component
  static __device__ __forceinline__ qc cmul(qc x, qc w) {
    constexpr int H = 1 << (BITS - 1);
    return {sat((x.re * w.re - x.im * w.im + H) >> BITS),
            sat((x.re * w.im + x.im * w.re + H) >> BITS)};
  }
  static __device__ __forceinline__ qc unpack(P p) { return {(int)p.re, (int)p.im}; }
  static __device__ __forceinline__ P pack(qc v) {
    P p;
    p.re = (decltype(p.re))v.re;
    p.im = (decltype(p.im))v.im;
    return p;
  }

  static __device__ __forceinline__ void dft4(qc *a) {
    const qc t0 = add(a[0], a[2]), t1 = sub(a[0], a[2]);
    const qc t2 = add(a[1], a[3]), t3 = negi(sub(a[1], a[3]));
    a[0] = add(t0, t2);
    a[2] = sub(t0, t2);
    a[1] = add(t1, t3);
    a[3] = sub(t1, t3);
  }

  // radix-R butterfly on shifted inputs; w16[k] = Q(W_16^k)
  template <int R>
  static __device__ __forceinline__ void dft(qc *a, const P *w16) {
    if constexpr (R == 2) {
      const qc x = add(a[0], a[1]), y = sub(a[0], a[1]);
      a[0] = x;
      a[1] = y;
    } else if constexpr (R == 4) {
      dft4(a);
    } else if constexpr (R == 8) {
      qc e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
      dft4(e);
      dft4(o);
      o[1] = cmul(o[1], unpack(w16[2]));
      o[2] = negi(o[2]);
      o[3] = cmul(o[3], unpack(w16[6]));
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        a[k] = add(e[k], o[k]);
        a[k + 4] = sub(e[k], o[k]);
      }
    } else if constexpr (R == 16) {
      qc f[4][4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int m = 0; m < 4; ++m) f[j][m] = a[j + 4 * m];
        dft4(f[j]);
      }
#pragma unroll
      for (int j = 1; j < 4; ++j)
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1) {
          const int e = j * k1;
          f[j][k1] = e == 4 ? negi(f[j][k1]) : cmul(f[j][k1], unpack(w16[e]));
        }
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) {
        qc g[4] = {f[0][k1], f[1][k1], f[2][k1], f[3][k1]};
        dft4(g);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = g[k2];
      }
    }
  }
};

// Runtime-chain Stockham over radices 2, 4, 8, 16 with LDS in the storage
// pair type. Stage 0 reads global memory, the last stage writes it.
template <int BITS, int TPT, int PTS>
struct FixedEngine {
  using O = QOps<BITS>;
  using P = QPair<BITS>;
  static constexpr int PAD = 5;

  template <int R, class Load, class Store>
  static __device__ __forceinline__ void stage(const GenericDesc *__restrict__ d, int s, P *lds,
                                               const P *__restrict__ tw,
                                               const P *__restrict__ w16, int t, bool valid,
                                               Load &load, Store &store) {
    constexpr int KMAX = (PTS + R - 1) / R;
    constexpr int SH = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
    asm volatile("" : "+v"(t));  // keep per-stage addressing out of the stage loop
    const int n = d->n;
    const int NB = n / R;
    const int L = d->L[s];
    const bool first = s == 0;
    const bool last = s == d->ns - 1;
    const P *tws = tw + d->twoff[s];
    qc v[KMAX][R];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        qc e = {0, 0};
        if (b < NB) {
          if (first) {
            if (valid) e = load(b + m * NB);
          } else {
            e = O::unpack(lds[lds_pad<PAD>(b + m * NB)]);
          }
        }
        v[k][m] = {O::template rshr<SH>(e.re), O::template rshr<SH>(e.im)};
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      const int j = b % L;
      if (!first && b < NB && j > 0) {
#pragma unroll
        for (int m = 1; m < R; ++m) v[k][m] = O::cmul(v[k][m], O::unpack(tws[(m - 1) * L + j]));
      }
      O::template dft<R>(v[k], w16);
    }
    if (!first && !last) __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      if (b < NB) {
        if (last) {
          if (valid) {
#pragma unroll
            for (int m = 0; m < R; ++m) store(b + m * NB, v[k][m]);
          }
        } else {
          const int j = b % L;
          const int base = (b / L) * L * R + j;
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad<PAD>(base + m * L)] = O::pack(v[k][m]);
        }
      }
    }
    if (!last) __syncthreads();
  }

  template <class Load, class Store>
  static __device__ __forceinline__ void run(const GenericDesc *__restrict__ d, P *lds,
                                             const P *tw, const P *w16, int t, bool valid,
                                             Load &&load, Store &&store) {
    const int ns = d->ns;
    if (ns == 0) {  // n == 1: identity
      if (t == 0 && valid) store(0, load(0));
      return;
    }
    for (int s = 0; s < ns; ++s) {
      switch (d->radix[s]) {
        case 2: stage<2>(d, s, lds, tw, w16, t, valid, load, store); break;
        case 4: stage<4>(d, s, lds, tw, w16, t, valid, load, store); break;
        case 8: stage<8>(d, s, lds, tw, w16, t, valid, load, store); break;
        case 16: stage<16>(d, s, lds, tw, w16, t, valid, load, store); break;
        default: break;
      }
    }
  }
};

// Batched fixed-point C2C. KArgs: in/out (storage pairs), tw = 16 codelet
// constants Q(W_16^k) followed by the stage twiddles, desc, count, strides
// and distances in complex elements, n, flags bit0 = backward; flags bit1:
// member offsets (complex elements) from the tables aux0 (input) and aux1
// (output) instead of member * distance.
template <int BITS, int TPT, int TPB, int PTS>
__global__ void __launch_bounds__(TPT *TPB) k_fixed(KArgs a) {
  using P = QPair<BITS>;
  extern __shared__ unsigned char plf_dyn_lds[];
  const int tb = threadIdx.x / TPT, t = threadIdx.x % TPT;
  const int64_t member = (int64_t)blockIdx.x * TPB + tb;
  const bool valid = member < a.count;
  P *lds = reinterpret_cast<P *>(plf_dyn_lds) + tb * lds_padded<5>((int)a.n);
  const P *tab = static_cast<const P *>(a.tw);
  const bool table = a.flags & 2;
  const int64_t ioff = !valid ? 0 : table ? static_cast<const int64_t *>(a.aux0)[member]
                                          : member * a.idist;
  const int64_t ooff = !valid ? 0 : table ? static_cast<const int64_t *>(a.aux1)[member]
                                          : member * a.odist;
  const P *in = static_cast<const P *>(a.in) + ioff;
  P *out = static_cast<P *>(a.out) + ooff;
  const bool bwd = a.flags & 1;
  const int64_t is = a.istride, os = a.ostride;
  FixedEngine<BITS, TPT, PTS>::run(
      a.desc, lds, tab + 16, tab, t, valid,
      [&](int e) {
        const P p = in[e * is];
        return bwd ? qc{(int)p.im, (int)p.re} : qc{(int)p.re, (int)p.im};
      },
      [&](int e, qc v) {
        P p;
        p.re = (decltype(p.re))(bwd ? v.im : v.re);
        p.im = (decltype(p.im))(bwd ? v.re : v.im);
        out[e * os] = p;
      });
}

// Real <-> complex staging for fixed-point R2C / C2R through a complex
// transform of length n (workspace: count * n contiguous pairs).
// Widen: ws[b][k] = (x[b*idist + k*istride], 0). KArgs: n, count, strides.
template <int BITS>
__global__ void __launch_bounds__(256) k_fixed_widen(KArgs a) {
  using P = QPair<BITS>;
  using T = decltype(P::re);
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * a.n) return;
  const int64_t b = i / a.n, k = i % a.n;
  const T *in = static_cast<const T *>(a.in);
  P p;
  p.re = in[b * a.idist + k * a.istride];
  p.im = 0;
  static_cast<P *>(a.out)[i] = p;
}

// Bins 0..n/2 of the complex result to the R2C output.
template <int BITS>
__global__ void __launch_bounds__(256) k_fixed_half(KArgs a) {
  using P = QPair<BITS>;
  const int64_t h = a.n / 2 + 1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * h) return;
  const int64_t b = i / h, k = i % h;
  static_cast<P *>(a.out)[b * a.odist + k * a.ostride] =
      static_cast<const P *>(a.in)[b * a.n + k];
}

// Hermitian extension of n/2+1 bins to n: X[n-k] = conj(X[k]) (negation
// saturating), imaginary parts of bins 0 and n/2 (even n) taken as zero.
template <int BITS>
__global__ void __launch_bounds__(256) k_fixed_hermitian(KArgs a) {
  using P = QPair<BITS>;
  using O = QOps<BITS>;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * a.n) return;
  const int64_t n = a.n, b = i / n, k = i % n;
  const P *in = static_cast<const P *>(a.in) + b * a.idist;
  P p;
  if (2 * k <= n) {
    p = in[k * a.istride];
    if (k == 0 || 2 * k == n) p.im = 0;
  } else {
    const P q = in[(n - k) * a.istride];
    p.re = q.re;
    p.im = (decltype(p.im))O::sat(-(int)q.im);
  }
  static_cast<P *>(a.out)[i] = p;
}

// Real parts of the complex result to the C2R output.
template <int BITS>
__global__ void __launch_bounds__(256) k_fixed_real(KArgs a) {
  using P = QPair<BITS>;
  using T = decltype(P::re);
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.count * a.n) return;
  const int64_t b = i / a.n, k = i % a.n;
  static_cast<T *>(a.out)[b * a.odist + k * a.ostride] = static_cast<const P *>(a.in)[i].re;
}

}  // namespace plf
