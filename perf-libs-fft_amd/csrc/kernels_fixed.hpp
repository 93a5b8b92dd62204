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

#include <type_traits>

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

// Scalar form: each component widened to int32, every result clamped.
template <int BITS>
struct QOps {
  static constexpr int HI = (1 << BITS) - 1, LO = -(1 << BITS);
  using P = QPair<BITS>;
  using V = qc;
  static __device__ __forceinline__ int sat(int v) { return v > HI ? HI : (v < LO ? LO : v); }
  static __device__ __forceinline__ V zero() { return {0, 0}; }
  static __device__ __forceinline__ V in(P p) { return {(int)p.re, (int)p.im}; }
  static __device__ __forceinline__ P out(V v) {
    P p;
    p.re = (decltype(p.re))v.re;
    p.im = (decltype(p.im))v.im;
    return p;
  }
  static __device__ __forceinline__ V swap(V v) { return {v.im, v.re}; }
  static __device__ __forceinline__ V add(V a, V b) { return {sat(a.re + b.re), sat(a.im + b.im)}; }
  static __device__ __forceinline__ V sub(V a, V b) { return {sat(a.re - b.re), sat(a.im - b.im)}; }
  // x * (-i) = (im, -re); the negation saturates (-LO > HI)
  static __device__ __forceinline__ V negi(V a) { return {a.im, sat(-a.re)}; }
  // srshr: (v + 2^(s-1)) >> s, arithmetic (floor), per component
  template <int S>
  static __device__ __forceinline__ V shr(V v) {
    if constexpr (S == 0) return v;
    else return {(v.re + (1 << (S - 1))) >> S, (v.im + (1 << (S - 1))) >> S};
  }
  // complex product, one rounding per component
  static __device__ __forceinline__ V cmul(V x, P wp) {
    constexpr int H = 1 << (BITS - 1);
    const V w = in(wp);
    return {sat((x.re * w.re - x.im * w.im + H) >> BITS),
            sat((x.re * w.im + x.im * w.re + H) >> BITS)};
  }
};

// Packed Q0.15 form: (re, im) in one 32-bit register as two int16 lanes,
// the same results bit for bit.
//   add/sub   v_pk_add_i16 / v_pk_sub_i16 with clamp (saturating)
//   shr<S>    floor(v / 2^S) + bit S-1 of v == floor((v + 2^(S-1)) / 2^S)
//   cmul      two v_dot2 (exact int32 sums of int16 products) + rounding
//             shift + saturating pack. Every product operand x has been
//             shifted right by at least one bit first (|x| <= 2^14), so the
//             16-bit negation of x.im is exact.
typedef short q15x2 __attribute__((ext_vector_type(2)));

struct QOps15P {
  using P = QPair<15>;
  using V = q15x2;
  static __device__ __forceinline__ V zero() { return V{0, 0}; }
  static __device__ __forceinline__ V in(P p) { return __builtin_bit_cast(V, p); }
  static __device__ __forceinline__ P out(V v) { return __builtin_bit_cast(P, v); }
  static __device__ __forceinline__ V swap(V v) { return __builtin_shufflevector(v, v, 1, 0); }
  static __device__ __forceinline__ V add(V a, V b) { return __builtin_elementwise_add_sat(a, b); }
  static __device__ __forceinline__ V sub(V a, V b) { return __builtin_elementwise_sub_sat(a, b); }
  static __device__ __forceinline__ V negi(V a) {
    const V n = __builtin_elementwise_sub_sat(V{0, 0}, a);
    return __builtin_shufflevector(a, n, 1, 2);  // (a.im, sat(-a.re))
  }
  template <int S>
  static __device__ __forceinline__ V shr(V v) {
    if constexpr (S == 0) {
      return v;
    } else {
      const V q = v >> (V)S;
      const V r = (v >> (V)(S - 1)) & (V)1;
      return q + r;
    }
  }
  static __device__ __forceinline__ V cmul(V x, P wp) {
    const V w = in(wp);
    const V xn = {x.x, (short)(-x.y)};
    const int re = __builtin_amdgcn_sdot2(xn, w, 1 << 14, false) >> 15;
    const int im = __builtin_amdgcn_sdot2(x, swap(w), 1 << 14, false) >> 15;
    const int lo = re > 32767 ? 32767 : (re < -32768 ? -32768 : re);
    const int hi = im > 32767 ? 32767 : (im < -32768 ? -32768 : im);
    return V{(short)lo, (short)hi};
  }
};

template <int BITS>
struct QSel {
  using type = QOps<BITS>;
};
template <>
struct QSel<15> {
  using type = QOps15P;
};

// Radix-2/4/8/16 butterflies on shifted inputs over either form;
// w16[k] = Q(W_16^k).
template <class O>
struct QBfly {
  using V = typename O::V;
  using P = typename O::P;
  static __device__ __forceinline__ void dft4(V *a) {
    const V t0 = O::add(a[0], a[2]), t1 = O::sub(a[0], a[2]);
    const V t2 = O::add(a[1], a[3]), t3 = O::negi(O::sub(a[1], a[3]));
    a[0] = O::add(t0, t2);
    a[2] = O::sub(t0, t2);
    a[1] = O::add(t1, t3);
    a[3] = O::sub(t1, t3);
  }

  template <int R>
  static __device__ __forceinline__ void dft(V *a, const P *w16) {
    if constexpr (R == 2) {
      const V x = O::add(a[0], a[1]), y = O::sub(a[0], a[1]);
      a[0] = x;
      a[1] = y;
    } else if constexpr (R == 4) {
      dft4(a);
    } else if constexpr (R == 8) {
      V e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
      dft4(e);
      dft4(o);
      o[1] = O::cmul(o[1], w16[2]);
      o[2] = O::negi(o[2]);
      o[3] = O::cmul(o[3], w16[6]);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        a[k] = O::add(e[k], o[k]);
        a[k + 4] = O::sub(e[k], o[k]);
      }
    } else if constexpr (R == 16) {
      V f[4][4];
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
          f[j][k1] = e == 4 ? O::negi(f[j][k1]) : O::cmul(f[j][k1], w16[e]);
        }
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) {
        V g[4] = {f[0][k1], f[1][k1], f[2][k1], f[3][k1]};
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
  using O = typename QSel<BITS>::type;
  using V = typename O::V;
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
    const int L = d->L[s];         // a power of two: radices 2, 4, 8, 16 only
    const int lgL = __builtin_ctz(L);
    const bool first = s == 0;
    const bool last = s == d->ns - 1;
    const P *tws = tw + d->twoff[s];
    V v[KMAX][R];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
#pragma unroll
      for (int m = 0; m < R; ++m) {
        V e = O::zero();
        if (b < NB) {
          if (first) {
            if (valid) e = load(b + m * NB);
          } else {
            e = O::in(lds[lds_pad<PAD>(b + m * NB)]);
          }
        }
        v[k][m] = O::template shr<SH>(e);
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      const int j = b & (L - 1);
      if (!first && b < NB && j > 0) {
#pragma unroll
        for (int m = 1; m < R; ++m) v[k][m] = O::cmul(v[k][m], tws[(m - 1) * L + j]);
      }
      QBfly<O>::template dft<R>(v[k], w16);
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
          const int j = b & (L - 1);
          const int base = ((b >> lgL) << lgL) * R + j;
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad<PAD>(base + m * L)] = O::out(v[k][m]);
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
  using O = typename FixedEngine<BITS, TPT, PTS>::O;
  using V = typename O::V;
  // unit strides (the common layout) index without 64-bit multiplies
  auto go = [&](auto unit) {
    constexpr bool U = decltype(unit)::value;
    FixedEngine<BITS, TPT, PTS>::run(
        a.desc, lds, tab + 16, tab, t, valid,
        [&](int e) {
          const V v = O::in(in[U ? (int64_t)e : e * is]);
          return bwd ? O::swap(v) : v;
        },
        [&](int e, V v) { out[U ? (int64_t)e : e * os] = O::out(bwd ? O::swap(v) : v); });
  };
  if (is == 1 && os == 1) go(std::true_type{});
  else go(std::false_type{});
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
