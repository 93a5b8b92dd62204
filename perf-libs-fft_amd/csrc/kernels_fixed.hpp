// kernels_fixed.hpp -- Q0.15 / Q0.7 fixed-point C2C for power-of-two n.
//
// The reference's fixed-point codelets (src/providers/jit/print_algo.cpp:
// 1142-1239 load, 2110-2196 twiddle, 1479-1560 store; src/providers/jit/
// emit.hpp:150-300) computed by a Stockham DIT over the chain
// [2^(e mod 4) if nonzero, then 16, 16, ...] (e = log2 n), one radix-R
// pass per stage:
//   1. each input is shifted right by log2 R with rounding (srshr);
//   2. on a twiddled column (j > 0) input m > 0 is multiplied by the stage
//      twiddle as one widening multiply-accumulate and one rounding narrow,
//      with the premultiplied pairs (Q cos t, Q(-sin t)) and
//      (Q sin t, Q cos t), t = ((dir 2 pi / (L R)) m) j;
//   3. the codelet's op list (fixed_codelets.hpp: the generator's Winograd
//      forms for 2, 4, 8 and split radix for 16) computes the forward DFT_R
//      with saturating adds, per-lane rounding multiplies and one-rounding
//      multiply-accumulates;
//   4. the backward transform stores codelet output k at slot (R - k) mod R
//      (print_algo.cpp:1486-1487) and uses the dir = +1 twiddles.
// The multiply-accumulates keep the 2B+2-bit accumulator of the widening
// instructions (32 bits for Q15, 16 for Q7, wrapping) and narrow with
// sat(floor((acc + 2^(B-1)) / 2^B)).
#pragma once

#include <stdint.h>

#include <type_traits>

#include "fft_device.hpp"
#include "fixed_codelets.hpp"
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

// One stage twiddle as its two premultiplied pairs: re = x . a, im = x . b
// with a = (Q cos t, Q(-sin t)), b = (Q sin t, Q cos t).
template <int BITS>
struct QTw {
  QPair<BITS> a, b;
};

struct qc {
  int re, im;
};

// Scalar form: each component widened to int32.
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
  static __device__ __forceinline__ V add(V a, V b) { return {sat(a.re + b.re), sat(a.im + b.im)}; }
  static __device__ __forceinline__ V sub(V a, V b) { return {sat(a.re - b.re), sat(a.im - b.im)}; }
  // srshr: floor((v + 2^(S-1)) / 2^S) per component
  template <int S>
  static __device__ __forceinline__ V shr(V v) {
    if constexpr (S == 0) return v;
    else return {(v.re + (1 << (S - 1))) >> S, (v.im + (1 << (S - 1))) >> S};
  }
  // the accumulator of the widening instructions: 2 BITS + 2 bits, wrapping
  static __device__ __forceinline__ int wrap(unsigned acc) {
    if constexpr (BITS == 7) return (int)(int16_t)(uint16_t)acc;
    else return (int)acc;
  }
  // rounding saturating narrow of a wrapped accumulator; floor of
  // (w + 2^(B-1)) / 2^B is (w >> B) plus bit B-1 of w, with no overflow
  static __device__ __forceinline__ int narrow(unsigned acc) {
    const int w = wrap(acc);
    return sat((w >> BITS) + ((w >> (BITS - 1)) & 1));
  }
  template <bool REV>
  static __device__ __forceinline__ V mul(V a, int q0, int q1) {
    const int l0 = REV ? a.im : a.re, l1 = REV ? a.re : a.im;
    return {narrow((unsigned)(l0 * q0)), narrow((unsigned)(l1 * q1))};
  }
  template <int SG, bool RA, bool RB>
  static __device__ __forceinline__ V mac(V a, int qa0, int qa1, V b, int qb0, int qb1) {
    const int a0 = RA ? a.im : a.re, a1 = RA ? a.re : a.im;
    const int b0 = RB ? b.im : b.re, b1 = RB ? b.re : b.im;
    const unsigned pa0 = (unsigned)(a0 * qa0), pa1 = (unsigned)(a1 * qa1);
    const unsigned pb0 = (unsigned)(b0 * qb0), pb1 = (unsigned)(b1 * qb1);
    return SG > 0 ? V{narrow(pa0 + pb0), narrow(pa1 + pb1)}
                  : V{narrow(pa0 - pb0), narrow(pa1 - pb1)};
  }
  static __device__ __forceinline__ V twiddle(V x, const QTw<BITS> &w) {
    return {narrow((unsigned)(x.re * (int)w.a.re + x.im * (int)w.a.im)),
            narrow((unsigned)(x.re * (int)w.b.re + x.im * (int)w.b.im))};
  }
};

// Packed Q0.15 form: (re, im) in one 32-bit register as two int16 lanes,
// the same results bit for bit.
//   add/sub   saturating packed 16-bit add / subtract
//   shr<S>    floor(v / 2^S) + bit S-1 of v == floor((v + 2^(S-1)) / 2^S)
//   products  v_dot2 (int16 x int16 products summed exactly into an int32
//             accumulator, modulo 2^32 like the 32-bit accumulator of the
//             widening instructions) with the rounding term 2^14 as its
//             addend, then >> 15 and saturation. Adding 2^14 before the
//             shift cannot change a result: a sum of two int16 x int16
//             products never lies in [2^31 - 2^14, 2^31) (the only product
//             above 2^30 - 2^15 is (-2^15)^2 = 2^30), so the wrapped sum plus
//             2^14 stays below 2^31.
typedef short q15x2 __attribute__((ext_vector_type(2)));

struct QOps15P {
  using P = QPair<15>;
  using V = q15x2;
  static __device__ __forceinline__ V zero() { return V{0, 0}; }
  static __device__ __forceinline__ V in(P p) { return __builtin_bit_cast(V, p); }
  static __device__ __forceinline__ P out(V v) { return __builtin_bit_cast(P, v); }
  static __device__ __forceinline__ V add(V a, V b) { return __builtin_elementwise_add_sat(a, b); }
  static __device__ __forceinline__ V sub(V a, V b) { return __builtin_elementwise_sub_sat(a, b); }
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
  static __device__ __forceinline__ int sat(int v) {
    return v > 32767 ? 32767 : (v < -32768 ? -32768 : v);
  }
  static __device__ __forceinline__ V pack(int re_acc, int im_acc) {
    return V{(short)sat(re_acc >> 15), (short)sat(im_acc >> 15)};
  }
  static __device__ __forceinline__ int dot(V x, V w, int c) {
    return __builtin_amdgcn_sdot2(x, w, c, false);
  }
  // constant pair selecting lane 0 (or lane 1 when REV) of the operand
  template <bool REV>
  static __device__ __forceinline__ V lane0(int q) {
    return REV ? V{0, (short)q} : V{(short)q, 0};
  }
  template <bool REV>
  static __device__ __forceinline__ V lane1(int q) {
    return REV ? V{(short)q, 0} : V{0, (short)q};
  }
  template <bool REV>
  static __device__ __forceinline__ V mul(V a, int q0, int q1) {
    return pack(dot(a, lane0<REV>(q0), 1 << 14), dot(a, lane1<REV>(q1), 1 << 14));
  }
  template <int SG, bool RA, bool RB>
  static __device__ __forceinline__ V mac(V a, int qa0, int qa1, V b, int qb0, int qb1) {
    const int pb0 = dot(b, lane0<RB>(qb0), 0), pb1 = dot(b, lane1<RB>(qb1), 0);
    const int c0 = SG > 0 ? (1 << 14) + pb0 : (1 << 14) - pb0;
    const int c1 = SG > 0 ? (1 << 14) + pb1 : (1 << 14) - pb1;
    return pack(dot(a, lane0<RA>(qa0), c0), dot(a, lane1<RA>(qa1), c1));
  }
  static __device__ __forceinline__ V twiddle(V x, const QTw<15> &w) {
    return pack(dot(x, in(w.a), 1 << 14), dot(x, in(w.b), 1 << 14));
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

template <class O, int BITS, int R>
__device__ __forceinline__ void fx_dft(typename O::V *a) {
  if constexpr (R == 2) fx_dft2<O, BITS>(a);
  else if constexpr (R == 4) fx_dft4<O, BITS>(a);
  else if constexpr (R == 8) fx_dft8<O, BITS>(a);
  else fx_dft16<O, BITS>(a);
}

// Runtime-chain Stockham over radices 2, 4, 8, 16 with LDS in the storage
// pair type. Stage 0 reads global memory, the last stage writes it.
template <int BITS, int TPT, int PTS, bool BWD>
struct FixedEngine {
  using O = typename QSel<BITS>::type;
  using V = typename O::V;
  using P = QPair<BITS>;
  using TW = QTw<BITS>;
  static constexpr int PAD = 5;

  // A stage's shape from the runtime descriptor ...
  struct RtStage {
    const GenericDesc *d;
    int s;
    __device__ __forceinline__ int n() const { return d->n; }
    __device__ __forceinline__ int L() const { return d->L[s]; }
    __device__ __forceinline__ int twoff() const { return d->twoff[s]; }
    __device__ __forceinline__ bool first() const { return s == 0; }
    __device__ __forceinline__ bool last() const { return s == d->ns - 1; }
  };
  // ... or from a compile-time chain (every index below folds to a constant)
  template <int N_, int L_, int TWOFF_, bool FIRST_, bool LAST_>
  struct CtStage {
    __device__ __forceinline__ constexpr int n() const { return N_; }
    __device__ __forceinline__ constexpr int L() const { return L_; }
    __device__ __forceinline__ constexpr int twoff() const { return TWOFF_; }
    __device__ __forceinline__ constexpr bool first() const { return FIRST_; }
    __device__ __forceinline__ constexpr bool last() const { return LAST_; }
  };

  template <int R, class SP, class Load, class Store>
  static __device__ __forceinline__ void stage(const SP &sp, P *lds,
                                               const TW *__restrict__ tw, int t, bool valid,
                                               Load &load, Store &store) {
    constexpr int KMAX = (PTS + R - 1) / R;
    constexpr int SH = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
    asm volatile("" : "+v"(t));  // keep per-stage addressing out of the stage loop
    const int n = sp.n();
    const int NB = n / R;
    const int L = sp.L();          // a power of two: radices 2, 4, 8, 16 only
    const int lgL = __builtin_ctz(L);
    const bool first = sp.first();
    const bool last = sp.last();
    const TW *tws = tw + sp.twoff();
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
        for (int m = 1; m < R; ++m) v[k][m] = O::twiddle(v[k][m], tws[(m - 1) * L + j]);
      }
      fx_dft<O, BITS, R>(v[k]);
    }
    if (!first && !last) __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int b = t + k * TPT;
      if (b < NB) {
        // backward: slot m takes codelet output (R - m) mod R
        V o[R];
#pragma unroll
        for (int m = 0; m < R; ++m) o[m] = v[k][BWD ? (R - m) & (R - 1) : m];
        if (last) {
          if (valid) {
#pragma unroll
            for (int m = 0; m < R; ++m) store(b + m * NB, o[m]);
          }
        } else {
          const int j = b & (L - 1);
          const int base = ((b >> lgL) << lgL) * R + j;
#pragma unroll
          for (int m = 0; m < R; ++m) lds[lds_pad<PAD>(base + m * L)] = O::out(o[m]);
        }
      }
    }
    if (!last) __syncthreads();
  }

  template <class Load, class Store>
  static __device__ __forceinline__ void run(const GenericDesc *__restrict__ d, P *lds,
                                             const TW *tw, int t, bool valid,
                                             Load &&load, Store &&store) {
    const int ns = d->ns;
    if (ns == 0) {  // n == 1: identity
      if (t == 0 && valid) store(0, load(0));
      return;
    }
    for (int s = 0; s < ns; ++s) {
      const RtStage sp{d, s};
      switch (d->radix[s]) {
        case 2: stage<2>(sp, lds, tw, t, valid, load, store); break;
        case 4: stage<4>(sp, lds, tw, t, valid, load, store); break;
        case 8: stage<8>(sp, lds, tw, t, valid, load, store); break;
        case 16: stage<16>(sp, lds, tw, t, valid, load, store); break;
        default: break;
      }
    }
  }

  // The same stages over a compile-time chain Rs (the product N, at least
  // two stages); the twiddle table layout is the runtime one's.
  template <int SI, int L, int TWOFF, int... Rs, class Load, class Store>
  static __device__ __forceinline__ void run_static_from(P *lds, const TW *tw, int t, bool valid,
                                                         Load &load, Store &store) {
    constexpr int NS = sizeof...(Rs);
    constexpr int RS[NS] = {Rs...};
    constexpr int N = (Rs * ...);
    constexpr int R = RS[SI];
    const CtStage<N, L, TWOFF, SI == 0, SI == NS - 1> sp{};
    stage<R>(sp, lds, tw, t, valid, load, store);
    if constexpr (SI + 1 < NS)
      run_static_from<SI + 1, L * R, (SI == 0 ? 0 : TWOFF + L * (R - 1)), Rs...>(lds, tw, t, valid,
                                                                                   load, store);
  }
  template <int... Rs, class Load, class Store>
  static __device__ __forceinline__ void run_static(P *lds, const TW *tw, int t, bool valid,
                                                    Load &&load, Store &&store) {
    run_static_from<0, 1, 0, Rs...>(lds, tw, t, valid, load, store);
  }
};

// Batched fixed-point C2C (BWD: the backward transform). KArgs: in/out
// (storage pairs), tw = the stage twiddles (QTw, for the plan's direction),
// desc, count, strides and distances in complex elements, n; flags bit1:
// member offsets (complex elements) from the tables aux0 (input) and aux1
// (output) instead of member * distance; flags bit2 (with bit1, unit
// strides): a gathered launch of several lengths, aux2 = one FxBlock per
// block (desc and tw are then the arrays the records index; n = the longest
// length, which sizes each member's LDS).
template <int BITS, int TPT, int TPB, int PTS, bool BWD>
__global__ void __launch_bounds__(TPT *TPB) k_fixed(KArgs a) {
  using P = QPair<BITS>;
  extern __shared__ unsigned char plf_dyn_lds[];
  const int tb = threadIdx.x / TPT, t = threadIdx.x % TPT;
  const FxBlock *rec = (a.flags & 4) ? static_cast<const FxBlock *>(a.aux2) + blockIdx.x : nullptr;
  const int64_t member = rec ? rec->member + tb : (int64_t)blockIdx.x * TPB + tb;
  const bool valid = rec ? tb < rec->count : member < a.count;
  const GenericDesc *desc = rec ? a.desc + rec->desc : a.desc;
  P *lds = reinterpret_cast<P *>(plf_dyn_lds) + tb * lds_padded<5>((int)a.n);
  const QTw<BITS> *tab = static_cast<const QTw<BITS> *>(a.tw) + (rec ? rec->tw_off : 0);
  const bool table = a.flags & 2;
  const int64_t ioff = !valid ? 0 : table ? static_cast<const int64_t *>(a.aux0)[member]
                                          : member * a.idist;
  const int64_t ooff = !valid ? 0 : table ? static_cast<const int64_t *>(a.aux1)[member]
                                          : member * a.odist;
  const P *in = static_cast<const P *>(a.in) + ioff;
  P *out = static_cast<P *>(a.out) + ooff;
  const int64_t is = a.istride, os = a.ostride;
  using E = FixedEngine<BITS, TPT, PTS, BWD>;
  using O = typename E::O;
  using V = typename O::V;
  // unit strides (the common layout) index without 64-bit multiplies
  auto go = [&](auto unit) {
    constexpr bool U = decltype(unit)::value;
    E::run(
        desc, lds, tab, t, valid,
        [&](int e) { return O::in(in[U ? (int64_t)e : e * is]); },
        [&](int e, V v) { out[U ? (int64_t)e : e * os] = O::out(v); });
  };
  if (is == 1 && os == 1) go(std::true_type{});
  else go(std::false_type{});
}

// k_fixed for one length with a compile-time chain Rs (unit strides,
// members at distance idist / odist, no offset tables). KArgs as k_fixed.
template <int BITS, int TPT, int TPB, int PTS, bool BWD, int... Rs>
__global__ void __launch_bounds__(TPT *TPB) k_fixed_static(KArgs a) {
  using P = QPair<BITS>;
  extern __shared__ unsigned char plf_dyn_lds[];
  constexpr int N = (Rs * ...);
  const int tb = threadIdx.x / TPT, t = threadIdx.x % TPT;
  const int64_t member = (int64_t)blockIdx.x * TPB + tb;
  const bool valid = member < a.count;
  P *lds = reinterpret_cast<P *>(plf_dyn_lds) + tb * lds_padded<5>(N);
  const QTw<BITS> *tab = static_cast<const QTw<BITS> *>(a.tw);
  const P *in = static_cast<const P *>(a.in) + (valid ? member * a.idist : 0);
  P *out = static_cast<P *>(a.out) + (valid ? member * a.odist : 0);
  using E = FixedEngine<BITS, TPT, PTS, BWD>;
  using O = typename E::O;
  using V = typename O::V;
  E::template run_static<Rs...>(
      lds, tab, t, valid, [&](int e) { return O::in(in[e]); },
      [&](int e, V v) { out[e] = O::out(v); });
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
