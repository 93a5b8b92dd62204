// fft_device.hpp -- device building blocks for the gfx950 FFT kernels.
//
// Complex values live in registers as {x = re, y = im}. Every butterfly here
// computes the FORWARD DFT (exponent sign -1). A backward transform swaps the
// real and imaginary parts on load and on store: swap(DFT(swap(x))) equals
// the unnormalized backward DFT, so one set of butterflies serves both
// directions at no arithmetic cost.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plf {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

// ---------------------------------------------------------------------------
// compile-time trigonometry (for butterfly constants)
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ce_sin_series(double a) {  // |a| <= pi/4
  double a2 = a * a, term = a, sum = a;
  for (int i = 1; i < 16; ++i) {
    term *= -a2 / double((2 * i) * (2 * i + 1));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos_series(double a) {  // |a| <= pi/4
  double a2 = a * a, term = 1.0, sum = 1.0;
  for (int i = 1; i < 16; ++i) {
    term *= -a2 / double((2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}
// cos and sin of 2*pi*k/R, reduced to [0, pi/4] with integer arithmetic on
// numerators over the common denominator 8R.
struct ce_sc {
  double c, s;
};
constexpr ce_sc ce_sincos_2pi(long long k, long long R) {
  long long num = ((k % R) + R) % R * 8;  // angle = 2*pi*num/(8R)
  const long long D = 8 * R;
  double ss = 1.0, cs = 1.0;
  if (2 * num > D) {  // (pi, 2pi): reflect
    num = D - num;
    ss = -1.0;
  }
  if (4 * num > D) {  // (pi/2, pi]
    num = D / 2 - num;
    cs = -cs;
  }
  bool swap = false;
  if (8 * num > D) {  // (pi/4, pi/2]
    num = D / 4 - num;
    swap = true;
  }
  const double a = 2.0 * kPi * double(num) / double(D);
  double c = ce_cos_series(a), s = ce_sin_series(a);
  if (swap) {
    double t = c;
    c = s;
    s = t;
  }
  return ce_sc{cs * c, ss * s};
}

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj_mul(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> mul_negi(cpx<T> a) {  // a * (-i)
  return {a.y, -a.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> mul_posi(cpx<T> a) {  // a * (+i)
  return {-a.y, a.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cswap(cpx<T> a) {
  return {a.y, a.x};
}

// a * exp(-2*pi*i*E/R), E and R compile-time; multiples of pi/4 take the
// short forms.
template <int R, int E, typename T>
__device__ __forceinline__ cpx<T> mul_wc(cpx<T> a) {
  constexpr int e = ((E % R) + R) % R;
  if constexpr (e == 0) {
    return a;
  } else if constexpr ((8 * e) % R == 0) {
    constexpr int o = 8 * e / R;
    constexpr T h = T(0.70710678118654752440084436210484904);
    if constexpr (o == 2) return mul_negi(a);
    else if constexpr (o == 4) return cpx<T>{-a.x, -a.y};
    else if constexpr (o == 6) return mul_posi(a);
    else if constexpr (o == 1) return cpx<T>{(a.x + a.y) * h, (a.y - a.x) * h};
    else if constexpr (o == 3) return cpx<T>{(a.y - a.x) * h, -(a.x + a.y) * h};
    else if constexpr (o == 5) return cpx<T>{-(a.x + a.y) * h, (a.x - a.y) * h};
    else return cpx<T>{(a.x - a.y) * h, (a.x + a.y) * h};  // o == 7
  } else {
    constexpr ce_sc w = ce_sincos_2pi(e, R);
    constexpr T c = T(w.c);
    constexpr T s = T(-w.s);
    return cpx<T>{a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

// exp(-2*pi*i*E/R) as a constant
template <int R, int E, typename T>
__device__ __forceinline__ cpx<T> root_c() {
  return mul_wc<R, E>(cpx<T>{T(1), T(0)});
}

// f(std::integral_constant<int, I>{}) for I = 0 .. N-1: unrolled loops whose
// index must be a compile-time constant (twiddle constants, codelet slots)
template <int I, int N, class F>
__device__ __forceinline__ void static_for_impl(F &f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for_impl<I + 1, N>(f);
  }
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl<0, N>(f);
}

// ---------------------------------------------------------------------------
// in-register forward DFT of length R on v[0..R)
// ---------------------------------------------------------------------------
constexpr int first_factor(int R) {
  if (R % 4 == 0 && R != 4) return 4;
  if (R % 2 == 0 && R != 2) return 2;
  for (int f = 3; f * f <= R; f += 2)
    if (R % f == 0) return f;
  return R;
}

template <int R, typename T>
struct Dft;

template <typename T>
struct Dft<1, T> {
  static __device__ __forceinline__ void run(cpx<T> *) {}
};

template <typename T>
struct Dft<2, T> {
  static __device__ __forceinline__ void run(cpx<T> *v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T>
struct Dft<3, T> {
  static __device__ __forceinline__ void run(cpx<T> *v) {
    constexpr T c = T(0.86602540378443864676372317075293618);  // sin(2pi/3)
    cpx<T> t1 = cadd(v[1], v[2]);
    cpx<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    cpx<T> d = csub(v[1], v[2]);
    cpx<T> t3 = {c * d.x, c * d.y};
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, mul_negi(t3));
    v[2] = cadd(t2, mul_posi(t3));
  }
};

template <typename T>
struct Dft<4, T> {
  static __device__ __forceinline__ void run(cpx<T> *v) {
    cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cpx<T> t2 = cadd(v[1], v[3]), t3 = mul_negi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

// Odd prime P > 3: symmetric-pair evaluation.
// y_k = A_k - i B_k, y_{P-k} = A_k + i B_k with
// A_k = v0 + sum_j (v_j + v_{P-j}) cos(2pi jk/P),
// B_k = sum_j (v_j - v_{P-j}) sin(2pi jk/P).
template <int P, typename T>
struct DftOddPrime {
  static __device__ __forceinline__ void run(cpx<T> *v) {
    constexpr int H = (P - 1) / 2;
    cpx<T> s[H], d[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      s[j - 1] = cadd(v[j], v[P - j]);
      d[j - 1] = csub(v[j], v[P - j]);
    }
    cpx<T> y0 = v[0];
#pragma unroll
    for (int j = 0; j < H; ++j) y0 = cadd(y0, s[j]);
    cpx<T> out[P];
    out[0] = y0;
    body<1>(v[0], s, d, out);
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = out[k];
  }
  template <int K>
  static __device__ __forceinline__ void body(cpx<T> v0, const cpx<T> *s,
                                              const cpx<T> *d, cpx<T> *out) {
    if constexpr (K <= (P - 1) / 2) {
      cpx<T> A = v0, B = {T(0), T(0)};
      accum<K, 1>(A, B, s, d);
      out[K] = cadd(A, mul_negi(B));
      out[P - K] = cadd(A, mul_posi(B));
      body<K + 1>(v0, s, d, out);
    }
  }
  template <int K, int J>
  static __device__ __forceinline__ void accum(cpx<T> &A, cpx<T> &B,
                                               const cpx<T> *s, const cpx<T> *d) {
    if constexpr (J <= (P - 1) / 2) {
      constexpr ce_sc w = ce_sincos_2pi((long long)J * K, P);
      constexpr T c = T(w.c), sn = T(w.s);
      A.x += s[J - 1].x * c;
      A.y += s[J - 1].y * c;
      B.x += d[J - 1].x * sn;
      B.y += d[J - 1].y * sn;
      accum<K, J + 1>(A, B, s, d);
    }
  }
};

// Composite R = A * B: n = B*n1 + n2, k = k1 + A*k2.
// Step 1: A-point DFTs over n1 for each n2; step 2: twiddle W_R^{n2*k1};
// step 3: B-point DFTs over n2 for each k1.
template <int R, int A, typename T>
struct DftComposite {
  static constexpr int B = R / A;
  static __device__ __forceinline__ void run(cpx<T> *v) {
    cpx<T> t[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      cpx<T> a[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) a[n1] = v[B * n1 + n2];
      Dft<A, T>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) t[n2 * A + k1] = a[k1];
    }
    twid<0>(t);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      cpx<T> b[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) b[n2] = t[n2 * A + k1];
      Dft<B, T>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = b[k2];
    }
  }
  template <int I>
  static __device__ __forceinline__ void twid(cpx<T> *t) {
    if constexpr (I < R) {
      constexpr int n2 = I / A, k1 = I % A;
      t[I] = mul_wc<R, n2 * k1>(t[I]);
      twid<I + 1>(t);
    }
  }
};

template <int R, typename T>
struct Dft {
  static __device__ __forceinline__ void run(cpx<T> *v) {
    constexpr int f = first_factor(R);
    if constexpr (f == R) {
      DftOddPrime<R, T>::run(v);
    } else {
      DftComposite<R, f, T>::run(v);
    }
  }
};

// ---------------------------------------------------------------------------
// storage <-> compute conversion
// ---------------------------------------------------------------------------
struct half_pair {
  uint16_t x, y;
};
struct bf16_pair {
  uint16_t x, y;
};

__device__ __forceinline__ float h2f(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}
__device__ __forceinline__ uint16_t f2h(float f) {  // round to nearest even
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}
__device__ __forceinline__ float bf2f(uint16_t b) {
  return __uint_as_float(uint32_t(b) << 16);
}
// float -> bf16, round to nearest even: gfx950's v_cvt_pk_bf16_f32
__device__ __forceinline__ uint16_t f2bf(float f) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
// two floats -> packed bf16 pair (lo in bits 0..15), one instruction
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}

// Storage traits: E is one complex element in memory, R the real scalar.
template <typename S>
struct StorageT;

template <>
struct StorageT<float> {
  using E = cpx<float>;
  using Rs = float;
  using C = float;
  typedef float V2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ cpx<float> ld(const E &e) { return e; }
  static __device__ __forceinline__ E st(cpx<float> v) { return v; }
  // streaming (non-temporal) element access for data touched exactly once
  static __device__ __forceinline__ cpx<float> ldnt(const E *p) {
    V2 v = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(p));
    return {v.x, v.y};
  }
  static __device__ __forceinline__ void stnt(E *p, cpx<float> c) {
    V2 v = {c.x, c.y};
    __builtin_nontemporal_store(v, reinterpret_cast<V2 *>(p));
  }
  static __device__ __forceinline__ float ldr(Rs r) { return r; }
  static __device__ __forceinline__ Rs str(float v) { return v; }
};
template <>
struct StorageT<double> {
  using E = cpx<double>;
  using Rs = double;
  using C = double;
  typedef double V2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ cpx<double> ld(const E &e) { return e; }
  static __device__ __forceinline__ E st(cpx<double> v) { return v; }
  static __device__ __forceinline__ cpx<double> ldnt(const E *p) {
    V2 v = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(p));
    return {v.x, v.y};
  }
  static __device__ __forceinline__ void stnt(E *p, cpx<double> c) {
    V2 v = {c.x, c.y};
    __builtin_nontemporal_store(v, reinterpret_cast<V2 *>(p));
  }
  static __device__ __forceinline__ double ldr(Rs r) { return r; }
  static __device__ __forceinline__ Rs str(double v) { return v; }
};
template <>
struct StorageT<half_pair> {
  using E = half_pair;
  using Rs = uint16_t;
  using C = float;
  static __device__ __forceinline__ cpx<float> ldnt(const E *p) {
    const uint32_t u = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
    return {h2f(uint16_t(u & 0xffffu)), h2f(uint16_t(u >> 16))};
  }
  static __device__ __forceinline__ void stnt(E *p, cpx<float> c) {
    const uint32_t u = uint32_t(f2h(c.x)) | (uint32_t(f2h(c.y)) << 16);
    __builtin_nontemporal_store(u, reinterpret_cast<uint32_t *>(p));
  }
  static __device__ __forceinline__ cpx<float> ld(const E &e) {
    return {h2f(e.x), h2f(e.y)};
  }
  static __device__ __forceinline__ E st(cpx<float> v) { return {f2h(v.x), f2h(v.y)}; }
  static __device__ __forceinline__ float ldr(Rs r) { return h2f(r); }
  static __device__ __forceinline__ Rs str(float v) { return f2h(v); }
};
template <>
struct StorageT<bf16_pair> {
  using E = bf16_pair;
  using Rs = uint16_t;
  using C = float;
  static __device__ __forceinline__ cpx<float> ldnt(const E *p) {
    const uint32_t u = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
    return {bf2f(uint16_t(u & 0xffffu)), bf2f(uint16_t(u >> 16))};
  }
  static __device__ __forceinline__ void stnt(E *p, cpx<float> c) {
    __builtin_nontemporal_store(f2bf2(c.x, c.y), reinterpret_cast<uint32_t *>(p));
  }
  static __device__ __forceinline__ cpx<float> ld(const E &e) {
    return {bf2f(e.x), bf2f(e.y)};
  }
  static __device__ __forceinline__ E st(cpx<float> v) {
    return {f2bf(v.x), f2bf(v.y)};
  }
  static __device__ __forceinline__ float ldr(Rs r) { return bf2f(r); }
  static __device__ __forceinline__ Rs str(float v) { return f2bf(v); }
};

// ---------------------------------------------------------------------------
// compile-time radix chain
// ---------------------------------------------------------------------------
template <int... Rs>
struct Chain {
  static constexpr int count = sizeof...(Rs);
  static constexpr int N = (Rs * ... * 1);
  static constexpr int radix(int s) {
    constexpr int a[] = {Rs...};
    return a[s];
  }
  static constexpr int L(int s) {  // product of the radices before stage s
    constexpr int a[] = {Rs...};
    int l = 1;
    for (int i = 0; i < s; ++i) l *= a[i];
    return l;
  }
  // offset (in complex entries) of stage s's twiddle block; stage 0 has none
  static constexpr int tw_offset(int s) {
    constexpr int a[] = {Rs...};
    int off = 0, l = a[0];
    for (int i = 1; i < s; ++i) {
      off += l * (a[i] - 1);
      l *= a[i];
    }
    return off;
  }
  static constexpr int tw_size() { return tw_offset(count); }
};

// LDS index with one pad slot every 2^PADSHIFT complex entries, to spread
// strided butterfly accesses over the banks.
template <int PADSHIFT>
__device__ __forceinline__ int lds_pad(int i) {
  if constexpr (PADSHIFT <= 0) return i;
  else return i + (i >> PADSHIFT);
}
template <int PADSHIFT>
constexpr int lds_padded(int n) {
  return PADSHIFT <= 0 ? n : n + (n >> PADSHIFT) + 1;
}
// lds_pad(base + m * STRIDE) for base >= 0. When STRIDE is a multiple of the
// padding period this is lds_pad(base) plus a constant step per m, which
// lets the compiler address a whole butterfly from one register with
// immediate offsets instead of one address register per element.
template <int PADSHIFT, int STRIDE>
__device__ __forceinline__ int lds_pad_step(int base, int m) {
  if constexpr (PADSHIFT > 0 && STRIDE % (1 << PADSHIFT) == 0)
    return lds_pad<PADSHIFT>(base) + m * (STRIDE + (STRIDE >> PADSHIFT));
  else
    return lds_pad<PADSHIFT>(base + m * STRIDE);
}

}  // namespace plf
