/*
 * oracle.c -- CPU restatement of the reference's 1D transform path.
 *
 * TEST INFRASTRUCTURE ONLY: linked by nothing in the product; loaded by
 * tests/, by the smoke check and by bench.py's cpu_baseline leg, always as
 * the checker or the timed CPU baseline. See oracle.h.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_CODELET 20

/* ------------------------------------------------------------------------ */
/* seeded inputs                                                             */
/* ------------------------------------------------------------------------ */

/* SplitMix64 (Steele, Lea, Flood 2014): the stream is state += golden gamma,
 * then two xor-shift-multiply mixing rounds. */
uint64_t orc_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* uniform[-1,1): 2*(u53*2^-53) - 1, exact in double. */
void orc_fill_uniform_f64(uint64_t seed, int64_t count, double *out) {
  uint64_t s = seed;
  for (int64_t i = 0; i < count; ++i) {
    uint64_t z = orc_splitmix64(&s);
    out[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
  }
}

/* uniform[-1,1): 2*(u24*2^-24) - 1, exact in float. */
void orc_fill_uniform_f32(uint64_t seed, int64_t count, float *out) {
  uint64_t s = seed;
  for (int64_t i = 0; i < count; ++i) {
    uint64_t z = orc_splitmix64(&s);
    out[i] = 2.0f * ((float)(z >> 40) * 0x1.0p-24f) - 1.0f;
  }
}

/* N(0,1) by Box-Muller in FP64; two draws per pair of outputs. */
void orc_fill_normal_f64(uint64_t seed, int64_t count, double *out) {
  uint64_t s = seed;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < count; i += 2) {
    uint64_t z1 = orc_splitmix64(&s);
    uint64_t z2 = orc_splitmix64(&s);
    double u1 = ((double)(z1 >> 11) + 1.0) * 0x1.0p-53; /* (0, 1] */
    double u2 = (double)(z2 >> 11) * 0x1.0p-53;         /* [0, 1) */
    double r = sqrt(-2.0 * log(u1));
    out[i] = r * cos(two_pi * u2);
    if (i + 1 < count) out[i + 1] = r * sin(two_pi * u2);
  }
}

/* ------------------------------------------------------------------------ */
/* integer helpers                                                           */
/* ------------------------------------------------------------------------ */

/* src/factorize.hpp:16-37: 2, 3, then 6k +/- 1 trial division. */
int64_t orc_smallest_prime_factor(int64_t n) {
  if (n % 2 == 0) return 2;
  if (n % 3 == 0) return 3;
  for (int64_t d = 5; d <= n / d; d += 6) {
    if (n % d == 0) return d;
    if (n % (d + 2) == 0) return d + 2;
  }
  return n;
}

/* src/factorize.hpp:74-87: smallest length >= n that the factors divide out. */
int64_t orc_next_length_factorable_by(int64_t n, const int *factors, int nf) {
  for (int64_t len = n;; ++len) {
    int64_t c = len;
    for (int i = 0; i < nf; ++i)
      while (c % factors[i] == 0) c /= factors[i];
    if (c == 1) return len;
  }
}

/* Bluestein padded length: smallest >= 2n-1 factorable by the AOT codelet
 * sizes 2..20 (src/planner/strategies/bluestein.hpp:42-47). */
int64_t orc_bluestein_length(int64_t n) {
  static const int f[] = {2,  3,  4,  5,  6,  7,  8,  9,  10, 11,
                          12, 13, 14, 15, 16, 17, 18, 19, 20};
  return orc_next_length_factorable_by(2 * n - 1, f, 19);
}

/* src/rader_generator.cpp:13-33: smallest g in 2..n-1 whose powers visit
 * every nonzero residue. */
int64_t orc_find_group_generator(int64_t n) {
  unsigned char *seen = (unsigned char *)malloc((size_t)n);
  for (int64_t cand = 2; cand < n; ++cand) {
    memset(seen, 0, (size_t)n);
    int ok = 1;
    int64_t x = 1;
    for (int64_t i = 1; i < n; ++i) {
      if (seen[x]) {
        ok = 0;
        break;
      }
      seen[x] = 1;
      x = x * cand % n;
    }
    if (ok) {
      free(seen);
      return cand;
    }
  }
  free(seen);
  return 0;
}

/* src/planner/strategies/rader.hpp:51-62: perm[k] = g^k mod n,
 * perm_rev_idx[g^k mod n] = (m-k) % m. perm_rev_idx[0] is set to 0. */
void orc_rader_permutations(int64_t n, int64_t *perm, int64_t *perm_rev_idx) {
  int64_t g = orc_find_group_generator(n);
  int64_t m = n - 1;
  int64_t x = 1;
  perm_rev_idx[0] = 0;
  for (int64_t k = 0; k < m; ++k) {
    perm[k] = x;
    perm_rev_idx[x] = (m - k) % m;
    x = x * g % n;
  }
}

/* src/planner/strategies/bluestein.hpp:66-79: k^2 mod 2n in integers. */
void orc_bluestein_phase(int64_t n, int64_t *phase) {
  for (int64_t k = 0; k < n; ++k)
    phase[k] = (int64_t)(((unsigned __int128)k * (unsigned __int128)k) %
                         (unsigned __int128)(2 * n));
}

static int orc_is_smooth(int64_t n, int64_t bound) {
  for (int64_t f = 2; f <= bound && n > 1; ++f)
    while (n % f == 0) n /= f;
  return n == 1;
}

/* First radix of the reference's descending CT list (40..2, skipping the
 * lengths absent from it) that is an AOT codelet size (<= 20) and a proper
 * divisor of n (src/planner/planner.cpp:104-137). */
static int64_t orc_ct_radix(int64_t n) {
  static const int list[] = {20, 19, 18, 17, 16, 15, 14, 13, 12, 11,
                             10, 9,  8,  7,  6,  5,  4,  3,  2};
  for (int i = 0; i < 19; ++i)
    if (list[i] < n && n % list[i] == 0) return list[i];
  return 0;
}

/* sin/cos of 2*pi*num/den with num reduced modulo den exactly, evaluated in
 * long double and rounded once to double. */
static void orc_sincos_2pi_frac(int64_t num, int64_t den, double *s,
                                double *c) {
  int64_t r = num % den;
  if (r < 0) r += den;
  if (2 * r > den) r -= den; /* angle in (-pi, pi] */
  long double a = 6.283185307179586476925286766559005768L * (long double)r /
                  (long double)den;
  *s = (double)sinl(a);
  *c = (double)cosl(a);
}

/* ------------------------------------------------------------------------ */
/* planned FFT, FP64 and FP32                                                */
/* ------------------------------------------------------------------------ */

#define REAL double
#define SFX _f64
#include "oracle_fft_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX _f32
#include "oracle_fft_impl.h"
#undef REAL
#undef SFX

/* Direct O(n^2) DFT in long double with exact angle reduction: the
 * definition in docs/plfft.md:98-109 (include/plfft.h:91-97). */
void orc_dft(int64_t n, int dir, const double *x, double *y) {
  long double *cs = (long double *)malloc(sizeof(long double) * 2 * (size_t)n);
  for (int64_t k = 0; k < n; ++k) {
    int64_t r = k;
    if (2 * r > n) r -= n;
    long double a = 6.283185307179586476925286766559005768L * (long double)r /
                    (long double)n;
    cs[2 * k] = cosl(a);
    cs[2 * k + 1] = dir * sinl(a);
  }
  for (int64_t k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    int64_t idx = 0;
    for (int64_t j = 0; j < n; ++j) {
      long double xr = x[2 * j], xi = x[2 * j + 1];
      sr += xr * cs[2 * idx] - xi * cs[2 * idx + 1];
      si += xr * cs[2 * idx + 1] + xi * cs[2 * idx];
      idx += k;
      if (idx >= n) idx -= n;
    }
    y[2 * k] = (double)sr;
    y[2 * k + 1] = (double)si;
  }
  free(cs);
}

void orc_fft(int64_t n, int dir, const double *x, double *y) {
  onode_f64 *p = plan_build_f64(n, dir);
  ocpx_f64 *scr = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)(p->scratch + 1));
  ocpx_f64 *tmp = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)n);
  plan_exec_f64(p, (const ocpx_f64 *)x, 1, tmp, scr);
  memcpy(y, tmp, sizeof(ocpx_f64) * (size_t)n);
  free(tmp);
  free(scr);
  plan_free_f64(p);
}

void orc_fft_f32(int64_t n, int dir, const float *x, float *y) {
  onode_f32 *p = plan_build_f32(n, dir);
  ocpx_f32 *scr = (ocpx_f32 *)malloc(sizeof(ocpx_f32) * (size_t)(p->scratch + 1));
  ocpx_f32 *tmp = (ocpx_f32 *)malloc(sizeof(ocpx_f32) * (size_t)n);
  plan_exec_f32(p, (const ocpx_f32 *)x, 1, tmp, scr);
  memcpy(y, tmp, sizeof(ocpx_f32) * (size_t)n);
  free(tmp);
  free(scr);
  plan_free_f32(p);
}

static int describe_rec(const onode_f64 *p, char *buf, int len, int pos) {
  if (!p) return pos;
  switch (p->kind) {
  case 0:
    pos += snprintf(buf + pos, pos < len ? (size_t)(len - pos) : 0, "(direct %lld)", (long long)p->n);
    break;
  case 1:
    pos += snprintf(buf + pos, pos < len ? (size_t)(len - pos) : 0, "(ct %lld r=%lld ", (long long)p->n, (long long)p->r);
    pos = describe_rec(p->sub, buf, len, pos);
    pos += snprintf(buf + pos, pos < len ? (size_t)(len - pos) : 0, ")");
    break;
  case 2:
  case 3:
    pos += snprintf(buf + pos, pos < len ? (size_t)(len - pos) : 0, "(%s %lld m=%lld ",
                    p->kind == 2 ? "rader" : "bluestein", (long long)p->n, (long long)p->m);
    pos = describe_rec(p->sub, buf, len, pos);
    pos += snprintf(buf + pos, pos < len ? (size_t)(len - pos) : 0, ")");
    break;
  }
  return pos;
}

int orc_fft_describe(int64_t n, char *buf, int len) {
  onode_f64 *p = plan_build_f64(n, -1);
  if (len > 0) buf[0] = 0;
  int pos = describe_rec(p, buf, len, 0);
  plan_free_f64(p);
  return pos;
}

/* Batch addressing of docs/plfft.md:136-170: member b, element k at
 * b*dist + k*stride (complex elements). */
void orc_c2c_batched(int64_t n, int dir, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y) {
  onode_f64 *p = plan_build_f64(n, dir);
  ocpx_f64 *scr = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)(p->scratch + 1));
  ocpx_f64 *tmp = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)n);
  const ocpx_f64 *xc = (const ocpx_f64 *)x;
  ocpx_f64 *yc = (ocpx_f64 *)y;
  for (int64_t b = 0; b < howmany; ++b) {
    plan_exec_f64(p, xc + b * idist, istride, tmp, scr);
    for (int64_t k = 0; k < n; ++k) yc[b * odist + k * ostride] = tmp[k];
  }
  free(tmp);
  free(scr);
  plan_free_f64(p);
}

/* FP32 port of the same planned transform, parallel over the batch (the
 * reference's OpenMP batch split, src/plfft_parallel.hpp:63-128). */
void orc_c2c_batched_f32(int64_t n, int dir, int64_t howmany, int64_t istride,
                         int64_t idist, int64_t ostride, int64_t odist,
                         const float *x, float *y, int threads) {
  onode_f32 *p = plan_build_f32(n, dir);
  const ocpx_f32 *xc = (const ocpx_f32 *)x;
  ocpx_f32 *yc = (ocpx_f32 *)y;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
  {
    ocpx_f32 *scr = (ocpx_f32 *)malloc(sizeof(ocpx_f32) * (size_t)(p->scratch + 1));
    ocpx_f32 *tmp = (ocpx_f32 *)malloc(sizeof(ocpx_f32) * (size_t)n);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t b = 0; b < howmany; ++b) {
      plan_exec_f32(p, xc + b * idist, istride, tmp, scr);
      for (int64_t k = 0; k < n; ++k) yc[b * odist + k * ostride] = tmp[k];
    }
    free(tmp);
    free(scr);
  }
  (void)threads;
  plan_free_f32(p);
}

/* R2C: the forward DFT of real input, bins 0..n/2 (docs/plfft.md:111-134).
 * Real side strides count reals; complex side strides count complex. */
void orc_r2c_batched(int64_t n, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y) {
  onode_f64 *p = plan_build_f64(n, -1);
  ocpx_f64 *scr = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)(p->scratch + 1));
  ocpx_f64 *in = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)n);
  ocpx_f64 *tmp = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)n);
  ocpx_f64 *yc = (ocpx_f64 *)y;
  for (int64_t b = 0; b < howmany; ++b) {
    for (int64_t j = 0; j < n; ++j) {
      in[j].re = x[b * idist + j * istride];
      in[j].im = 0;
    }
    plan_exec_f64(p, in, 1, tmp, scr);
    for (int64_t k = 0; k <= n / 2; ++k) yc[b * odist + k * ostride] = tmp[k];
  }
  free(tmp);
  free(in);
  free(scr);
  plan_free_f64(p);
}

/* C2R: backward DFT of the Hermitian extension X[n-k] = conj(X[k]); the
 * imaginary parts of X[0] (and X[n/2], n even) do not contribute. */
void orc_c2r_batched(int64_t n, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y) {
  onode_f64 *p = plan_build_f64(n, 1);
  ocpx_f64 *scr = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)(p->scratch + 1));
  ocpx_f64 *in = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)n);
  ocpx_f64 *tmp = (ocpx_f64 *)malloc(sizeof(ocpx_f64) * (size_t)n);
  const ocpx_f64 *xc = (const ocpx_f64 *)x;
  for (int64_t b = 0; b < howmany; ++b) {
    for (int64_t k = 0; k <= n / 2; ++k) in[k] = xc[b * idist + k * istride];
    in[0].im = 0;
    if (n % 2 == 0) in[n / 2].im = 0;
    for (int64_t k = n / 2 + 1; k < n; ++k) {
      in[k].re = in[n - k].re;
      in[k].im = -in[n - k].im;
    }
    plan_exec_f64(p, in, 1, tmp, scr);
    for (int64_t j = 0; j < n; ++j) y[b * odist + j * ostride] = tmp[j].re;
  }
  free(tmp);
  free(in);
  free(scr);
  plan_free_f64(p);
}

/* ------------------------------------------------------------------------ */
/* real-to-real                                                              */
/* ------------------------------------------------------------------------ */

/* cos / sin of pi * num / den, num reduced mod 2*den exactly. */
static long double cospi_frac(int64_t num, int64_t den) {
  int64_t r = num % (2 * den);
  if (r < 0) r += 2 * den;
  return cosl(3.141592653589793238462643383279502884L * (long double)r / (long double)den);
}
static long double sinpi_frac(int64_t num, int64_t den) {
  int64_t r = num % (2 * den);
  if (r < 0) r += 2 * den;
  return sinl(3.141592653589793238462643383279502884L * (long double)r / (long double)den);
}

/* FFTW's unnormalized R2R definitions (include/plfft.h:105-134,
 * docs/plfft.md:186-212), evaluated directly. Returns 0 on bad input. */
int orc_r2r_direct(int kind, int64_t n, const double *x, double *y) {
  if (n < 1 || kind < 0 || kind > 10) return 0;
  if (kind == 0 && n < 2) return 0;
  for (int64_t k = 0; k < n; ++k) {
    long double s = 0;
    switch (kind) {
    case 0: /* DCT-I, REDFT00 */
      s = x[0] + ((k & 1) ? -1.0L : 1.0L) * x[n - 1];
      for (int64_t j = 1; j < n - 1; ++j) s += 2.0L * x[j] * cospi_frac(j * k, n - 1);
      break;
    case 1: /* DCT-II, REDFT10 */
      for (int64_t j = 0; j < n; ++j) s += 2.0L * x[j] * cospi_frac((2 * j + 1) * k, 2 * n);
      break;
    case 2: /* DCT-III, REDFT01 */
      s = x[0];
      for (int64_t j = 1; j < n; ++j) s += 2.0L * x[j] * cospi_frac(j * (2 * k + 1), 2 * n);
      break;
    case 3: /* DCT-IV, REDFT11 */
      for (int64_t j = 0; j < n; ++j)
        s += 2.0L * x[j] * cospi_frac((2 * j + 1) * (2 * k + 1), 4 * n);
      break;
    case 4: /* DST-I, RODFT00 */
      for (int64_t j = 0; j < n; ++j) s += 2.0L * x[j] * sinpi_frac((j + 1) * (k + 1), n + 1);
      break;
    case 5: /* DST-II, RODFT10 */
      for (int64_t j = 0; j < n; ++j) s += 2.0L * x[j] * sinpi_frac((2 * j + 1) * (k + 1), 2 * n);
      break;
    case 6: /* DST-III, RODFT01 */
      s = ((k & 1) ? -1.0L : 1.0L) * x[n - 1];
      for (int64_t j = 0; j < n - 1; ++j)
        s += 2.0L * x[j] * sinpi_frac((j + 1) * (2 * k + 1), 2 * n);
      break;
    case 7: /* DST-IV, RODFT11 */
      for (int64_t j = 0; j < n; ++j)
        s += 2.0L * x[j] * sinpi_frac((2 * j + 1) * (2 * k + 1), 4 * n);
      break;
    case 8: /* DHT: cas(2 pi j k / n) */
      for (int64_t j = 0; j < n; ++j)
        s += x[j] * (cospi_frac(2 * j * k, n) + sinpi_frac(2 * j * k, n));
      break;
    case 9: /* R2HC */
      if (2 * k <= n) {
        for (int64_t j = 0; j < n; ++j) s += x[j] * cospi_frac(2 * j * k, n);
      } else {
        int64_t kk = n - k; /* Im X[kk] */
        for (int64_t j = 0; j < n; ++j) s -= x[j] * sinpi_frac(2 * j * kk, n);
      }
      break;
    case 10: { /* HC2R: x_k = sum_j X_j exp(+2 pi i j k / n) */
      s = x[0];
      for (int64_t j = 1; 2 * j < n; ++j)
        s += 2.0L * (x[j] * cospi_frac(2 * j * k, n) - x[n - j] * sinpi_frac(2 * j * k, n));
      if (n % 2 == 0) s += ((k & 1) ? -1.0L : 1.0L) * x[n / 2];
      break;
    }
    }
    y[k] = (double)s;
  }
  return 1;
}

/* DCT-II / DST-II / DCT-III / DST-III through one length-n real FFT.
 * DCT-II follows src/r2r_buffer_helper.hpp:318-335 (even/odd-reversed
 * permutation) and :688-701 (y_k = 2 Re(V_k w_k), y_{n-k} = -2 Im(V_k w_k),
 * w_k = exp(-i pi k / 2n)); DST-II follows :457-470 (signed interleave) and
 * :830-857 (y_{i-1} = 2 Re(V_i / (sin t_i, -cos t_i)), t_i = pi i / 2n,
 * y_{n-1} = 2 Re V_0). DCT-III / DST-III use the inverse (Makhoul) steps. */
int orc_r2r_fast(int kind, int64_t n, const double *x, double *y) {
  if (n < 1) return 0;
  double *v = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  double *V = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  const long double pi = 3.141592653589793238462643383279502884L;
  int ok = 1;
  if (kind == 1 || kind == 5) {
    for (int64_t i = 0; i < n; ++i) {
      double val;
      int64_t j;
      if (kind == 1) {
        j = i;
        val = (i < (n + 1) / 2) ? x[2 * i] : x[2 * (n - i) - 1];
      } else {
        j = (i % 2) ? n - (i + 1) / 2 : i / 2;
        val = (i % 2) ? -x[i] : x[i];
        (void)0;
      }
      if (kind == 1) {
        v[2 * j] = val;
      } else {
        v[2 * j] = val;
      }
      v[2 * j + 1] = 0;
    }
    orc_fft(n, -1, v, V);
    if (kind == 1) {
      y[0] = 2 * V[0];
      for (int64_t k = 1; k <= n / 2; ++k) {
        long double wr = cosl(pi * k / (2 * n)), wi = -sinl(pi * k / (2 * n));
        long double jr = 2 * (V[2 * k] * wr - V[2 * k + 1] * wi);
        long double ji = 2 * (V[2 * k] * wi + V[2 * k + 1] * wr);
        y[k] = (double)jr;
        y[n - k] = (double)(-ji);
      }
    } else {
      for (int64_t i = 1; i < n; ++i) {
        long double tr, ti;
        if (i <= n / 2) {
          tr = V[2 * i];
          ti = V[2 * i + 1];
        } else {
          tr = V[2 * (n - i)];
          ti = -V[2 * (n - i) + 1];
        }
        long double t = pi * i / (2 * n);
        long double dr = sinl(t), di = -cosl(t); /* divide by (dr + i di) */
        long double den = dr * dr + di * di;
        long double qr = (tr * dr + ti * di) / den;
        y[i - 1] = (double)(2 * qr);
      }
      y[n - 1] = 2 * V[0];
    }
  } else if (kind == 2 || kind == 6) {
    /* DCT-III: inverse of the DCT-II steps. Build V_k = (y_k - i y_{n-k}) *
     * conj(w_k) / 2 style spectrum, inverse FFT, un-permute. For DST-III,
     * DST-III(x)_k = (-1)^k DCT-III(reverse(x))_k. */
    const double *src = x;
    double *rev = NULL;
    if (kind == 6) {
      rev = (double *)malloc(sizeof(double) * (size_t)n);
      for (int64_t i = 0; i < n; ++i) rev[i] = x[n - 1 - i];
      src = rev;
    }
    /* DCT-III(u)_j = sum_k c_k u_k cos(pi k (2j+1)/2n) with c_0=1, c_k=2:
     * Z_k = u_k - i u_{n-k} (u_n = 0), Z_k *= exp(i pi k / 2n),
     * z = backward FFT(Z) (full complex), v_j = Re z_j;
     * output[2i] = v_i, output[2(n-i)-1] = v_i for the upper half. */
    for (int64_t k = 0; k < n; ++k) {
      double ur = src[k];
      double ui = (k == 0) ? 0.0 : -src[n - k];
      long double c = cosl(pi * k / (2 * n)), s = sinl(pi * k / (2 * n));
      v[2 * k] = (double)(ur * c - ui * s);
      v[2 * k + 1] = (double)(ur * s + ui * c);
    }
    orc_fft(n, 1, v, V);
    for (int64_t i = 0; i < n; ++i) {
      int64_t dst = (i < (n + 1) / 2) ? 2 * i : 2 * (n - i) - 1;
      y[dst] = V[2 * i];
    }
    if (kind == 6) {
      for (int64_t k = 1; k < n; k += 2) y[k] = -y[k];
      free(rev);
    }
  } else {
    ok = orc_r2r_direct(kind, n, x, y);
  }
  free(v);
  free(V);
  return ok;
}

void orc_r2r_batched(int kind, int64_t n, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y) {
  double *in = (double *)malloc(sizeof(double) * (size_t)n);
  double *out = (double *)malloc(sizeof(double) * (size_t)n);
  for (int64_t b = 0; b < howmany; ++b) {
    for (int64_t j = 0; j < n; ++j) in[j] = x[b * idist + j * istride];
    if (!orc_r2r_fast(kind, n, in, out)) memset(out, 0, sizeof(double) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) y[b * odist + k * ostride] = out[k];
  }
  free(in);
  free(out);
}

/* ------------------------------------------------------------------------ */
/* fixed point                                                               */
/* ------------------------------------------------------------------------ */

/* src/plfft_convert.hpp:40-62: the value is first rounded to float, scaled
 * by 2^bits, truncated toward zero with int32 saturation (fcvtzs #bits), then
 * saturated to the storage width (sqxtn). 1.0 maps to 2^bits - 1. */
int32_t orc_q_convert(double v, int bits) {
  float f = (float)v;
  double scaled = (double)f * (double)(1LL << bits);
  int64_t t;
  if (scaled != scaled) t = 0;
  else if (scaled >= 2147483647.0) t = 2147483647LL;
  else if (scaled <= -2147483648.0) t = -2147483648LL;
  else t = (int64_t)scaled; /* C truncates toward zero */
  int64_t hi = (1LL << bits) - 1, lo = -(1LL << bits);
  if (bits == 7) { /* int32 -> int16 -> int8, each saturating */
    if (t > 32767) t = 32767;
    if (t < -32768) t = -32768;
  }
  if (t > hi) t = hi;
  if (t < lo) t = lo;
  return (int32_t)t;
}

/* ------------------------------------------------------------------------ */
/* fixed-point C2C (Q15 / Q7), power-of-two n                                */
/* ------------------------------------------------------------------------ */
/*
 * The reference's fixed-point transform (src/providers/jit/print_algo.cpp:
 * 1142-1239 input load, 2110-2196 twiddle multiply; src/providers/jit/
 * emit.hpp:150-430 saturating ops) is a Cooley-Tukey DIT over radix-2/4/8/16
 * codelets (src/providers/aot/aot_kernel_data.cpp:290-295) in which every
 * codelet first rounds-and-shifts its inputs right by log2(radix) (srshr),
 * then multiplies by the stage twiddles with one rounding per component
 * (smull/smlal + sqrshrn), then runs its butterflies with saturating
 * add/sub and rounded constant multiplies. Twiddles and constants are
 * sat(trunc(float(x) * 2^bits)) (src/plfft_convert.hpp:40-62), so 1.0 maps to
 * 2^bits - 1. The total effect is X/n.
 *
 * This restatement fixes one concrete op graph with those primitives, which
 * is the one the GPU kernels execute: a Stockham DIT with the radix chain
 * [2^(e mod 4) if nonzero, then 16, 16, ...] (e = log2 n), stage twiddle
 * W_{LR}^{j m} skipped when j m = 0, and the codelets below. The exact bits
 * of the reference's generated codelets (WFTA / split-radix op order, IR
 * rewrites) are not reproducible here, so bit parity with the reference is
 * unpinned; this function is pinned by the reference's own fixed-point test
 * semantics instead (tests/test_oracle.py: delta <-> constant known pairs and
 * DFT/n within 6 log2(n) LSB, test/test_utils.hpp:132-151).
 */
typedef struct {
  int64_t re, im;
} qc_t;

static int orc_q_bits_g = 15;
/* Test hook: plants a known mistake so that tests can show their checks
 * reject it (1: truncating shifts instead of rounding ones, 2: conjugated
 * stage twiddles). 0 in normal use. */
static int orc_fixed_plant_g = 0;
void orc_fixed_set_plant(int mode) { orc_fixed_plant_g = mode; }

static int64_t q_sat(int64_t v) {
  const int64_t hi = (1LL << orc_q_bits_g) - 1, lo = -(1LL << orc_q_bits_g);
  return v > hi ? hi : (v < lo ? lo : v);
}
static qc_t q_add(qc_t a, qc_t b) { qc_t r = {q_sat(a.re + b.re), q_sat(a.im + b.im)}; return r; }
static qc_t q_sub(qc_t a, qc_t b) { qc_t r = {q_sat(a.re - b.re), q_sat(a.im - b.im)}; return r; }
/* (x) * (-i) = (im, -re), the negation saturating */
static qc_t q_mul_negi(qc_t a) { qc_t r = {a.im, q_sat(-a.re)}; return r; }
/* floor(v / 2^s) */
static int64_t q_floor_shift(int64_t v, int s) {
  return v >= 0 ? (v >> s) : -((-v + (1LL << s) - 1) >> s);
}
/* rounding shift right (srshr): floor((v + 2^(s-1)) / 2^s) */
static int64_t q_rshr(int64_t v, int s) {
  if (s == 0) return v;
  if (orc_fixed_plant_g == 1) return q_floor_shift(v, s);
  return q_floor_shift(v + (1LL << (s - 1)), s);
}
/* complex multiply with one rounding per component (widening MAC + rounding
 * saturating narrow by `bits`) */
static qc_t q_cmul(qc_t x, qc_t w) {
  const int b = orc_q_bits_g;
  qc_t r;
  r.re = q_sat(q_rshr(x.re * w.re - x.im * w.im, b));
  r.im = q_sat(q_rshr(x.re * w.im + x.im * w.re, b));
  return r;
}

/* Q value of exp(-2 pi i num/den): (q(cos), q(-sin)) */
static qc_t q_root(int64_t num, int64_t den) {
  double s, c;
  orc_sincos_2pi_frac(num, den, &s, &c);
  qc_t r = {orc_q_convert(c, orc_q_bits_g), orc_q_convert(-s, orc_q_bits_g)};
  return r;
}

static void q_dft4(qc_t *a) {
  qc_t t0 = q_add(a[0], a[2]), t1 = q_sub(a[0], a[2]);
  qc_t t2 = q_add(a[1], a[3]), t3 = q_mul_negi(q_sub(a[1], a[3]));
  a[0] = q_add(t0, t2);
  a[2] = q_sub(t0, t2);
  a[1] = q_add(t1, t3);
  a[3] = q_sub(t1, t3);
}

/* radix-R butterfly on inputs already shifted; w16[k] = Q(W_16^k) */
static void q_dft(int R, qc_t *a, const qc_t *w16) {
  if (R == 1) return;
  if (R == 2) {
    qc_t x = q_add(a[0], a[1]), y = q_sub(a[0], a[1]);
    a[0] = x;
    a[1] = y;
    return;
  }
  if (R == 4) {
    q_dft4(a);
    return;
  }
  if (R == 8) {
    /* two radix-4 on even/odd inputs, internal twiddles W_8^k, radix-2 */
    qc_t e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
    q_dft4(e);
    q_dft4(o);
    o[1] = q_cmul(o[1], w16[2]);
    o[2] = q_mul_negi(o[2]);
    o[3] = q_cmul(o[3], w16[6]);
    for (int k = 0; k < 4; ++k) {
      a[k] = q_add(e[k], o[k]);
      a[k + 4] = q_sub(e[k], o[k]);
    }
    return;
  }
  /* R == 16: radix-4 over j (stride 4), internal W_16^{j k1}, radix-4 over k1 */
  qc_t f[4][4];
  for (int j = 0; j < 4; ++j) {
    for (int m = 0; m < 4; ++m) f[j][m] = a[j + 4 * m];
    q_dft4(f[j]);
  }
  for (int j = 1; j < 4; ++j)
    for (int k1 = 1; k1 < 4; ++k1) {
      const int e = j * k1;
      f[j][k1] = e == 4 ? q_mul_negi(f[j][k1]) : q_cmul(f[j][k1], w16[e]);
    }
  for (int k1 = 0; k1 < 4; ++k1) {
    qc_t g[4] = {f[0][k1], f[1][k1], f[2][k1], f[3][k1]};
    q_dft4(g);
    for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = g[k2];
  }
}

int orc_fixed_chain(int64_t n, int *chain) {
  if (n < 1 || (n & (n - 1))) return -1;
  int e = 0;
  while ((1LL << e) < n) ++e;
  int ns = 0;
  if (e % 4) chain[ns++] = 1 << (e % 4);
  for (int i = 0; i < e / 4; ++i) chain[ns++] = 16;
  return ns;
}

int orc_fixed_c2c(int64_t n, int dir, int bits, const int32_t *x, int32_t *y) {
  int chain[64];
  const int ns = orc_fixed_chain(n, chain);
  if (ns < 0 || (bits != 15 && bits != 7)) return -1;
  orc_q_bits_g = bits;
  qc_t *a = (qc_t *)malloc(sizeof(qc_t) * (size_t)n);
  qc_t *b = (qc_t *)malloc(sizeof(qc_t) * (size_t)n);
  qc_t w16[16];
  for (int k = 0; k < 16; ++k) w16[k] = q_root(k, 16);
  /* backward = swap o forward o swap */
  for (int64_t k = 0; k < n; ++k) {
    a[k].re = dir > 0 ? x[2 * k + 1] : x[2 * k];
    a[k].im = dir > 0 ? x[2 * k] : x[2 * k + 1];
  }
  int64_t L = 1;
  for (int s = 0; s < ns; ++s) {
    const int R = chain[s];
    int sh = 0;
    while ((1 << sh) < R) ++sh;
    const int64_t NB = n / R;
    for (int64_t bb = 0; bb < NB; ++bb) {
      const int64_t j = bb % L;
      qc_t v[16];
      for (int m = 0; m < R; ++m) {
        v[m].re = q_rshr(a[bb + m * NB].re, sh);
        v[m].im = q_rshr(a[bb + m * NB].im, sh);
      }
      if (j > 0)
        for (int m = 1; m < R; ++m)
          v[m] = q_cmul(v[m], q_root(orc_fixed_plant_g == 2 ? -j * m : j * m, L * R));
      q_dft(R, v, w16);
      const int64_t base = (bb / L) * L * R + j;
      for (int m = 0; m < R; ++m) b[base + m * L] = v[m];
    }
    qc_t *t = a;
    a = b;
    b = t;
    L *= R;
  }
  for (int64_t k = 0; k < n; ++k) {
    y[2 * k] = (int32_t)(dir > 0 ? a[k].im : a[k].re);
    y[2 * k + 1] = (int32_t)(dir > 0 ? a[k].re : a[k].im);
  }
  free(a);
  free(b);
  return 0;
}
