/*
 * oracle_fft_impl.h -- precision-generic body of the oracle's planned FFT.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Included by oracle.c once with
 * REAL=double / SFX=_f64 and once with REAL=float / SFX=_f32.
 *
 * The plan tree restates the reference planner's first-fit choice under
 * ESTIMATE for an AOT build (kernel ns 2..20):
 *   - direct codelet for n <= 20
 *     (src/planner/strategies/direct.hpp:26-97, src/providers/aot/aot_kernel_data.cpp:297-298);
 *   - Cooley-Tukey DIT with r = the first radix in the descending list
 *     {40..2} that is a codelet size and divides n
 *     (src/planner/planner.cpp:104-137, src/planner/strategies/cooley_tukey.hpp:622-679):
 *     stage1 = r interleaved m-point sub-transforms, stage2 = twiddled r-point
 *     DFTs with W[i][k] = exp(dir*2*pi*i*k/n) (cooley_tukey.hpp:82-143);
 *   - Rader for prime n whose n-1 is 20-smooth
 *     (src/planner/strategies/rader.hpp:36-124, 816-875);
 *   - Bluestein otherwise (src/planner/strategies/bluestein.hpp:42-146);
 *   - Cooley-Tukey with a prime radix above 20 for composite n without a
 *     codelet divisor (cooley_tukey_rader_ab, rader.hpp:877-1040), the prime
 *     radix evaluated directly.
 * Twiddles and chirps are computed in double from integer-exact angles and
 * rounded to REAL on store (src/generate_twiddles.hpp:72-136).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define CPX CAT(ocpx, SFX)
#define NODE CAT(onode, SFX)

typedef struct {
  REAL re, im;
} CPX;

typedef struct NODE NODE;
struct NODE {
  int kind; /* 0 direct, 1 Cooley-Tukey, 2 Rader, 3 Bluestein */
  int dir;
  int64_t n, r, m;
  NODE *sub;      /* CT: m-point; Rader/Bluestein: forward sub-plan */
  NODE *sub_bwd;  /* Rader/Bluestein: backward sub-plan */
  NODE *radix;    /* CT with prime radix > 20: r-point plan */
  CPX *tw;        /* direct: W_n^k (n); CT: W_n^{i k} ((r-1)*m) */
  CPX *wr;        /* CT: W_r^q (r), for the direct r-point DFT */
  CPX *chirp;     /* Bluestein: w[k] (n) */
  CPX *bfft;      /* Rader: FFT(b) (m); Bluestein: FFT(b) (n_pad) */
  int64_t *perm, *perm_rev;
  int64_t scratch; /* complex elements of scratch this node needs */
};

static CPX CAT(cmul, SFX)(CPX a, CPX b) {
  CPX c;
  c.re = a.re * b.re - a.im * b.im;
  c.im = a.re * b.im + a.im * b.re;
  return c;
}

static CPX CAT(unit_root, SFX)(int dir, int64_t num, int64_t den) {
  /* exp(dir * 2*pi*i * num/den), num reduced mod den exactly */
  CPX w;
  double s, c;
  orc_sincos_2pi_frac(num, den, &s, &c);
  w.re = (REAL)c;
  w.im = (REAL)(dir * s);
  return w;
}

static NODE *CAT(plan_build, SFX)(int64_t n, int dir);
static void CAT(plan_exec, SFX)(const NODE *p, const CPX *x, int64_t xs, CPX *y,
                                CPX *scr);

static NODE *CAT(plan_build, SFX)(int64_t n, int dir) {
  NODE *p = (NODE *)calloc(1, sizeof(NODE));
  p->n = n;
  p->dir = dir;
  if (n <= ORC_MAX_CODELET) {
    p->kind = 0;
    p->tw = (CPX *)malloc(sizeof(CPX) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) p->tw[k] = CAT(unit_root, SFX)(dir, k, n);
    p->scratch = n;
    return p;
  }
  int64_t r = orc_ct_radix(n);
  if (r == 0 && orc_smallest_prime_factor(n) != n) r = orc_smallest_prime_factor(n);
  if (r) {
    p->kind = 1;
    p->r = r;
    p->m = n / r;
    p->sub = CAT(plan_build, SFX)(p->m, dir);
    if (r > ORC_MAX_CODELET) p->radix = CAT(plan_build, SFX)(r, dir);
    p->tw = (CPX *)malloc(sizeof(CPX) * (size_t)((r - 1) * p->m));
    for (int64_t i = 1; i < r; ++i)
      for (int64_t k = 0; k < p->m; ++k)
        p->tw[(i - 1) * p->m + k] = CAT(unit_root, SFX)(dir, i * k, n);
    p->wr = (CPX *)malloc(sizeof(CPX) * (size_t)r);
    for (int64_t q = 0; q < r; ++q) p->wr[q] = CAT(unit_root, SFX)(dir, q, r);
    int64_t s = p->sub->scratch;
    int64_t need = 2 * r + (p->radix ? p->radix->scratch : 0);
    if (need > s) s = need;
    p->scratch = n + s;
    return p;
  }
  /* n prime and larger than every codelet */
  int64_t m = n - 1;
  if (orc_is_smooth(m, ORC_MAX_CODELET)) {
    p->kind = 2;
    p->m = m;
    p->sub = CAT(plan_build, SFX)(m, -1);
    p->sub_bwd = CAT(plan_build, SFX)(m, 1);
    p->perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
    p->perm_rev = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    orc_rader_permutations(n, p->perm, p->perm_rev);
    /* b[(m-k)%m] = exp(dir*2*pi*i*perm[k]/n) / m  (rader.hpp:64-77) */
    CPX *b = (CPX *)malloc(sizeof(CPX) * (size_t)m);
    for (int64_t k = 0; k < m; ++k) {
      double s, c;
      orc_sincos_2pi_frac(p->perm[k], n, &s, &c);
      b[(m - k) % m].re = (REAL)(c / (double)m);
      b[(m - k) % m].im = (REAL)(dir * s / (double)m);
    }
    p->bfft = (CPX *)malloc(sizeof(CPX) * (size_t)m);
    CPX *scr = (CPX *)malloc(sizeof(CPX) * (size_t)(p->sub->scratch + 1));
    CAT(plan_exec, SFX)(p->sub, b, 1, p->bfft, scr);
    free(scr);
    free(b);
    p->scratch = 2 * m + p->sub->scratch;
    return p;
  }
  /* Bluestein (bluestein.hpp:42-110) */
  {
    int64_t np = orc_bluestein_length(n);
    p->kind = 3;
    p->m = np;
    p->sub = CAT(plan_build, SFX)(np, -1);
    p->sub_bwd = CAT(plan_build, SFX)(np, 1);
    p->chirp = (CPX *)malloc(sizeof(CPX) * (size_t)n);
    double *wd = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int64_t k = 0; k < n; ++k) {
      /* w[k] = exp(i*dir*pi*(k^2 mod 2n)/n) = exp(dir*2*pi*i*(k^2 mod 2n)/(2n)) */
      int64_t ph = (int64_t)(((unsigned __int128)k * (unsigned __int128)k) % (unsigned __int128)(2 * n));
      double s, c;
      orc_sincos_2pi_frac(ph, 2 * n, &s, &c);
      wd[2 * k] = c;
      wd[2 * k + 1] = dir * s;
      p->chirp[k].re = (REAL)c;
      p->chirp[k].im = (REAL)(dir * s);
    }
    CPX *b = (CPX *)calloc((size_t)np, sizeof(CPX));
    const double rn = 1.0 / (double)np;
    /* b[0] = conj(w0)/np; b[k] = b[np-k] = conj(w[k])/np (bluestein.hpp:81-98) */
    b[0].re = (REAL)(wd[0] * rn);
    b[0].im = (REAL)(-wd[1] * rn);
    for (int64_t k = 1; k < n; ++k) {
      CPX v;
      v.re = (REAL)(wd[2 * k] * rn);
      v.im = (REAL)(-wd[2 * k + 1] * rn);
      b[k] = v;
      b[np - k] = v;
    }
    free(wd);
    p->bfft = (CPX *)malloc(sizeof(CPX) * (size_t)np);
    CPX *scr = (CPX *)malloc(sizeof(CPX) * (size_t)(p->sub->scratch + 1));
    CAT(plan_exec, SFX)(p->sub, b, 1, p->bfft, scr);
    free(scr);
    free(b);
    p->scratch = 2 * np + p->sub->scratch;
    return p;
  }
}

static void CAT(plan_free, SFX)(NODE *p) {
  if (!p) return;
  CAT(plan_free, SFX)(p->sub);
  CAT(plan_free, SFX)(p->sub_bwd);
  CAT(plan_free, SFX)(p->radix);
  free(p->tw);
  free(p->wr);
  free(p->chirp);
  free(p->bfft);
  free(p->perm);
  free(p->perm_rev);
  free(p);
}

/* y[0..n) (contiguous) = transform of x[0], x[xs], ..., x[(n-1)xs]. */
static void CAT(plan_exec, SFX)(const NODE *p, const CPX *x, int64_t xs, CPX *y,
                                CPX *scr) {
  const int64_t n = p->n;
  switch (p->kind) {
  case 0: { /* direct codelet: O(n^2) with the W_n table */
    for (int64_t k = 0; k < n; ++k) {
      REAL sr = 0, si = 0;
      int64_t idx = 0;
      for (int64_t j = 0; j < n; ++j) {
        CPX v = CAT(cmul, SFX)(x[j * xs], p->tw[idx]);
        sr += v.re;
        si += v.im;
        idx += k;
        if (idx >= n) idx -= n;
      }
      scr[k].re = sr;
      scr[k].im = si;
    }
    for (int64_t k = 0; k < n; ++k) y[k] = scr[k];
    return;
  }
  case 1: { /* Cooley-Tukey DIT: stage1 into scr, stage2 into y */
    const int64_t r = p->r, m = p->m;
    CPX *tmp = scr;
    CPX *deeper = scr + n;
    for (int64_t i = 0; i < r; ++i)
      CAT(plan_exec, SFX)(p->sub, x + i * xs, xs * r, tmp + i * m, deeper);
    CPX *v = deeper;       /* r inputs */
    CPX *vo = deeper + r;  /* r outputs */
    for (int64_t k = 0; k < m; ++k) {
      v[0] = tmp[k];
      for (int64_t i = 1; i < r; ++i)
        v[i] = CAT(cmul, SFX)(tmp[i * m + k], p->tw[(i - 1) * m + k]);
      if (p->radix) {
        CAT(plan_exec, SFX)(p->radix, v, 1, vo, deeper + 2 * r);
      } else {
        for (int64_t q = 0; q < r; ++q) {
          REAL sr = 0, si = 0;
          int64_t idx = 0;
          for (int64_t i = 0; i < r; ++i) {
            CPX t = CAT(cmul, SFX)(v[i], p->wr[idx]);
            sr += t.re;
            si += t.im;
            idx += q;
            if (idx >= r) idx -= r;
          }
          vo[q].re = sr;
          vo[q].im = si;
        }
      }
      for (int64_t q = 0; q < r; ++q) y[k + q * m] = vo[q];
    }
    return;
  }
  case 2: { /* Rader: pack, convolve, unpack (rader.hpp:110-124,145-223) */
    const int64_t m = p->m;
    CPX *a = scr;
    CPX *a2 = scr + m;
    CPX *deeper = scr + 2 * m;
    CPX x0 = x[0];
    REAL s0r = x0.re, s0i = x0.im;
    for (int64_t k = 0; k < m; ++k) {
      a[k] = x[p->perm[k] * xs];
      s0r += a[k].re;
      s0i += a[k].im;
    }
    CAT(plan_exec, SFX)(p->sub, a, 1, a2, deeper);
    for (int64_t k = 0; k < m; ++k) a2[k] = CAT(cmul, SFX)(a2[k], p->bfft[k]);
    CAT(plan_exec, SFX)(p->sub_bwd, a2, 1, a, deeper);
    y[0].re = s0r;
    y[0].im = s0i;
    for (int64_t k = 1; k < n; ++k) {
      CPX c = a[p->perm_rev[k]];
      y[k].re = x0.re + c.re;
      y[k].im = x0.im + c.im;
    }
    return;
  }
  case 3: { /* Bluestein: a = x.w, a = ifft(fft(a).B), y = a.w */
    const int64_t np = p->m;
    CPX *a = scr;
    CPX *a2 = scr + np;
    CPX *deeper = scr + 2 * np;
    for (int64_t k = 0; k < n; ++k) a[k] = CAT(cmul, SFX)(x[k * xs], p->chirp[k]);
    for (int64_t k = n; k < np; ++k) a[k].re = a[k].im = 0;
    CAT(plan_exec, SFX)(p->sub, a, 1, a2, deeper);
    for (int64_t k = 0; k < np; ++k) a2[k] = CAT(cmul, SFX)(a2[k], p->bfft[k]);
    CAT(plan_exec, SFX)(p->sub_bwd, a2, 1, a, deeper);
    for (int64_t k = 0; k < n; ++k) y[k] = CAT(cmul, SFX)(a[k], p->chirp[k]);
    return;
  }
  }
}

#undef CPX
#undef NODE
