/*
 * oracle.h -- CPU restatement of the reference's 1D transform path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code; only tests/, the smoke check and bench.py's cpu_baseline
 * leg use it, and only as the checker (or as the timed CPU baseline).
 *
 * Complex data are interleaved (re, im) scalars. Every function cites the
 * reference file:line whose behaviour it restates.
 */
#ifndef PLFFT_ORACLE_H_
#define PLFFT_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded inputs (counter-based, bit-identical everywhere) ---------- */
uint64_t orc_splitmix64(uint64_t *state);
void orc_fill_uniform_f64(uint64_t seed, int64_t count, double *out);
void orc_fill_uniform_f32(uint64_t seed, int64_t count, float *out);
void orc_fill_normal_f64(uint64_t seed, int64_t count, double *out);

/* ---- integer helpers -------------------------------------------------- */
int64_t orc_smallest_prime_factor(int64_t n);
int64_t orc_next_length_factorable_by(int64_t n, const int *factors, int nf);
int64_t orc_find_group_generator(int64_t n);
void orc_rader_permutations(int64_t n, int64_t *perm, int64_t *perm_rev_idx);
void orc_bluestein_phase(int64_t n, int64_t *phase);
int64_t orc_bluestein_length(int64_t n);

/* ---- floating-point transforms (FP64) --------------------------------- */
void orc_dft(int64_t n, int dir, const double *x, double *y);
void orc_fft(int64_t n, int dir, const double *x, double *y);
void orc_fft_f32(int64_t n, int dir, const float *x, float *y);
int orc_fft_describe(int64_t n, char *buf, int len);

void orc_c2c_batched(int64_t n, int dir, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y);
void orc_c2c_batched_f32(int64_t n, int dir, int64_t howmany, int64_t istride,
                         int64_t idist, int64_t ostride, int64_t odist,
                         const float *x, float *y, int threads);
void orc_r2c_batched(int64_t n, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y);
void orc_c2r_batched(int64_t n, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y);

/* Real-to-real, FFTW definitions evaluated directly (O(n^2), long double). */
int orc_r2r_direct(int kind, int64_t n, const double *x, double *y);
/* DCT-II / DST-II through a length-n R2C, following the reference's
 * Makhoul-style pre/post processing. */
int orc_r2r_fast(int kind, int64_t n, const double *x, double *y);
void orc_r2r_batched(int kind, int64_t n, int64_t howmany, int64_t istride,
                     int64_t idist, int64_t ostride, int64_t odist,
                     const double *x, double *y);

/* ---- fixed point (Q0.15 / Q0.7) --------------------------------------- */
int32_t orc_q_convert(double v, int bits);
/* Radix chain of the fixed-point transform of n (power of two); returns the
 * number of stages, or -1 if n is not a power of two. */
int orc_fixed_chain(int64_t n, int *chain);
/* One forward radix-R codelet (R = 2, 4, 8, 16) of the fixed-point
 * transform on R pairs already shifted and twiddled; returns 0, or -1 for
 * an unsupported R or bits. */
int orc_fixed_codelet(int R, int bits, const int32_t *x, int32_t *y);
/* Fixed-point C2C of n = 2^e points, bits 15 (Q0.15) or 7 (Q0.7); x and y
 * hold interleaved values in int32 (within the storage range). dir -1
 * forward, +1 backward. Output = DFT(x)/n rounded stage by stage. */
int orc_fixed_c2c(int64_t n, int dir, int bits, const int32_t *x, int32_t *y);
/* Test hook for tests/test_oracle.py: plant a known mistake (1 truncating
 * shifts, 2 conjugated stage twiddles; 0 = none). */
void orc_fixed_set_plant(int mode);

#ifdef __cplusplus
}
#endif

#endif
