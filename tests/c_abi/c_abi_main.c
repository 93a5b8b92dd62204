/*
 * A plain C99 caller of the 23 functions of include/plfft.h, compiled with
 * gcc and linked against libplfft_amd.so: the role of the reference's
 * test/c_compat (a C program linking the library without C++). It exercises
 * every function and, when a GPU is present, a forward + backward C2C round
 * trip on host buffers (staged by the library) plus an R2C/C2R and an R2R
 * plan. Exit status 0 on success, 1 on a failed check; prints "no device"
 * and exits 0 when plan creation reports that no GPU is available.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "plfft.h"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main(void) {
  const int64_t n = 10;
  plfft_config_t *cf = NULL, *cb = NULL, *cr = NULL, *cc = NULL, *cd = NULL;
  plfft_plan_t *pf = NULL, *pb = NULL, *pr = NULL, *pc = NULL, *pd = NULL;

  /* status strings (include/plfft.h: plfft_status_to_string) */
  CHECK(strcmp(plfft_status_to_string(PLFFT_OK), "success") == 0);
  CHECK(strcmp(plfft_status_to_string((plfft_status_t)99), "unknown status") == 0);

  /* creators and defaults */
  CHECK(plfft_config_create(&cf, n, PLFFT_FORWARD) == PLFFT_OK);
  CHECK(plfft_config_create(&cb, n, PLFFT_BACKWARD) == PLFFT_OK);
  CHECK(plfft_config_create_r2c(&cr, n) == PLFFT_OK);
  CHECK(plfft_config_create_c2r(&cc, n) == PLFFT_OK);
  CHECK(plfft_config_create_r2r(&cd, n, PLFFT_R2R_DCT_2) == PLFFT_OK);
  CHECK(plfft_config_get_data_type(cf) == PLFFT_DATA_TYPE_FP32);
  CHECK(plfft_config_get_io_alias(cf) == PLFFT_IO_NO_ALIAS);
  CHECK(plfft_config_get_sme_mode(cf) == PLFFT_SME_DISABLED);
  CHECK(plfft_config_get_rigor(cf) == PLFFT_MEASURE);
  CHECK(plfft_config_get_howmany(cf) == 1);
  CHECK(plfft_config_get_istride(cf) == 1 && plfft_config_get_ostride(cf) == 1);
  CHECK(plfft_config_get_idist(cf) == n && plfft_config_get_odist(cf) == n);

  /* setters store without validation */
  plfft_config_set_data_type(cf, PLFFT_DATA_TYPE_FP64);
  plfft_config_set_data_type(cb, PLFFT_DATA_TYPE_FP64);
  plfft_config_set_io_alias(cf, PLFFT_IO_MAY_ALIAS);
  plfft_config_set_sme_mode(cf, PLFFT_SME_DISABLED);
  plfft_config_set_rigor(cf, PLFFT_ESTIMATE);
  plfft_config_set_rigor(cb, PLFFT_ESTIMATE);
  plfft_config_set_batch(cf, 1, 1, n, 1, n);
  CHECK(plfft_config_get_data_type(cf) == PLFFT_DATA_TYPE_FP64);
  CHECK(plfft_config_get_io_alias(cf) == PLFFT_IO_MAY_ALIAS);
  CHECK(plfft_config_get_rigor(cf) == PLFFT_ESTIMATE);

  /* plans */
  plfft_status_t st = plfft_plan_create_from_config(&pf, cf);
  if (st == PLFFT_PLAN_CREATION_FAILED) {
    CHECK(pf == NULL);
    puts("no device");
    plfft_config_destroy(cf);
    plfft_config_destroy(cb);
    plfft_config_destroy(cr);
    plfft_config_destroy(cc);
    plfft_config_destroy(cd);
    plfft_plan_destroy(NULL);
    return 0;
  }
  CHECK(st == PLFFT_OK && pf != NULL);
  CHECK(plfft_plan_create_from_config(&pb, cb) == PLFFT_OK);
  CHECK(plfft_plan_create_from_config(&pr, cr) == PLFFT_OK);
  CHECK(plfft_plan_create_from_config(&pc, cc) == PLFFT_OK);
  CHECK(plfft_plan_create_from_config(&pd, cd) == PLFFT_OK);
  plfft_config_destroy(cf);
  plfft_config_destroy(cb);
  plfft_config_destroy(cr);
  plfft_config_destroy(cc);
  plfft_config_destroy(cd);

  /* C2C round trip in FP64: in[i] = i, backward(forward(in)) = n * in,
     the check of test/c_compat/c_test.cpp */
  double in[2 * 10], out[2 * 10], back[2 * 10];
  for (int i = 0; i < n; ++i) {
    in[2 * i] = i;
    in[2 * i + 1] = 0.0;
  }
  plfft_plan_execute(pf, in, out);
  plfft_plan_execute(pb, out, back);
  const double tol = 10 * 2 * 2.220446049250313e-16 * 4 * n * log2((double)n);
  for (int i = 0; i < n; ++i) {
    const double res = back[2 * i] / n, ref = in[2 * i];
    const double diff = fabs(res - ref);
    CHECK((ref > tol ? diff / ref : diff) <= tol);
    CHECK(fabs(back[2 * i + 1]) / n <= tol);
  }
  /* X[0] of the forward transform is the sum 0 + 1 + ... + 9 */
  CHECK(fabs(out[0] - 45.0) < 1e-12 && fabs(out[1]) < 1e-12);

  /* R2C then C2R in FP32: n * x back */
  float xr[10], spec[2 * 6], yr[10];
  for (int i = 0; i < n; ++i) xr[i] = (float)((i % 3) - 1);
  plfft_plan_execute(pr, xr, spec);
  plfft_plan_execute(pc, spec, yr);
  for (int i = 0; i < n; ++i) CHECK(fabs(yr[i] / n - xr[i]) < 1e-5f);

  /* DCT-II: y[0] = 2 * sum x (test/r2r_test.hpp first-element identity) */
  float dct[10];
  plfft_plan_execute(pd, xr, dct);
  float s = 0;
  for (int i = 0; i < n; ++i) s += xr[i];
  CHECK(fabs(dct[0] - 2 * s) < 1e-5f);

  plfft_plan_destroy(pf);
  plfft_plan_destroy(pb);
  plfft_plan_destroy(pr);
  plfft_plan_destroy(pc);
  plfft_plan_destroy(pd);
  plfft_plan_destroy(NULL);
  puts("ok");
  return 0;
}
