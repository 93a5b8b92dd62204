/*
 * plfft.h -- one-dimensional FFT plan/execute C interface, MI355X build.
 *
 * Drop-in for the reference interface include/plfft.h:47-532 of
 * arm/perf-libs-fft: the same 23 functions, the same enumerator values and the
 * same transform, layout, batching, aliasing and scaling rules. Only the
 * meaning of the execute pointers is widened: they may be device pointers
 * (hipMalloc / managed memory) or ordinary host pointers. Host buffers are
 * staged through device memory; device buffers are used in place. Execution is
 * synchronous, as in the reference.
 *
 * GPU-only extensions (stream-ordered execute, adjoint execute, workspace
 * query, mixed-length batches, the BF16 storage type) live in plfft_gpu.h so
 * that this header stays ABI-identical.
 */
#ifndef PLFFT_H_
#define PLFFT_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque configuration; replaces include/plfft.h:47 (plfft_config_t). */
typedef struct plfft_config_t plfft_config_t;

/* Opaque executable plan; replaces include/plfft.h:55 (plfft_plan_t). */
typedef struct plfft_plan_t plfft_plan_t;

/* Status codes; values are ABI (include/plfft.h:58-89). */
typedef enum plfft_status_t {
  PLFFT_OK = 0,
  PLFFT_NULL_ARGUMENT = 1,
  PLFFT_INVALID_SIZE = 2,
  PLFFT_UNSUPPORTED_BATCH = 3,
  PLFFT_INVALID_TRANSFORM_KIND = 4,
  PLFFT_INVALID_DATA_TYPE = 5,
  PLFFT_INVALID_IO_ALIAS = 6,
  PLFFT_PLAN_CREATION_FAILED = 7,
  PLFFT_INVALID_SME_MODE = 8,
  PLFFT_SME_NOT_ENABLED = 9,
  PLFFT_SME_UNSUPPORTED_R2R = 10,
  PLFFT_FP16_NOT_AVAILABLE = 11,
  PLFFT_SME_NOT_AVAILABLE = 12,
  PLFFT_SME_UNSUPPORTED_FIXED_POINT = 13,
  PLFFT_INVALID_PLANNER_RIGOR = 14
} plfft_status_t;

/* C2C direction: sign of the exponent (include/plfft.h:92-97). */
typedef enum plfft_direction_t {
  PLFFT_FORWARD = -1,
  PLFFT_BACKWARD = 1
} plfft_direction_t;

/* Unnormalized real-to-real kinds, FFTW conventions (include/plfft.h:105-134).
 * R2HC stores Re X[k] at k (0 <= k <= n/2) and Im X[k] at n-k
 * (1 <= k <= (n-1)/2); HC2R inverts it. DCT_1 needs n > 1. */
typedef enum plfft_r2r_kind_t {
  PLFFT_R2R_DCT_1 = 0,
  PLFFT_R2R_DCT_2 = 1,
  PLFFT_R2R_DCT_3 = 2,
  PLFFT_R2R_DCT_4 = 3,
  PLFFT_R2R_DST_1 = 4,
  PLFFT_R2R_DST_2 = 5,
  PLFFT_R2R_DST_3 = 6,
  PLFFT_R2R_DST_4 = 7,
  PLFFT_R2R_DHT = 8,
  PLFFT_R2R_R2HC = 9,
  PLFFT_R2R_HC2R = 10
} plfft_r2r_kind_t;

/* Scalar storage type of every real component (include/plfft.h:143-154).
 * FP16 is IEEE binary16 (uint16 bit patterns on the host side); Q7/Q15 are
 * signed fixed point with an implicit 1/n scaling per transform. The
 * extension value PLFFT_DATA_TYPE_BF16 (= 5) is defined in plfft_gpu.h. */
typedef enum plfft_data_type_t {
  PLFFT_DATA_TYPE_FP32 = 0,
  PLFFT_DATA_TYPE_FP64 = 1,
  PLFFT_DATA_TYPE_FP16 = 2,
  PLFFT_DATA_TYPE_Q7 = 3,
  PLFFT_DATA_TYPE_Q15 = 4
} plfft_data_type_t;

/* Input/output overlap permission (include/plfft.h:157-170). With MAY_ALIAS a
 * batch member's own input and output may overlap (exactly or partially);
 * members never overlap each other's output. */
typedef enum plfft_io_alias_t {
  PLFFT_IO_NO_ALIAS = 0,
  PLFFT_IO_MAY_ALIAS = 1
} plfft_io_alias_t;

/* SME request (include/plfft.h:178-190). This build has no SME: enabling it
 * makes plan creation return PLFFT_SME_NOT_ENABLED, as a reference build
 * without SME support does. */
typedef enum plfft_sme_mode_t {
  PLFFT_SME_DISABLED = 0,
  PLFFT_SME_ENABLED = 1
} plfft_sme_mode_t;

/* Planning effort (include/plfft.h:193-201). ESTIMATE uses the heuristic or a
 * cached tuning result; MEASURE and above time candidate GPU kernels within
 * the reference's per-rigor planning-time budget. */
typedef enum plfft_rigor_t {
  PLFFT_ESTIMATE = 0,
  PLFFT_MEASURE = 1,
  PLFFT_PATIENT = 2,
  PLFFT_EXHAUSTIVE = 3,
} plfft_rigor_t;

/* include/plfft.h:212 -- static string, "unknown status" for other values. */
const char *plfft_status_to_string(plfft_status_t status);

/* include/plfft.h:235 -- C2C configuration; invalid direction gives
 * PLFFT_INVALID_TRANSFORM_KIND and leaves *config NULL. */
plfft_status_t plfft_config_create(plfft_config_t **config, int64_t n,
                                   plfft_direction_t direction);

/* include/plfft.h:256 -- forward R2C: n reals -> n/2+1 complex. */
plfft_status_t plfft_config_create_r2c(plfft_config_t **config, int64_t n);

/* include/plfft.h:277 -- backward C2R: n/2+1 complex -> n reals. */
plfft_status_t plfft_config_create_c2r(plfft_config_t **config, int64_t n);

/* include/plfft.h:300 -- R2R of the given kind. */
plfft_status_t plfft_config_create_r2r(plfft_config_t **config, int64_t n,
                                       plfft_r2r_kind_t r2r_kind);

/* include/plfft.h:312 -- NULL is accepted. */
void plfft_config_destroy(plfft_config_t *config);

/* include/plfft.h:324-394 -- setters store without validation. */
void plfft_config_set_data_type(plfft_config_t *config,
                                plfft_data_type_t precision);
plfft_data_type_t plfft_config_get_data_type(const plfft_config_t *config);
void plfft_config_set_io_alias(plfft_config_t *config, plfft_io_alias_t alias);
plfft_io_alias_t plfft_config_get_io_alias(const plfft_config_t *config);
void plfft_config_set_sme_mode(plfft_config_t *config,
                               plfft_sme_mode_t sme_mode);
plfft_sme_mode_t plfft_config_get_sme_mode(const plfft_config_t *config);
void plfft_config_set_rigor(plfft_config_t *config, plfft_rigor_t rigor);
plfft_rigor_t plfft_config_get_rigor(const plfft_config_t *config);

/* include/plfft.h:418 -- element k of member b lives at
 * input[b*idist + k*istride] / output[b*odist + k*ostride], counted in
 * logical elements (real scalars on the real side, complex pairs on the
 * complex side). */
void plfft_config_set_batch(plfft_config_t *config, int64_t howmany,
                            int64_t istride, int64_t idist, int64_t ostride,
                            int64_t odist);

/* include/plfft.h:427-467 */
int64_t plfft_config_get_howmany(const plfft_config_t *config);
int64_t plfft_config_get_istride(const plfft_config_t *config);
int64_t plfft_config_get_idist(const plfft_config_t *config);
int64_t plfft_config_get_ostride(const plfft_config_t *config);
int64_t plfft_config_get_odist(const plfft_config_t *config);

/* include/plfft.h:498 -- validates in the reference order
 * (src/plfft.cpp:427-517), then builds the GPU plan on the current HIP
 * device. No usable GPU gives PLFFT_PLAN_CREATION_FAILED. */
plfft_status_t plfft_plan_create_from_config(plfft_plan_t **plan,
                                             const plfft_config_t *config);

/* include/plfft.h:509 -- NULL is accepted. */
void plfft_plan_destroy(plfft_plan_t *plan);

/* include/plfft.h:531 -- synchronous execute. Pointers may be device or host
 * memory; a plan may be executed concurrently from several threads. */
void plfft_plan_execute(const plfft_plan_t *plan, const void *input,
                        void *output);

#ifdef __cplusplus
}
#endif

#endif /* PLFFT_H_ */
