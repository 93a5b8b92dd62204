/*
 * plfft_gpu.h -- MI355X extensions to the plfft.h interface.
 *
 * None of these has a counterpart in the reference; they expose what a GPU
 * caller needs beyond the synchronous drop-in execute: stream ordering,
 * caller-owned workspaces (for hipGraph capture), adjoint execution, plan
 * introspection, the BF16 storage type and mixed-length batches.
 */
#ifndef PLFFT_GPU_H_
#define PLFFT_GPU_H_

#include <stddef.h>
#include <stdint.h>

#include "plfft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* bfloat16 storage with FP32 arithmetic (the reference has no bf16). Pass to
 * plfft_config_set_data_type(). */
#define PLFFT_DATA_TYPE_BF16 ((plfft_data_type_t)5)

/* Device bytes of scratch that one execution of `plan` needs (0 for
 * single-kernel plans). */
size_t plfft_plan_get_workspace_size(const plfft_plan_t *plan);

/* Enqueue the transform on `stream` (a hipStream_t; NULL is the default
 * stream) and return without waiting. `input`/`output` must be device
 * memory. `workspace` is device memory of plfft_plan_get_workspace_size()
 * bytes, or NULL to use the plan's own workspace (then concurrent async
 * executions of this plan on different streams race on it). Allocates
 * nothing when `workspace` is given or the plan needs none, so it can be
 * captured into a hipGraph. Returns PLFFT_NULL_ARGUMENT for null plan or
 * buffers, PLFFT_PLAN_CREATION_FAILED if the launch fails. */
plfft_status_t plfft_plan_execute_async(const plfft_plan_t *plan,
                                        const void *input, void *output,
                                        void *workspace, void *stream);

/* Adjoint of the planned linear map with respect to the real inner product
 * <a, b> = sum Re(a_k conj(b_k)): `input` has the layout of the plan's
 * OUTPUT, `output` the layout of the plan's INPUT (strides, distances and
 * element counts swapped accordingly). C2C forward -> C2C backward; R2C ->
 * Hermitian-weighted C2R; C2R -> R2C with doubled interior bins;
 * DCT-II <-> DCT-III and DST-II <-> DST-III with the end weight. */
void plfft_plan_execute_adjoint(const plfft_plan_t *plan, const void *input,
                                void *output);
plfft_status_t plfft_plan_execute_adjoint_async(const plfft_plan_t *plan,
                                                const void *input, void *output,
                                                void *workspace, void *stream);

/* Human-readable plan tree (the reference's plan_to_string). Writes at most
 * `len` bytes including the terminator; returns the full length. */
size_t plfft_plan_describe(const plfft_plan_t *plan, char *buf, size_t len);

/* Text of the last plan-creation failure on this thread ("" if none). */
const char *plfft_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PLFFT_GPU_H_ */
