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

/* Bytes of the fixed device scratch array behind plfft_plan_execute_noalloc:
 * 64 MiB on every device that runs it, held until the process ends. */
#define PLFFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory: code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc. Same transform and stream ordering as
 * plfft_plan_execute_async(plan, input, output, <workspace>, stream), but the
 * plan's scratch is a fixed array of PLFFT_NOALLOC_SCRATCH_BYTES bytes per
 * device inside this library; the call allocates nothing and issues kernel
 * launches only. Plans that stream their batch through the scratch in
 * chunks (four-step) shrink the chunk to fit it. `input`/`output` must be
 * device memory. Every call on one device shares that device's array, so
 * calls that overlap in time on one device (other streams, other threads)
 * race on it: order them on one stream or serialize them. Returns
 * PLFFT_NULL_ARGUMENT (null plan or buffers), PLFFT_INVALID_SIZE when the
 * plan needs more scratch than the array holds (see plfft_last_error()), or
 * PLFFT_PLAN_CREATION_FAILED when the scratch lookup or a launch fails. */
plfft_status_t plfft_plan_execute_noalloc(const plfft_plan_t *plan, const void *input,
                                          void *output, void *stream);

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

/* ------------------------------------------------------------------------
 * Mixed-length batches: one plan transforms `count` complex-to-complex
 * members of different lengths (powers of two, other smooth lengths and
 * primes alike) in a fixed number of kernel launches: at most two for the
 * one-pass and four-step members (one per dependency level), one per group
 * of one-block Bluestein transforms of equal padded length (whole members,
 * and the rows of members split as n1 x n2 with a large prime factor in n2;
 * at most seven groups per kind), and three per group of chirp-z members of
 * equal padded length (at most ten groups for lengths up to 65536). The
 * reference has no such object: a
 * reference caller creates one plan per length (plfft_plan_create_from_config,
 * include/plfft.h:491-501) and executes them one after another; this plan
 * replaces that loop.
 *
 * Member i reads n[i] contiguous complex elements at input + in_offset[i] and
 * writes its transform to output + out_offset[i] (offsets in complex elements
 * of the data type; NULL packs the members back to back). Members may be
 * transformed in place (equal offsets, input == output); member regions must
 * not otherwise overlap. Unnormalized, like every plfft transform. Lengths up
 * to 131072 are accepted for any n, and up to 262144 for n whose prime
 * factors are all at most 7. data_type: FP32, FP64, FP16 or BF16 (FP32
 * arithmetic for the 16-bit types), or Q15 / Q7 with power-of-two lengths up
 * to 32768 (one launch per kernel tier, at most five; other lengths fail, as the
 * reference's fixed-point plans do).
 * ---------------------------------------------------------------------- */
typedef struct plfft_mixed_plan_t plfft_mixed_plan_t;

/* Returns PLFFT_NULL_ARGUMENT (null plan or n), PLFFT_INVALID_SIZE (count < 1,
 * a length < 1, a negative offset), PLFFT_INVALID_TRANSFORM_KIND (direction),
 * PLFFT_INVALID_DATA_TYPE, PLFFT_INVALID_PLANNER_RIGOR, or
 * PLFFT_PLAN_CREATION_FAILED (unsupported length, no device; see
 * plfft_last_error()). */
plfft_status_t plfft_mixed_plan_create(plfft_mixed_plan_t **plan, int64_t count,
                                       const int64_t *n, const int64_t *in_offset,
                                       const int64_t *out_offset,
                                       plfft_direction_t direction,
                                       plfft_data_type_t data_type,
                                       plfft_rigor_t rigor);
void plfft_mixed_plan_destroy(plfft_mixed_plan_t *plan);

/* Device bytes of scratch per execution: the intermediate rows of the
 * four-step and chirp-z members (0 when every member fits one pass). */
size_t plfft_mixed_plan_get_workspace_size(const plfft_mixed_plan_t *plan);

/* Complex elements spanned by the input (which = 0) or output (which = 1). */
int64_t plfft_mixed_plan_get_span(const plfft_mixed_plan_t *plan, int which);

/* Synchronous execute: device or host buffers (host buffers are staged).
 * Returns PLFFT_OK or PLFFT_NULL_ARGUMENT. */
plfft_status_t plfft_mixed_plan_execute(const plfft_mixed_plan_t *plan, const void *input,
                                        void *output);

/* Stream-ordered execute with device buffers; `workspace` as for
 * plfft_plan_execute_async (NULL: the plan's own). Allocates nothing when
 * `workspace` is given or the plan needs none: kernel launches only (at
 * most three), no memset, so it can be captured into a hipGraph. Returns PLFFT_OK,
 * PLFFT_NULL_ARGUMENT, or PLFFT_PLAN_CREATION_FAILED if the launch fails. */
plfft_status_t plfft_mixed_plan_execute_async(const plfft_mixed_plan_t *plan,
                                              const void *input, void *output,
                                              void *workspace, void *stream);

/* Schedule summary (members per algorithm, tasks, grid, LDS, workspace). */
size_t plfft_mixed_plan_describe(const plfft_mixed_plan_t *plan, char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* PLFFT_GPU_H_ */
