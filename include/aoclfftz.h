/* aoclfftz.h -- plan-create / execute / destroy interface of the MI355X FFT
 * library.  Drop-in replacement for the reference's api/aoclfftz.h: every
 * struct below has the same field order, types and layout, every entry point
 * the same name, signature and error behaviour, so a program built against
 * the reference links against this library unchanged.
 *
 * Differences a caller may observe:
 *   - in / out may be device pointers (hipMalloc / torch) or host pointers;
 *     host buffers are staged through device memory inside the call;
 *   - execute and execute_io return with results ready (synchronous), as in
 *     the reference, unless a stream has been bound with
 *     aoclfftz_gpu_set_stream (aoclfftz_gpu.h);
 *   - pthr_fft (OpenMP threads) and cntrl_params.opt_level select CPU code
 *     paths in the reference and are accepted and ignored here.
 */
#ifndef AOCLFFTZ_H
#define AOCLFFTZ_H

#include <stddef.h>

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AOCLFFTZ_LIBRARY_VERSION "AOCL-FFTZ 6.0.0"

/* Error codes; values fixed by the reference (api/aoclfftz.h:60-68). */
typedef enum
{
    AOCLFFTZ_TIME_OUT = -5,
    AOCLFFTZ_MEMORY_FAILURE,
    AOCLFFTZ_INVALID_INPUT,
    AOCLFFTZ_SETUP_FAILURE,
    AOCLFFTZ_EXECUTION_FAILURE,
    AOCLFFTZ_SUCCESS
} aoclfftz_error_type;

/* Logging verbosity (api/aoclfftz.h:74-80). */
typedef enum
{
    AOCLFFTZ_LOG_NONE = 0,
    AOCLFFTZ_LOG_INFO = 1,
    AOCLFFTZ_LOG_TRACE = 2,
    AOCLFFTZ_LOG_DEBUG = 3,
} aoclfftz_logger_mode;

/* Transform flags (api/aoclfftz.h:86-102).
 *   fft_type            0 complex, 1 real
 *   fft_direction       0 forward (exp(-2 pi i jk/N)), 1 backward (+i),
 *                       both unnormalised
 *   storage_order       must be 0 (in-order)
 *   fft_placement       0 in-place (in == out), 1 out-of-place
 *   transpose_mode      must be 0
 *   bit_reproducibility 0 / 1; GPU results are bitwise reproducible run to
 *                       run in either mode (no atomics in the arithmetic) */
typedef struct aoclfftz_flags
{
    FFTZ_UINT8 fft_type;
    FFTZ_UINT8 fft_direction;
    FFTZ_UINT8 storage_order;
    FFTZ_UINT8 fft_placement;
    FFTZ_UINT8 transpose_mode;
    FFTZ_UINT8 bit_reproducibility;
} aoclfftz_flags_t;

/* One tensor axis: length and strides in elements (complex elements for
 * complex data, real elements on the real side of R2C / C2R).  dims[0] is the
 * innermost axis. (api/aoclfftz.h:108-125) */
typedef struct aoclfftz_dim
{
    FFTZ_INT32 n;
    FFTZ_INT32 in_stride;
    FFTZ_INT32 out_stride;
} aoclfftz_dim_t;

typedef struct aoclfftz_dim_64_
{
    FFTZ_INTP n;
    FFTZ_INTP in_stride;
    FFTZ_INTP out_stride;
} aoclfftz_dim_t_64_;

/* CPU thread parameters (api/aoclfftz.h:130-140); accepted, unused on GPU. */
typedef struct aoclfftz_smp_pfft
{
    FFTZ_INT32 num_threads;
    FFTZ_UINT32 dynamic_load_model;
} aoclfftz_smp_pfft_t;

/* Control parameters (api/aoclfftz.h:146-165). */
typedef struct aoclfftz_cntrl_params
{
    FFTZ_INT32 opt_level;
    FFTZ_INT32 opt_off;
    aoclfftz_logger_mode logger_mode;
    FFTZ_INT32 measure_stats;
} aoclfftz_cntrl_params_t;

/* Problem descriptors (api/aoclfftz.h:174-311).  The library copies what it
 * needs at setup; the descriptor and its dims/vecs arrays may be freed after
 * the setup call returns. */
typedef struct
{
    FFTZ_FLOAT *in;
    FFTZ_FLOAT *out;
    FFTZ_INT32 vec_rank;
    FFTZ_INT32 dim_rank;
    aoclfftz_dim_t *dims;
    aoclfftz_dim_t *vecs;
    aoclfftz_flags_t flags;
    aoclfftz_smp_pfft_t pthr_fft;
    aoclfftz_cntrl_params_t cntrl_params;
} aoclfftz_prob_desc_f;

typedef struct
{
    FFTZ_DOUBLE *in;
    FFTZ_DOUBLE *out;
    FFTZ_INT32 vec_rank;
    FFTZ_INT32 dim_rank;
    aoclfftz_dim_t *dims;
    aoclfftz_dim_t *vecs;
    aoclfftz_flags_t flags;
    aoclfftz_smp_pfft_t pthr_fft;
    aoclfftz_cntrl_params_t cntrl_params;
} aoclfftz_prob_desc_d;

typedef struct
{
    FFTZ_FLOAT *in;
    FFTZ_FLOAT *out;
    FFTZ_INT32 vec_rank;
    FFTZ_INT32 dim_rank;
    aoclfftz_dim_t_64_ *dims;
    aoclfftz_dim_t_64_ *vecs;
    aoclfftz_flags_t flags;
    aoclfftz_smp_pfft_t pthr_fft;
    aoclfftz_cntrl_params_t cntrl_params;
} aoclfftz_prob_desc_f_64_;

typedef struct
{
    FFTZ_DOUBLE *in;
    FFTZ_DOUBLE *out;
    FFTZ_INT32 vec_rank;
    FFTZ_INT32 dim_rank;
    aoclfftz_dim_t_64_ *dims;
    aoclfftz_dim_t_64_ *vecs;
    aoclfftz_flags_t flags;
    aoclfftz_smp_pfft_t pthr_fft;
    aoclfftz_cntrl_params_t cntrl_params;
} aoclfftz_prob_desc_d_64_;

/* Plan creation.  Validates the descriptor exactly as the reference does
 * (api/validate_problem.h:346-417), builds the GPU plan (pass list, twiddle
 * and chirp tables in device memory) and returns an opaque handle, or NULL
 * with an "[ERROR]" line on stderr.
 * Replace api/aoclfftz.h:343 (setup_f), :371 (setup_d), :399 (setup_f_64_),
 * :428 (setup_d_64_). */
FFTZ_VOID *aoclfftz_setup_f(aoclfftz_prob_desc_f *problem);
FFTZ_VOID *aoclfftz_setup_d(aoclfftz_prob_desc_d *problem);
FFTZ_VOID *aoclfftz_setup_f_64_(aoclfftz_prob_desc_f_64_ *problem);
FFTZ_VOID *aoclfftz_setup_d_64_(aoclfftz_prob_desc_d_64_ *problem);

/* Execute on the setup-time buffers.  AOCLFFTZ_EXECUTION_FAILURE for a NULL
 * handle or when setup left in/out NULL.  Not thread-safe on one handle.
 * Replaces api/aoclfftz.h:452. */
aoclfftz_error_type aoclfftz_execute(FFTZ_VOID *handle);

/* Execute on caller buffers; safe to call concurrently on one handle (each
 * call gets its own stream and device scratch).  Replaces
 * api/aoclfftz.h:484. */
aoclfftz_error_type aoclfftz_execute_io(FFTZ_VOID *handle, FFTZ_VOID *in,
                                        FFTZ_VOID *out);

/* Free every device and host resource of the handle.  Replaces
 * api/aoclfftz.h:499. */
FFTZ_VOID aoclfftz_destroy(FFTZ_VOID *handle);

/* Version string.  Replaces api/aoclfftz.h:507. */
const FFTZ_CHAR *aoclfftz_version(FFTZ_VOID);

#ifdef __cplusplus
}
#endif

#endif /* AOCLFFTZ_H */
