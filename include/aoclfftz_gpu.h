/* aoclfftz_gpu.h -- GPU-only extensions of the MI355X FFT library.
 *
 * None of these has a counterpart in the reference (a CPU library has no
 * streams or device buffers).  They are for callers that keep data resident
 * in HBM and want asynchronous execution; a program that only uses
 * aoclfftz.h never needs them.
 */
#ifndef AOCLFFTZ_GPU_H
#define AOCLFFTZ_GPU_H

#include "aoclfftz.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bind a HIP stream (hipStream_t passed as void*) to the handle.  Afterwards
 * aoclfftz_execute / aoclfftz_execute_io enqueue on that stream and return
 * without waiting.  In this mode in / out MUST be device memory (hipMalloc,
 * managed, or torch device tensors): the call does not inspect the pointers.
 * Passing NULL restores the default synchronous behaviour, where host
 * buffers are also accepted and staged.  Calls that share
 * one handle in stream mode are ordered by the stream; do not use one handle
 * from several streams at once. */
aoclfftz_error_type aoclfftz_gpu_set_stream(FFTZ_VOID *handle, FFTZ_VOID *stream);

/* Execute the plan on count (in, out) buffer pairs in one call, in order, with
 * the semantics of aoclfftz_execute_io for each pair (asynchronous in stream
 * mode).  Saves the per-call overhead of foreign-function callers. */
aoclfftz_error_type aoclfftz_gpu_execute_many(FFTZ_VOID *handle, FFTZ_VOID **ins, FFTZ_VOID **outs,
                                             FFTZ_INT32 count);

/* Execute the plan on (in, out) without allocating anything: no hipMalloc,
 * hipFree, stream, event or graph is created by the call.  No counterpart in
 * the reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).
 *
 * in / out must be device memory.  Results are those of aoclfftz_execute_io
 * on the same buffers (the same kernels run).  With a stream bound
 * (aoclfftz_gpu_set_stream) the kernels are enqueued there and the call
 * returns without waiting; otherwise they run on the null stream and the
 * call waits for them.  Plans that need work memory (four-step and
 * multi-pass transforms, uneven batches) use a static per-GPU scratch array
 * of 64 MiB of work memory plus 256 bytes of queue counters, held by every
 * GPU the library runs on until the process ends; a plan needing more
 * returns AOCLFFTZ_INVALID_INPUT with a message and launches nothing.  The
 * array is looked up on the first call on each device (with that device
 * current) and must be the plan's device.  All calls on one GPU share its
 * array and it keeps the last call's contents: calls that overlap on one GPU
 * (from several threads, or on different streams) race and must be ordered
 * by the caller. */
aoclfftz_error_type aoclfftz_gpu_execute_noalloc(FFTZ_VOID *handle, FFTZ_VOID *in, FFTZ_VOID *out);

/* Wait for all work the handle has enqueued. */
aoclfftz_error_type aoclfftz_gpu_synchronize(FFTZ_VOID *handle);

/* Capture the handle's launches into one hipGraph per (in, out) pair and
 * replay it (enable = 1), or launch kernels directly (enable = 0). */
aoclfftz_error_type aoclfftz_gpu_set_graph(FFTZ_VOID *handle, FFTZ_INT32 enable);

/* Operand precision of the radix-16 stages of complex64 passes of length 256
 * or 4096 (the all-16 radix plan; the batched N = 4096 and N = 256 cases):
 *   0  plain FP32 Stockham codelets on the vector ALUs (default);
 *   1  each radix-16 stage as a DFT-matrix GEMM on the matrix cores with
 *      FP32 operands;
 *   2  the same GEMM with operands split into bf16 hi + lo, three products
 *      (hi*hi + hi*lo + lo*hi), FP32 accumulation;
 *   3  operands scaled per column by a power of two and split into fp16
 *      hi + lo, three products, FP32 accumulation.
 * Re-plans the handle.  AOCLFFTZ_INVALID_INPUT (plan unchanged) when no pass
 * of the plan qualifies.  Not to be called while the handle executes. */
aoclfftz_error_type aoclfftz_gpu_set_compute_mode(FFTZ_VOID *handle, FFTZ_INT32 mode);

/* Autotuner.  Times candidate plans of the handle's problem on the given
 * device buffers (reps executes each; in / out are overwritten; an in-place
 * handle takes in == out) and keeps the fastest.  For an ordinary plan the
 * search runs in steps, each starting from the best so far: row-tile width x
 * column-tile lines; the four-step split of a long 1-D axis (2 or 3 factors);
 * the radix order of each pass length (other multisets of the supported
 * radices, descending and ascending); complex values per thread (8 or 16).
 * For an uneven batch it times workgroups per CU of each tile class.  The
 * choice is also stored in a process-wide cache keyed by the problem, which
 * later setups of the same problem use without tuning again.  No counterpart
 * in the reference, whose plan choice is a cost model over the same kinds of
 * decision (selector/selector.c; radix plan core/common/pow2_radix_decompose.c:122-198,
 * radix search selector/complex/selector_ct_dft.c:104-328). */
aoclfftz_error_type aoclfftz_gpu_autotune(FFTZ_VOID *handle, FFTZ_VOID *in, FFTZ_VOID *out, FFTZ_INT32 reps);

/* Write / read the tuned cache as text (one problem per line:
 * "key row_t col_c wg_small wg_medium [e=E] [r=L:r.r.r/L:r.r] [f=f.f.f]",
 * the optional fields being complex values per thread, radix orders per pass
 * length and the four-step split; '#' starts a comment line).  Return
 * the number of entries, or -1 when the file cannot be opened or a line is
 * malformed (then nothing is imported).  The counterpart of FFTW wisdom files
 * for this library; the reference's wisdom calls are stubs
 * (wrappers/fftw/src/c/fftw_misc_wrapper.c:300-367), this library's FFTW
 * wrapper routes them here. */
FFTZ_INT32 aoclfftz_gpu_tune_export(const FFTZ_CHAR *path);
FFTZ_INT32 aoclfftz_gpu_tune_import(const FFTZ_CHAR *path);

/* The same text in memory.  _export_string returns a malloc'd,
 * NUL-terminated string (release with free()), NULL only when out of memory;
 * _import_string returns entries stored or -1 as above; _clear empties the
 * cache (later setups use the planner's defaults).  No counterparts in the
 * reference; they back fftw_export_wisdom_to_string / import_wisdom_from_string
 * / forget_wisdom and the callback forms. */
FFTZ_CHAR *aoclfftz_gpu_tune_export_string(FFTZ_VOID);
FFTZ_INT32 aoclfftz_gpu_tune_import_string(const FFTZ_CHAR *text);
FFTZ_VOID aoclfftz_gpu_tune_clear(FFTZ_VOID);

/* Human-readable plan summary (pass list with lengths, radices, tiles).  The
 * string lives as long as the handle. */
const FFTZ_CHAR *aoclfftz_gpu_plan_info(FFTZ_VOID *handle);

/* Algorithmic flop count of one execute: 5 N log2 N per complex transform,
 * 2.5 N log2 N per real transform, times the batch. */
FFTZ_DOUBLE aoclfftz_gpu_plan_flops(FFTZ_VOID *handle);

/* Device bytes held by the plan (tables) and per execution slot (work
 * buffer). */
FFTZ_INT64 aoclfftz_gpu_plan_device_bytes(FFTZ_VOID *handle);

/* Per-launch timing.  While enabled, every execute records a HIP event
 * before its first kernel and after each kernel, on the stream the kernels
 * run on (direct launches; the graph path is bypassed).  profile_read waits
 * for those events and writes the kernel durations in milliseconds, in
 * launch order, execute after execute; it returns how many were recorded
 * (which may exceed cap) and recycles the events. */
aoclfftz_error_type aoclfftz_gpu_profile(FFTZ_VOID *handle, FFTZ_INT32 enable);
FFTZ_INT32 aoclfftz_gpu_profile_read(FFTZ_VOID *handle, float *ms, FFTZ_INT32 cap);

/* Number of kernel launches per execute. */
FFTZ_INT32 aoclfftz_gpu_plan_launches(FFTZ_VOID *handle);

/* Diagnostic: validate a descriptor (aoclfftz_prob_desc_f/_d, or the _64_
 * variants when ilp64 != 0; the buffer type does not matter), plan it
 * without touching the device and write a JSON description of every launch
 * (lengths, strides, line geometry, flags) into buf.  Returns the full JSON
 * length, or -1 on an invalid / unsupported problem.  Needs no GPU. */
FFTZ_INT64 aoclfftz_gpu_plan_json(FFTZ_VOID *problem, FFTZ_INT32 precision_bytes, FFTZ_INT32 ilp64,
                                  FFTZ_CHAR *buf, FFTZ_INT64 buflen);

/* Uneven batch: many 1-D C2C transforms of different lengths in one
 * persistent launch per tile class, fed by a device work queue.  Item i is
 * `batch` transforms of length n, `dist` elements apart (0 means n), starting
 * at element offsets in_offset / out_offset of the buffers later passed to
 * aoclfftz_execute_io (complex elements; precision_bytes 4 or 8;
 * direction 0 forward, 1 backward, unnormalised).  Lengths with a prime
 * factor above 13 run Bluestein inside the same launch.  The handle works
 * with aoclfftz_execute_io, aoclfftz_destroy and the stream / graph /
 * profiling calls above.  No counterpart in the reference, which plans and
 * executes each (n, batch) separately. */
typedef struct
{
    FFTZ_INT64 n;
    FFTZ_INT64 batch;
    FFTZ_INT64 in_offset;
    FFTZ_INT64 out_offset;
    FFTZ_INT64 dist;
} aoclfftz_gpu_batch_item;

FFTZ_VOID *aoclfftz_gpu_batch_setup(const aoclfftz_gpu_batch_item *items, FFTZ_INT64 count,
                                    FFTZ_INT32 precision_bytes, FFTZ_INT32 direction);

/* Dry-run of aoclfftz_gpu_batch_setup (no GPU needed): JSON with the size
 * classes and every tile of the work queue.  Returns the JSON length or -1. */
FFTZ_INT64 aoclfftz_gpu_batch_plan_json(const aoclfftz_gpu_batch_item *items, FFTZ_INT64 count,
                                        FFTZ_INT32 precision_bytes, FFTZ_INT32 direction, FFTZ_CHAR *buf,
                                        FFTZ_INT64 buflen);

/* Serving runtime: independent transform requests sharing one GPU.
 *
 * A server keeps one plan (and its tables) resident per distinct problem and
 * runs requests on up to max_inflight execution slots at once.  A request is
 * admitted when a slot is free and its HBM footprint -- input and output
 * extents plus the plan's work buffer, in bytes -- fits hbm_budget_bytes
 * next to the requests already running; admission is first come, first
 * served, and requests that arrive while others run are admitted as slots
 * and memory free up.  A progress thread owned by the server does the
 * admitting and completion tracking; it sleeps while the server is idle and
 * polls (yielding the core) while requests are in flight.  Plans stay
 * resident until the server is destroyed (no eviction).  No counterpart in the reference (a
 * CPU library executes one call at a time on the caller's thread). */

/* NULL on bad arguments or when no device is visible. */
FFTZ_VOID *aoclfftz_gpu_server_create(FFTZ_INT64 hbm_budget_bytes, FFTZ_INT32 max_inflight);

/* Queue one request.  problem: an ILP64 descriptor (aoclfftz_prob_desc_f_64_
 * or aoclfftz_prob_desc_d_64_, per precision_bytes 4 / 8) whose in / out are
 * device buffers that stay valid until the request completes.  Returns the
 * request id (>= 0), -1 for an invalid problem, -2 when the footprint alone
 * exceeds the budget.  Thread-safe; returns without waiting. */
FFTZ_INT64 aoclfftz_gpu_server_submit(FFTZ_VOID *server, FFTZ_VOID *problem, FFTZ_INT32 precision_bytes);

/* Open-loop load generator: submits problems[i] (as aoclfftz_gpu_server_submit,
 * precision_bytes[i] 4 or 8) at arrival_s[i] seconds after the call starts,
 * spinning on the host clock in between, and writes the request ids (or the
 * submit error codes) to ids[i].  Each request's latency is measured from its
 * scheduled arrival, so a generator that falls behind shows up as latency,
 * not as a lower offered rate; *max_late_s (optional) receives the largest
 * delay of a submission behind its schedule.  Returns count, or -1 on bad
 * arguments (arrival times must not decrease).  Blocks the calling thread
 * until the last arrival.  No counterpart in the reference. */
FFTZ_INT64 aoclfftz_gpu_server_submit_timed(FFTZ_VOID *server, FFTZ_VOID **problems,
                                           const FFTZ_INT32 *precision_bytes, const FFTZ_DOUBLE *arrival_s,
                                           FFTZ_INT64 count, FFTZ_INT64 *ids, FFTZ_DOUBLE *max_late_s);

/* 1 when request id has completed (or failed), 0 while queued or running,
 * -1 for an unknown id. */
FFTZ_INT32 aoclfftz_gpu_server_test(FFTZ_VOID *server, FFTZ_INT64 id);

/* Block until request id completes; writes its latency (submit to
 * completion, seconds) and forgets the id.  0 on success, -1 on failure or
 * an unknown id. */
FFTZ_INT32 aoclfftz_gpu_server_wait(FFTZ_VOID *server, FFTZ_INT64 id, FFTZ_DOUBLE *latency_s);

/* Statistics since creation or the last reset, written to out[0..cap):
 * completed, failed, p50 latency (ms), p99 latency (ms), max latency (ms),
 * transforms/s, GFLOP/s (5 N log2 N), peak admitted footprint (bytes),
 * plans resident, peak queue length, span (s, first submit to last
 * completion).  Returns the number of values (11). */
FFTZ_INT32 aoclfftz_gpu_server_stats(FFTZ_VOID *server, FFTZ_DOUBLE *out, FFTZ_INT32 cap);
FFTZ_VOID aoclfftz_gpu_server_reset_stats(FFTZ_VOID *server);

/* Drains queued and running requests, then frees plans and the server. */
FFTZ_VOID aoclfftz_gpu_server_destroy(FFTZ_VOID *server);

#ifdef __cplusplus
}
#endif

#endif /* AOCLFFTZ_GPU_H */
