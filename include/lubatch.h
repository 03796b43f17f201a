/*
 * lubatch.h -- C ABI of the MI355X (gfx950) batched small-matrix LU inversion
 * library (liblubatch.so).
 *
 * Every entry point takes plain pointers and sizes.  Matrix data is FP32,
 * row-major, contiguous with stride n*n between matrices, and lives in device
 * memory (HBM) owned by the caller.  Inversion is in place.  Functions return 0
 * on success or a LUB_E* status; they never print and never exit the process.
 *
 * Reference interfaces replaced (sumukhashridhar/matrixInversion):
 *   the kernel  template<T,matrixSize,threadsPerMatrix,matricesPerBlock,numMatrices>
 *               __global__ void batched_lu_subwarp(T* A)
 *               (parallel_pivot/luBatchedInplace.cuh:127-129,
 *                serial_pivot/luBatchedInplace.cuh:104-106,
 *                templated/luBatchedInplace.cuh:78-81)
 *   its runtime-argument form
 *               batched_lu_subwarp(T* A, int matrixSize, int numMatrices,
 *                                  int threadsPerMatrix, int matricesPerBlock)
 *               (no_templ/luBatchedInplace.cuh:82-85)
 *   and the launch in each variant's main()
 *               (e.g. parallel_pivot/luBatchedInplace.cu:118-133).
 */
#ifndef LUBATCH_H
#define LUBATCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LUB_OK            0
#define LUB_EINVAL_N      1   /* n outside [1, LUB_MAX_N]                     */
#define LUB_EINVAL_BATCH  2   /* batch < 0                                    */
#define LUB_ENULL         3   /* required pointer is NULL                     */
#define LUB_EFLAGS        4   /* unknown flag bits or pivot rule              */
#define LUB_EHIP          5   /* a HIP runtime call failed                    */
#define LUB_ENODEVICE     6   /* no gfx950 device / library not initialised   */
#define LUB_EALIGN        7   /* pointer not 4-byte aligned                   */
#define LUB_EINVAL_ARG    8   /* other invalid argument (ragged tables, ...)  */
#define LUB_EBUSY         9   /* server: no room under the budget right now;
                                 retry after lu_server_progress / _release    */
#define LUB_ETOOBIG      10   /* batch above a fixed capacity
                                 (lu_batched_inverse_ragged_noalloc)          */

#define LUB_MAX_N 32

/* ---- flags ------------------------------------------------------------- */
/* Pivot rule, bits 0..3.  The REF rules reproduce the reference: at step k the
 * pivot is searched in column k of the row-permuted *input* matrix, which the
 * Doolittle factorisation has not updated yet, so the pivot vector is a pure
 * function of the input values. */
#define LUB_PIVOT_NONE      0x0  /* no pivoting (templated/, no_templ/)                  */
#define LUB_PIVOT_PARALLEL  0x1  /* REF rule, first maximum |a| wins ties, cross-lane
                                    DPP argmax (parallel_pivot/ technique, serial_pivot/
                                    tie rule); default                                   */
#define LUB_PIVOT_SERIAL    0x2  /* REF rule, first-max (serial_pivot/'s rule): the same
                                    pivots as LUB_PIVOT_PARALLEL and the same kernel,
                                    whose exact fallback is serial_pivot's lane-serial
                                    scan                                                 */
#define LUB_PIVOT_TREE      0x3  /* REF rule with parallel_pivot/'s exact shared-memory
                                    tree (luBatchedInplace.cuh:12-44): left slot wins
                                    ties and, for n not a power of two, some lanes are
                                    never reduced -- bit-compatible with that kernel     */
#define LUB_PIVOT_GEPP      0x4  /* LAPACK getrf's rule: largest |a| of the *updated* column,
                                    first position on ties (no counterpart in the
                                    reference, whose rules search the un-updated column);
                                    pivots depend on the arithmetic, ipiv is getrf's    */
#define LUB_PIVOT_MASK      0xF

/* Lane layout, bits 4..5 (a tuning override; AUTO picks per n, from the
 * autotuned table when one is loaded).  Results do not depend on it. */
#define LUB_LAYOUT_AUTO     0x00
#define LUB_LAYOUT_ROW16    0x10 /* one matrix per 16-lane DPP row, 4 per wave (n <= 32) */
#define LUB_LAYOUT_QUAD4    0x20 /* one matrix per 4-lane quad, 16 per wave (n <= 16)   */
#define LUB_LAYOUT_MASK     0x30

/* ---- entry points ------------------------------------------------------ */

/* Invert `batch` n x n FP32 matrices at device pointer A in place.
 * piv: device int32[batch*n] receiving perm[p] = original row at position p
 *      (the reference's `pivots` array, never exported by the reference itself;
 *      P A = L U with (P A)[p][:] = A[perm[p]][:]), or NULL.  With
 *      LUB_PIVOT_NONE a non-NULL piv receives the identity.
 * Synchronous: runs on the calling thread's default stream and waits. */
int lu_batched_inverse(float* A, int* piv, int n, int batch, int flags);

/* Same, enqueued on `stream` (a hipStream_t; NULL = default stream); returns
 * after the launch without waiting.  A and piv must stay valid until the
 * stream reaches the kernel. */
int lu_batched_inverse_async(float* A, int* piv, int n, int batch, int flags, void* stream);

/* 64-bit batch form of lu_batched_inverse_async. */
int lu_batched_inverse_async64(float* A, int* piv, int n, int64_t batch, int flags, void* stream);

/* ---- ragged (mixed-n) batches -----------------------------------------
 * Matrix i has size ns[i] (1..32) and starts at float offset offsets[i] of
 * A (row-major, contiguous, inverted in place); its pivot vector goes to
 * piv + piv_offsets[i] (piv may be NULL).  All five arrays are device
 * memory.  The batch is planned and inverted on the device by one
 * persistent kernel (size classes largest first, per-class lane layout and
 * matrices per work item from the tuning table); results equal a plain launch
 * of each matrix's n.  Matrices with n outside 1..32 are skipped and counted
 * (lu_ragged_invalid_count).  Replaces, for mixed sizes, one reference build
 * and launch per MATRIXSIZE (parallel_pivot/run.py:198-234). */
typedef struct lub_ragged_plan lub_ragged_plan;

/* Workspace for batches of up to max_batch matrices (device memory of about
 * 16.1 bytes per matrix: two int64 slots, and 33 int32 per chunk of 1024);
 * a plan may be used by one stream at a time. */
int lu_ragged_plan_create(lub_ragged_plan** plan, int64_t max_batch);
int lu_ragged_plan_destroy(lub_ragged_plan* plan);

/* Enqueue the planning kernels and the persistent inversion kernel on
 * `stream`.  Allocates nothing and never synchronises, so the call can be
 * captured into a hipGraph and replayed. */
int lu_batched_inverse_ragged_plan(lub_ragged_plan* plan, float* A, const int64_t* offsets, const int* ns, int* piv,
                                   const int64_t* piv_offsets, int64_t batch, int flags, void* stream);

/* Convenience form with a library-owned plan per host thread (created or
 * grown on demand, which synchronises the device once).  Calls from one
 * thread share that plan: issue them on one stream. */
int lu_batched_inverse_ragged(float* A, const int64_t* offsets, const int* ns, int* piv, const int64_t* piv_offsets,
                              int batch, int flags, void* stream);

/* The same ragged inversion with no allocation at all, for callers that
 * cannot call hipMalloc (code under stream capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  The
 * planning workspace is a fixed device array inside the library, one per GPU,
 * sized for LUB_NOALLOC_MAX_BATCH matrices; a larger batch returns
 * LUB_ETOOBIG before any device work.  Every call on one GPU shares that
 * GPU's array, so calls that overlap on one GPU (other streams, other host
 * threads) race: serialise them, e.g. on one stream.  The first call on a
 * device looks up the array's address (no allocation); afterwards the call
 * only enqueues kernels and never synchronises. */
#define LUB_NOALLOC_MAX_BATCH 1048576
int lu_batched_inverse_ragged_noalloc(float* A, const int64_t* offsets, const int* ns, int* piv,
                                      const int64_t* piv_offsets, int64_t batch, int flags, void* stream);

/* Matrices of the last enqueued batch whose n was outside 1..32 (waits for it). */
int lu_ragged_invalid_count(const lub_ragged_plan* plan, int* out);

/* ---- tuning ------------------------------------------------------------
 * Per-n launch choices used when flags carry LUB_LAYOUT_AUTO: layout[n]
 * (0 = 16-lane rows, 1 = quads; quads apply to n <= 16 in the plain kernel,
 * n <= 12 in the ragged one) and grab[n] (ragged: wave tiles per work item,
 * >= 1, used up to 16 matrices per item).  Arrays of 33 entries indexed by n
 * (entry 0 unused). */
int lu_tuning_set(const signed char* layout, const signed char* grab);
int lu_tuning_get(signed char* layout, signed char* grab);

/* LDS row padding of the plain kernels, per n (33 entries indexed by n):
 * pad[n] = 1 pads each row of a staged n x n matrix by 4 floats (16 bytes),
 * 0 leaves rows unpadded; only n % 4 == 0 may be padded (LUB_EINVAL_ARG
 * otherwise).  Default: padded where n % 4 == 0, except n = 28.  Results
 * do not depend on it.  Replaces the reference's bank swizzle of the
 * shared-memory rows (swzl/luBatchedInplace.cuh:12-20), a compile-time
 * choice there. */
int lu_tuning_set_lds_pad(const signed char* pad);
int lu_tuning_get_lds_pad(signed char* pad);

/* ---- LAPACK-style outputs -----------------------------------------------
 * The same inversion with getrf/getri-shaped outputs: ipiv[k] (1-based) is
 * the row swapped with row k at step k, info is 0, or k + 1 for the first
 * pivot U(k,k) of P*A = LU that is exactly zero (as LAPACK) or has no finite
 * reciprocal (inf, NaN, |U(k,k)| below ~3e-39; LAPACK would report 0 there);
 * the inverse then holds inf/NaN.  The
 * interchanges are those of the chosen pivot rule (the reference's first-max
 * on the un-updated column: parallel_pivot/luBatchedInplace.cuh:12-44), so
 * they may differ from LAPACK partial pivoting; applied in order they give
 * this factorisation's P.  ipiv: batch x n ints, info: batch ints (device). */
int lu_batched_inverse_lapack(float* A, int* ipiv, int* info, int n, int64_t batch, int flags, void* stream);

/* In-place conversion of permutation pivot vectors (the `piv` output of
 * lu_batched_inverse*) to 1-based interchange sequences. */
int lu_perm_to_ipiv(int* piv, int n, int64_t batch, void* stream);

/* ---- refinement ---------------------------------------------------------
 * One step of iterative refinement, X <- X + X (I - A X), for `batch` n x n
 * FP32 matrices (A and X contiguous, row-major): the residual I - A X is
 * accumulated in FP64, the correction X (I - A X) comes from FP32 matrix-core
 * (MFMA) products.  The reference has no refinement step: its inverse is the
 * FP32 LU result (parallel_pivot/luBatchedInplace.cuh:84-125). */
int lu_batched_refine(const float* A, float* X, int n, int64_t batch, void* stream);

/* Out-of-place inverse plus one refinement step: X <- A^-1 (the plain path
 * with `flags`, pivots to piv as in lu_batched_inverse), then
 * lu_batched_refine(A, X).  A is left unchanged (A != X). */
int lu_batched_inverse_refined(const float* A, float* X, int* piv, int n, int64_t batch, int flags, void* stream);

/* ---- serving ------------------------------------------------------------
 * A continuous-batching server over the ragged path.  The reference has no
 * serving layer (one batch per process: parallel_pivot/run.py:198-234).
 *
 * Requests are admitted into a device arena of budget_bytes (the HBM
 * admission budget) and stay resident until released.  Queued requests of
 * one pivot rule are coalesced into one ragged batch whenever one of the
 * server's launch slots (three, each a stream and a ragged plan) is free, so
 * up to three batches -- of one rule or of several -- run at once, and a
 * request that arrives while a batch runs starts without waiting for it.
 * Everything the server uses is allocated by lu_server_create.
 * A result is bit-identical to inverting the same matrices alone.
 * One host thread drives a server; no call blocks except destroy. */
typedef struct lub_server lub_server;

/* max_batch: most matrices in one launched batch (and in one request);
 * ring_slots (>= max_batch): queued + running matrices per pivot rule. */
int lu_server_create(lub_server** server, int64_t budget_bytes, int64_t max_batch, int64_t ring_slots);
int lu_server_destroy(lub_server* server);

/* Admit `count` matrices of sizes ns[0..count) (host array), packed
 * row-major back to back at src (device or host memory), to be inverted
 * with `flags` (pivot rule | layout).  The data is copied in asynchronously;
 * src may be reused once the request is done.  LUB_EBUSY: not now. */
int lu_server_submit(lub_server* server, const float* src, const int* ns, int count, int flags, uint64_t* ticket);

/* Retire finished batches, launch queued work on idle streams; counts of
 * launched batches and completed requests (either pointer may be NULL). */
int lu_server_progress(lub_server* server, int* launched, int* completed);

/* state: 0 queued, 1 running, 2 done.  result / piv: device pointers to the
 * request's inverses (same packing as the input) and pivot vectors (n ints
 * per matrix; NULL for LUB_PIVOT_NONE); valid until release.  latency_ms:
 * submit to the progress call that saw completion (-1 before done). */
int lu_server_query(lub_server* server, uint64_t ticket, int* state, float** result, int** piv, double* latency_ms);

/* Tickets that finished since the last poll, in completion order: up to
 * `capacity` of them go to tickets[], their number to *count (the rest wait
 * for the next poll). */
int lu_server_poll(lub_server* server, uint64_t* tickets, int capacity, int* count);

/* Give a done request's memory back. */
int lu_server_release(lub_server* server, uint64_t ticket);

/* Matrices queued and running; arena bytes held (admitted, not reclaimed). */
int lu_server_stats(lub_server* server, int64_t* queued, int64_t* running, int64_t* resident_bytes);

/* Batches launched since create, and the most batches of one pivot rule that
 * were running at the same time (either pointer may be NULL). */
int lu_server_counters(lub_server* server, int64_t* batches, int* peak_rule_inflight);

/* Device-side inversion check, the reference's verifyInv (parallel_pivot/
 * verify.hpp:50-103) on the GPU.  For each of `batch` matrices, A (the
 * original matrices) times X (their computed inverses) is formed in FP32 in
 * the reference's summation order; the matrix counts as correct when every
 * entry is within `thr` of the identity (the reference uses 1e-3).
 *   *d_correct  (device uint64)  += number of correct matrices
 *   *d_max_bits (device uint32)  = max(*d_max_bits, bits of max |A X - I|)
 *                                  (non-negative FP32 bit pattern; NaN -> 0x7f800001)
 * The caller zeroes both before the first call.  Enqueued on `stream`. */
int lu_batched_residual(const float* A, const float* X, int n, int64_t batch, float thr,
                        unsigned long long* d_correct, unsigned int* d_max_bits, void* stream);

/* Human-readable text of a status code (static storage). */
const char* lu_status_string(int status);

/* Text of the last HIP runtime failure seen by the calling thread ("" if none). */
const char* lu_last_error(void);

/* Library version, e.g. "0.2.0". */
const char* lu_version(void);

/* Hash of the sources the library was built from (16 hex digits); counter
 * profiles record it, so that a profile of other code can be told apart. */
const char* lu_build_id(void);

#ifdef __cplusplus
}
#endif

#endif /* LUBATCH_H */
