/*
 * lubatch_util.h -- device plumbing exported by liblubatch.so for callers that
 * have no HIP bindings of their own (the Python ctypes harness, tests, bench).
 * None of these are on the inversion path; they wrap the HIP runtime calls a
 * host driver needs (the reference's main() makes the same calls through the
 * CUDA runtime: cudaMalloc/cudaMemcpy/cudaEvent*, parallel_pivot/
 * luBatchedInplace.cu:112-135).  Status returns use the LUB_* codes.
 */
#ifndef LUBATCH_UTIL_H
#define LUBATCH_UTIL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int   lubx_device_count(void);
/* Properties of device `dev`; any output pointer may be NULL. */
int   lubx_device_info(int dev, char* name, int name_len, char* arch, int arch_len, int* cu_count,
                       int* clock_khz, size_t* total_mem, int* mem_clock_khz, int* mem_bus_width,
                       size_t* lds_per_block);

void* lubx_malloc(size_t bytes);            /* device memory, NULL on failure */
int   lubx_free(void* p);
void* lubx_host_alloc(size_t bytes);        /* pinned host memory, NULL on failure */
int   lubx_host_free(void* p);
int   lubx_memcpy_htod(void* dst, const void* src, size_t bytes);
int   lubx_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int   lubx_memcpy_dtod_async(void* dst, const void* src, size_t bytes, void* stream);
int   lubx_memset(void* p, int value, size_t bytes);
int   lubx_device_sync(void);

void* lubx_stream_create(void);             /* default-priority, non-blocking */
int   lubx_stream_destroy(void* stream);
int   lubx_stream_sync(void* stream);

void* lubx_event_create(void);
int   lubx_event_destroy(void* ev);
int   lubx_event_record(void* ev, void* stream);
int   lubx_event_sync(void* ev);
/* Milliseconds between two recorded events (waits for `stop`); < 0 on error. */
float lubx_event_elapsed_ms(void* start, void* stop);

/* hipGraph capture of work enqueued on `stream` between begin and end; the
 * instantiated graph replays it with one launch. */
int   lubx_graph_capture_begin(void* stream);
int   lubx_graph_capture_end(void* stream, void** graph_exec);
int   lubx_graph_launch(void* graph_exec, void* stream);
int   lubx_graph_destroy(void* graph_exec);

/* Gather / scatter of `count` n x n matrices (nn = n*n floats each) between a
 * packed buffer (matrix i at float offset offsets[i]) and a contiguous one:
 * gather packed -> contiguous, scatter contiguous -> packed.  Benchmark
 * support: the work a caller of the one-n entry point does to invert a
 * class of a mixed-size batch in place.  The contiguous side must be
 * 16-byte aligned. */
int   lubx_gather_f32(float* dst, const float* packed, const int64_t* offsets, int nn, int64_t count, void* stream);
int   lubx_scatter_f32(float* packed, const float* src, const int64_t* offsets, int nn, int64_t count, void* stream);

/* Plain streaming copy kernel (dwordx4, grid-stride), used to calibrate the
 * FETCH_SIZE / WRITE_SIZE counters against a known byte count. */
int   lubx_copy_f32(float* dst, const float* src, size_t count, void* stream);

/* Benchmark driver for lu_server_*: replays a request schedule from a host
 * loop with no interpreter in it.  Request i is due arrival_s[i] seconds
 * after the call (ascending): count[i] matrices packed at address src[i],
 * sizes ns[ns_off[i] ..], flags[i].  latency_ms[i] = arrival to the loop pass
 * that saw it done; stats[0] = batches launched, stats[1] = LUB_EBUSY
 * refusals.  Requests submitted before the call are not released; their
 * completions are taken off the poll queue (lu_server_query still reports
 * them).  LUB_EHIP if the stream has not drained 30 s after its last
 * arrival. */
struct lub_server;
int   lubx_serve_replay(struct lub_server* server, int n_req, const double* arrival_s, const uint64_t* src,
                        const int* ns, const int64_t* ns_off, const int* count, const int* flags,
                        double* latency_ms, int64_t* stats);

#ifdef __cplusplus
}
#endif

#endif /* LUBATCH_UTIL_H */
