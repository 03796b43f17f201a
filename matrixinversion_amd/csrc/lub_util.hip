// lub_util.hip -- device plumbing for the ctypes harness (include/lubatch_util.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "lubatch.h"
#include "lubatch_util.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));  // 4-byte aligned: packed offsets

inline int rc(hipError_t e) { return e == hipSuccess ? LUB_OK : LUB_EHIP; }

void copy_n(char* dst, int len, const char* src) {
  if (dst == nullptr || len <= 0) return;
  std::strncpy(dst, src, static_cast<size_t>(len - 1));
  dst[len - 1] = '\0';
}

__global__ __launch_bounds__(256) void copy_kernel(f32x4* __restrict__ dst, const f32x4* __restrict__ src,
                                                   size_t n4) {
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n4; i += size_t(gridDim.x) * 256)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ __launch_bounds__(256) void copy_scalar_kernel(float* __restrict__ dst, const float* __restrict__ src,
                                                          size_t begin, size_t count) {
  for (size_t i = begin + blockIdx.x * size_t(256) + threadIdx.x; i < count; i += size_t(gridDim.x) * 256)
    dst[i] = src[i];
}

// Element f of the contiguous side is element f % nn of matrix f / nn, at
// float offset offsets[f / nn] of the packed side.  Each thread moves four
// consecutive contiguous-side floats (one dwordx4 there; the packed side is
// only 4-byte aligned, so element-wise) with one division per four elements;
// I is uint32_t when every index fits, else int64_t.
template <bool TO_CONTIGUOUS, typename I>
__global__ __launch_bounds__(256) void gather_kernel(float* __restrict__ dst, const float* __restrict__ src,
                                                     const int64_t* __restrict__ offsets, int nn, int64_t total) {
  const I quads = static_cast<I>((total + 3) / 4);
  const I n = static_cast<I>(nn);
  for (I q = blockIdx.x * I(blockDim.x) + threadIdx.x; q < quads; q += I(gridDim.x) * blockDim.x) {
    const I f0 = 4 * q;
    I i = f0 / n;
    I e = f0 - i * n;
    int64_t base = offsets[i];
    const bool whole = int64_t(f0) + 4 <= total;
    if (TO_CONTIGUOUS) {
      f32x4 v;
      if (whole && e + 4 <= n) {  // inside one matrix: one dword-aligned dwordx4
        v = *reinterpret_cast<const f32x4u*>(src + base + e);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (e == n) { ++i; e = 0; base = (int64_t(f0) + k < total) ? offsets[i] : 0; }
          v[k] = (int64_t(f0) + k < total) ? src[base + e] : 0.0f;
          ++e;
        }
      }
      if (whole) *reinterpret_cast<f32x4*>(dst + f0) = v;
      else
        for (int k = 0; int64_t(f0) + k < total; ++k) dst[f0 + k] = v[k];
    } else {
      f32x4 v;
      if (whole) v = *reinterpret_cast<const f32x4*>(src + f0);
      else
        for (int k = 0; k < 4; ++k) v[k] = (int64_t(f0) + k < total) ? src[f0 + k] : 0.0f;
      if (whole && e + 4 <= n) {
        *reinterpret_cast<f32x4u*>(dst + base + e) = v;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (e == n) { ++i; e = 0; base = (int64_t(f0) + k < total) ? offsets[i] : 0; }
          if (int64_t(f0) + k < total) dst[base + e] = v[k];
          ++e;
        }
      }
    }
  }
}

template <bool TO_CONTIGUOUS>
int launch_gather(float* dst, const float* src, const int64_t* offsets, int nn, int64_t count, void* stream) {
  const int64_t total = int64_t(nn) * count;
  const int64_t quads = (total + 3) / 4;
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((quads + 255) / 256, 1 << 20));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (total + 4 < (int64_t(1) << 32))
    gather_kernel<TO_CONTIGUOUS, uint32_t><<<blocks, 256, 0, s>>>(dst, src, offsets, nn, total);
  else
    gather_kernel<TO_CONTIGUOUS, int64_t><<<blocks, 256, 0, s>>>(dst, src, offsets, nn, total);
  return hipGetLastError() == hipSuccess ? LUB_OK : LUB_EHIP;
}

}  // namespace

extern "C" {

int lubx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int lubx_device_info(int dev, char* name, int name_len, char* arch, int arch_len, int* cu_count, int* clock_khz,
                     size_t* total_mem, int* mem_clock_khz, int* mem_bus_width, size_t* lds_per_block) {
  hipDeviceProp_t p;
  hipError_t e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) return LUB_EHIP;
  copy_n(name, name_len, p.name);
  copy_n(arch, arch_len, p.gcnArchName);
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (clock_khz) *clock_khz = p.clockRate;
  if (total_mem) *total_mem = p.totalGlobalMem;
  if (mem_clock_khz) *mem_clock_khz = p.memoryClockRate;
  if (mem_bus_width) *mem_bus_width = p.memoryBusWidth;
  if (lds_per_block) *lds_per_block = p.sharedMemPerBlock;
  return LUB_OK;
}

void* lubx_malloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

int lubx_free(void* p) { return rc(hipFree(p)); }

void* lubx_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

int lubx_host_free(void* p) { return rc(hipHostFree(p)); }

// The blocking helpers below return only once the device holds the bytes.
// hipMemset (and hipMemcpy from pageable memory) may return while the null
// stream is still writing, and work on a non-blocking stream does not wait
// for the null stream: without the wait a kernel launched next on such a
// stream can read the buffer half-filled.
int lubx_memcpy_htod(void* dst, const void* src, size_t bytes) {
  int r = rc(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return r != LUB_OK ? r : rc(hipStreamSynchronize(nullptr));
}

int lubx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return rc(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int lubx_memcpy_dtod_async(void* dst, const void* src, size_t bytes, void* stream) {
  return rc(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
}

int lubx_memset(void* p, int value, size_t bytes) {
  int r = rc(hipMemset(p, value, bytes));
  return r != LUB_OK ? r : rc(hipStreamSynchronize(nullptr));
}

int lubx_device_sync(void) { return rc(hipDeviceSynchronize()); }

void* lubx_stream_create(void) {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return s;
}

int lubx_stream_destroy(void* stream) { return rc(hipStreamDestroy(static_cast<hipStream_t>(stream))); }

int lubx_stream_sync(void* stream) { return rc(hipStreamSynchronize(static_cast<hipStream_t>(stream))); }

void* lubx_event_create(void) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int lubx_event_destroy(void* ev) { return rc(hipEventDestroy(static_cast<hipEvent_t>(ev))); }

int lubx_event_record(void* ev, void* stream) {
  return rc(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream)));
}

int lubx_event_sync(void* ev) { return rc(hipEventSynchronize(static_cast<hipEvent_t>(ev))); }

float lubx_event_elapsed_ms(void* start, void* stop) {
  if (hipEventSynchronize(static_cast<hipEvent_t>(stop)) != hipSuccess) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)) != hipSuccess)
    return -1.0f;
  return ms;
}

int lubx_graph_capture_begin(void* stream) {
  return rc(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
}

int lubx_graph_capture_end(void* stream, void** graph_exec) {
  if (graph_exec == nullptr) return LUB_ENULL;
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(static_cast<hipStream_t>(stream), &g) != hipSuccess) return LUB_EHIP;
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return LUB_EHIP;
  *graph_exec = ex;
  return LUB_OK;
}

int lubx_graph_launch(void* graph_exec, void* stream) {
  return rc(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), static_cast<hipStream_t>(stream)));
}

int lubx_graph_destroy(void* graph_exec) { return rc(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec))); }

int lubx_gather_f32(float* dst, const float* packed, const int64_t* offsets, int nn, int64_t count, void* stream) {
  if (nn < 1 || count < 0) return LUB_EINVAL_ARG;
  if (count == 0) return LUB_OK;
  if (dst == nullptr || packed == nullptr || offsets == nullptr) return LUB_ENULL;
  if (reinterpret_cast<uintptr_t>(dst) & 15) return LUB_EALIGN;  // the contiguous side moves as dwordx4
  return launch_gather<true>(dst, packed, offsets, nn, count, stream);
}

int lubx_scatter_f32(float* packed, const float* src, const int64_t* offsets, int nn, int64_t count, void* stream) {
  if (nn < 1 || count < 0) return LUB_EINVAL_ARG;
  if (count == 0) return LUB_OK;
  if (packed == nullptr || src == nullptr || offsets == nullptr) return LUB_ENULL;
  if (reinterpret_cast<uintptr_t>(src) & 15) return LUB_EALIGN;
  return launch_gather<false>(packed, src, offsets, nn, count, stream);
}

int lubx_copy_f32(float* dst, const float* src, size_t count, void* stream) {
  if (count == 0) return LUB_OK;
  if (dst == nullptr || src == nullptr) return LUB_ENULL;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) == 0;
  size_t n4 = aligned ? count / 4 : 0;
  if (n4 > 0) {
    size_t blocks = std::min<size_t>((n4 + 255) / 256, size_t(1) << 20);
    copy_kernel<<<static_cast<unsigned>(blocks), 256, 0, s>>>(reinterpret_cast<f32x4*>(dst),
                                                             reinterpret_cast<const f32x4*>(src), n4);
  }
  if (n4 * 4 < count) {
    size_t rest = count - n4 * 4;
    size_t blocks = std::min<size_t>((rest + 255) / 256, size_t(1) << 20);
    copy_scalar_kernel<<<static_cast<unsigned>(blocks), 256, 0, s>>>(dst, src, n4 * 4, count);
  }
  return rc(hipGetLastError());
}

}  // extern "C"
