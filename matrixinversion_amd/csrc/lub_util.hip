// lub_util.hip -- device plumbing for the ctypes harness (include/lubatch_util.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "lubatch.h"
#include "lubatch_util.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

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

int lubx_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return rc(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int lubx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return rc(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int lubx_memcpy_dtod_async(void* dst, const void* src, size_t bytes, void* stream) {
  return rc(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
}

int lubx_memset(void* p, int value, size_t bytes) { return rc(hipMemset(p, value, bytes)); }

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
