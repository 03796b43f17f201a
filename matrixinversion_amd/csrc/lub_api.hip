// lub_api.hip -- the C ABI of liblubatch.so (include/lubatch.h).
//
// Replaces the reference's per-variant launch code (e.g. parallel_pivot/
// luBatchedInplace.cu:8-11 for the grid arithmetic, :118-133 for the launch):
// n, batch and the pivot rule are runtime arguments here, dispatched to the
// per-n kernel instantiations; errors come back as status codes instead of
// CUDA_CHECK's print-and-exit (luBatchedInplace.cuh:9).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "lubatch.h"
#include "lub_internal.h"

namespace {

lub::KernelTable g_table;
std::once_flag g_table_once;
thread_local char g_last_error[256] = "";

void set_last_error(const char* what, hipError_t e) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
}

const lub::KernelTable& table() {
  std::call_once(g_table_once, [] {
    std::memset(&g_table, 0, sizeof(g_table));
    lub_register_1_8(&g_table);
    lub_register_9_16(&g_table);
    lub_register_17_20(&g_table);
    lub_register_21_24(&g_table);
    lub_register_25_28(&g_table);
    lub_register_29_32(&g_table);
  });
  return g_table;
}

// Workgroups are one wave (64 lanes) and own a tile of 4 matrices; the grid
// strides over tiles.  2^22 workgroups is far above what the 256 CUs hold at
// once, so the hardware dispatcher balances the work; the cap only keeps the
// dispatch within the 32-bit grid size for very large batches.
constexpr int64_t kMaxGrid = int64_t(1) << 22;

int validate(const float* A, const int* piv, int n, int64_t batch, int flags) {
  if (n < 1 || n > LUB_MAX_N) return LUB_EINVAL_N;
  if (batch < 0) return LUB_EINVAL_BATCH;
  if ((flags & ~LUB_PIVOT_MASK) != 0 || (flags & LUB_PIVOT_MASK) > LUB_PIVOT_TREE) return LUB_EFLAGS;
  if (batch == 0) return LUB_OK;
  if (A == nullptr) return LUB_ENULL;
  if ((reinterpret_cast<uintptr_t>(A) & 3u) != 0 || (reinterpret_cast<uintptr_t>(piv) & 3u) != 0) return LUB_EALIGN;
  return LUB_OK;
}

}  // namespace

extern "C" {

int lu_batched_inverse_async64(float* A, int* piv, int n, int64_t batch, int flags, void* stream) {
  int rc = validate(A, piv, n, batch, flags);
  if (rc != LUB_OK || batch == 0) return rc;
  const void* fn = table().plain[n][flags & LUB_PIVOT_MASK];
  if (fn == nullptr) return LUB_ENODEVICE;
  int64_t ntiles = (batch + 3) / 4;
  unsigned grid = static_cast<unsigned>(std::min(ntiles, kMaxGrid));
  int vec = (reinterpret_cast<uintptr_t>(A) & 15u) == 0 ? 1 : 0;
  void* args[] = {&A, &piv, &batch, &ntiles, &vec};
  hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(64), args, 0, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    set_last_error("hipLaunchKernel", e);
    return LUB_EHIP;
  }
  return LUB_OK;
}

int lu_batched_inverse_async(float* A, int* piv, int n, int batch, int flags, void* stream) {
  return lu_batched_inverse_async64(A, piv, n, static_cast<int64_t>(batch), flags, stream);
}

int lu_batched_inverse(float* A, int* piv, int n, int batch, int flags) {
  int rc = lu_batched_inverse_async64(A, piv, n, static_cast<int64_t>(batch), flags, nullptr);
  if (rc != LUB_OK || batch == 0) return rc;
  hipError_t e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    set_last_error("hipStreamSynchronize", e);
    return LUB_EHIP;
  }
  return LUB_OK;
}

const char* lu_status_string(int status) {
  switch (status) {
    case LUB_OK: return "ok";
    case LUB_EINVAL_N: return "n outside [1, 32]";
    case LUB_EINVAL_BATCH: return "batch < 0";
    case LUB_ENULL: return "required pointer is NULL";
    case LUB_EFLAGS: return "unknown flags or pivot rule";
    case LUB_EHIP: return "HIP runtime error";
    case LUB_ENODEVICE: return "no kernel for this device";
    case LUB_EALIGN: return "pointer not 4-byte aligned";
    case LUB_EINVAL_ARG: return "invalid argument";
    default: return "unknown status";
  }
}

const char* lu_version(void) { return "0.1.0"; }

// Text of the last HIP failure seen by this thread (empty if none).
const char* lu_last_error(void) { return g_last_error; }

}  // extern "C"
