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
#include <cstdlib>
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
    lub_register_ragged_0(&g_table);
    lub_register_ragged_1(&g_table);
    lub_register_ragged_3(&g_table);
    lub_register_ragged_4(&g_table);
    lub_register_refined_17_24(&g_table);
    lub_register_refined_25_32(&g_table);
    g_table.ragged[LUB_PIVOT_SERIAL] = g_table.ragged[LUB_PIVOT_PARALLEL];  // same pivots, same kernel
  });
  return g_table;
}

// Per-n launch choices.  Defaults from MI355X measurements at 1M matrices
// (quads for n <= 12; one wave tile per ragged queue descriptor), replaced
// by lu_tuning_set (layout, grab) and lu_tuning_set_lds_pad (pad): the
// autotuner's table.
std::mutex g_tuning_mutex;
lub::Tuning g_tuning = [] {
  lub::Tuning t;
  for (int n = 0; n <= lub::kMaxN; ++n) {
    // measured on MI355X (python -m matrixinversion_amd.autotune): quads win
    // up to n = 12; ragged work items of 8 row tiles (4 at n = 32) beat
    // single tiles by 5-10 % for n >= 13; tuned on the mixed batch
    // (--objective mix) no other choice wins by more than 0.5 %, n = 30
    // taking 16 row tiles; LDS rows padded for n % 4 == 0 except n = 28
    t.layout[n] = n <= 12 ? 1 : 0;
    t.grab[n] = n <= 12 ? 1 : (n == 32 ? 4 : (n == 30 ? 16 : 8));
    t.pad[n] = lub::initial_pad(n);
  }
  return t;
}();

// Workgroups are one wave (64 lanes) and own a tile of 4 matrices at a time;
// the kernels stride over tiles.  Two launch shapes:
//   one workgroup per tile (default): the hardware dispatcher starts a new
//     wave whenever one finishes, so the loads of one wave overlap the
//     arithmetic of the others on the same SIMD;
//   persistent: the grid is `waves_per_cu` workgroups per CU (capped by the
//     kernel's occupancy) and each wave prefetches its next tile into
//     registers during the current tile's Gauss-Jordan.
//   several tiles per workgroup: grid = ceil(tiles / T), for small n where a
//     single tile is too little work to amortise a wave launch.
// LUB_WAVES_PER_CU=<k> selects the persistent shape with k waves per CU;
// LUB_TILES_PER_WG=<T> the grouped one (defaults: one workgroup per tile).
constexpr int64_t kMaxGrid = int64_t(1) << 22;  // keeps grid*64 within 32 bits

int env_int(const char* name) {
  const char* e = std::getenv(name);
  return e ? std::atoi(e) : 0;
}

int env_waves_per_cu() {
  static int v = env_int("LUB_WAVES_PER_CU");
  return v;
}

int env_tiles_per_wg() {
  static int v = env_int("LUB_TILES_PER_WG");
  return v;
}

struct Residency {
  int device = -1;
  int waves = 0;  // resident single-wave workgroups for the whole device
};
std::mutex g_res_mutex;
Residency g_res[LUB_MAX_N + 1][lub::kNumPivot][lub::kNumLayout];

int resident_workgroups(const void* fn, int n, int pv, int layout, int cap_per_cu) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(g_res_mutex);
  Residency& r = g_res[n][pv][layout];
  if (r.device != dev) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0) != hipSuccess) return 0;
    r.device = dev;
    r.waves = std::max(1, cus * std::max(1, std::min(per_cu, cap_per_cu)));
  }
  return r.waves;
}

// Lanes per matrix for each n: 4-lane quads (16 matrices per wave) where the
// 16-lane row would leave most lanes idle, 16-lane rows otherwise (tuning
// table; measured on MI355X at 1M matrices: quads are 3x faster at n <= 4,
// 1.5x at n = 8, even at n = 12, slower at n = 16).  LUB_LANES=4|16 forces
// one layout (where it exists).
int default_layout(int n) {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  return (g_tuning.layout[n] == 1 && n <= lub::kLayoutMaxN[1]) ? 1 : 0;
}

int choose_layout(int n, int flags) {
  switch (flags & LUB_LAYOUT_MASK) {
    case LUB_LAYOUT_ROW16: return 0;
    case LUB_LAYOUT_QUAD4: return 1;
    default: break;
  }
  static int forced = env_int("LUB_LANES");
  if (forced == 4 && n <= lub::kLayoutMaxN[1]) return 1;
  if (forced == 16) return 0;
  return default_layout(n);
}

int validate(const float* A, const int* piv, int n, int64_t batch, int flags) {
  if (n < 1 || n > LUB_MAX_N) return LUB_EINVAL_N;
  if (batch < 0) return LUB_EINVAL_BATCH;
  if ((flags & ~(LUB_PIVOT_MASK | LUB_LAYOUT_MASK)) != 0 || (flags & LUB_PIVOT_MASK) > LUB_PIVOT_GEPP) return LUB_EFLAGS;
  if ((flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_MASK) return LUB_EFLAGS;
  if ((flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_QUAD4 && n > lub::kLayoutMaxN[1]) return LUB_EFLAGS;
  if (batch == 0) return LUB_OK;
  if (A == nullptr) return LUB_ENULL;
  if ((reinterpret_cast<uintptr_t>(A) & 3u) != 0 || (reinterpret_cast<uintptr_t>(piv) & 3u) != 0) return LUB_EALIGN;
  return LUB_OK;
}

}  // namespace

namespace lub {
Tuning current_tuning() {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  return g_tuning;
}
const KernelTable& kernel_table() { return table(); }

// The plain launch; info (nullable) receives LAPACK-style per-matrix info.
int enqueue_plain(float* A, int* piv, int* info, int n, int64_t batch, int flags, void* stream) {
  int rc = validate(A, piv, n, batch, flags);
  if (rc != LUB_OK || batch == 0) return rc;
  const int layout = choose_layout(n, flags);
  bool flip = false;
  if (n % 4 == 0) {
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    flip = g_tuning.pad[n] != lub::geo_pad(n);
  }
  const void* fn = (flip ? table().plain_flip : table().plain)[n][flags & LUB_PIVOT_MASK][layout];
  if (fn == nullptr) return LUB_ENODEVICE;
  const int per_tile = 64 / lub::kLayoutLanes[layout];  // matrices per wave tile
  int64_t ntiles = (batch + per_tile - 1) / per_tile;
  int64_t cap = kMaxGrid;
  if (int t = env_tiles_per_wg(); t > 1) cap = std::min<int64_t>(cap, (ntiles + t - 1) / t);
  if (int wpc = env_waves_per_cu(); wpc > 0) {
    cap = resident_workgroups(fn, n, flags & LUB_PIVOT_MASK, layout, wpc);
    if (cap <= 0) return LUB_ENODEVICE;
  }
  unsigned grid = static_cast<unsigned>(std::min<int64_t>(ntiles, cap));
  int vec = (reinterpret_cast<uintptr_t>(A) & 15u) == 0 ? 1 : 0;
  void* args[] = {&A, &piv, &info, &batch, &ntiles, &vec};
  hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(64), args, 0, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) {
    set_last_error("hipLaunchKernel", e);
    return LUB_EHIP;
  }
  return LUB_OK;
}
}  // namespace lub



extern "C" {

int lu_tuning_set(const signed char* layout, const signed char* grab) {
  if (layout == nullptr || grab == nullptr) return LUB_ENULL;
  for (int n = 0; n <= lub::kMaxN; ++n)
    if (layout[n] < 0 || layout[n] > 1 || grab[n] < 1) return LUB_EINVAL_ARG;
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  for (int n = 0; n <= lub::kMaxN; ++n) {
    g_tuning.layout[n] = static_cast<int8_t>(layout[n]);
    g_tuning.grab[n] = static_cast<int8_t>(grab[n]);
  }
  return LUB_OK;
}

int lu_tuning_set_lds_pad(const signed char* pad) {
  if (pad == nullptr) return LUB_ENULL;
  for (int n = 0; n <= lub::kMaxN; ++n)
    if (pad[n] < 0 || pad[n] > 1 || (pad[n] == 1 && (n % 4 != 0 || n == 0))) return LUB_EINVAL_ARG;
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  for (int n = 0; n <= lub::kMaxN; ++n) g_tuning.pad[n] = static_cast<int8_t>(pad[n]);
  return LUB_OK;
}

int lu_tuning_get_lds_pad(signed char* pad) {
  if (pad == nullptr) return LUB_ENULL;
  lub::Tuning t = lub::current_tuning();
  for (int n = 0; n <= lub::kMaxN; ++n) pad[n] = t.pad[n];
  return LUB_OK;
}

int lu_tuning_get(signed char* layout, signed char* grab) {
  if (layout == nullptr || grab == nullptr) return LUB_ENULL;
  lub::Tuning t = lub::current_tuning();
  for (int n = 0; n <= lub::kMaxN; ++n) {
    layout[n] = t.layout[n];
    grab[n] = t.grab[n];
  }
  return LUB_OK;
}

int lu_batched_inverse_async64(float* A, int* piv, int n, int64_t batch, int flags, void* stream) {
  return lub::enqueue_plain(A, piv, nullptr, n, batch, flags, stream);
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
    case LUB_EBUSY: return "no room under the server budget now";
    case LUB_ETOOBIG: return "batch above the entry point's fixed capacity";
    default: return "unknown status";
  }
}

const char* lu_version(void) { return "0.2.0"; }

// Hash of the sources this library was built from (set by the build).
#ifndef LUB_BUILD_ID
#define LUB_BUILD_ID "unknown"
#endif
const char* lu_build_id(void) { return LUB_BUILD_ID; }

// Text of the last HIP failure seen by this thread (empty if none).
const char* lu_last_error(void) { return g_last_error; }

}  // extern "C"
