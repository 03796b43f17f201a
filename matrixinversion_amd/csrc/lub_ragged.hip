// lub_ragged.hip -- C ABI of the ragged (mixed-n) path (include/lubatch.h).
//
// A plan owns the workspace of the device-side planning kernels; enqueueing
// a batch on a plan allocates nothing and synchronises nothing, so the whole
// sequence (hist, scan, scatter, invert) can be captured in a hipGraph
// once and replayed.  lu_batched_inverse_ragged keeps one plan per host
// thread and grows it when a larger batch arrives.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>

#include "lu_ragged.hpp"
#include "lub_internal.h"
#include "lubatch.h"

struct lub_ragged_plan {
  int device = -1;
  int64_t capacity = 0;  // matrices
  char* ws = nullptr;    // lub::RaggedWs layout
  int cus = 0;           // compute units of `device`
  int resident[lub::kNumPivot] = {};  // resident waves of each pivot rule's kernel on `device`
  int* header() const { return reinterpret_cast<int*>(ws); }
  int* bhist() const { return reinterpret_cast<int*>(ws + lub::RaggedWs::bhist_offset()); }
  int64_t* soffs() const { return reinterpret_cast<int64_t*>(ws + lub::RaggedWs::slots_offset(capacity)); }
  int64_t* spoffs() const { return soffs() + capacity; }
};

namespace lub {

// Planning kernels (one translation unit: they are not templates).  The
// block table is class-major: bhist[c * nblocks + b] for chunk b of this call
// (nblocks chunks of CHUNK matrices).

constexpr int kPlanThreads = 256;
constexpr int kPlanItems = RaggedWs::CHUNK / kPlanThreads;  // matrices per thread
constexpr int kScanKeep = 8;  // table entries per scan lane kept in registers
static_assert(kPlanItems * kPlanThreads == RaggedWs::CHUNK, "a planning block covers one chunk");

// Counts per n of one chunk (LDS histogram), written to the chunk's column
// of the block table: [n] for n = 1..32, [0] for n outside 1..32.
__global__ __launch_bounds__(kPlanThreads) void ragged_hist(const int* __restrict__ ns, int64_t batch,
                                                            int* __restrict__ bhist, int64_t nblocks) {
  __shared__ int hist[RaggedWs::BSTRIDE];
  if (threadIdx.x < RaggedWs::BSTRIDE) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t b0 = int64_t(blockIdx.x) * RaggedWs::CHUNK;
#pragma unroll
  for (int k = 0; k < kPlanItems; ++k) {
    const int64_t i = b0 + k * kPlanThreads + threadIdx.x;
    if (i < batch) {
      const int n = ns[i];
      atomicAdd(&hist[(n >= 1 && n <= 32) ? n : 0], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < RaggedWs::BSTRIDE) bhist[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

// Workgroup c - 1 scans class c's row of the block table: each chunk's count
// becomes the chunk's first slot within the class (an exclusive prefix over
// chunks), and the class total goes to the header.  Thread t owns a
// consecutive range of chunks; the 256 range sums are scanned in LDS.
__global__ __launch_bounds__(kPlanThreads) void ragged_scan(int* __restrict__ ws, int* __restrict__ bhist,
                                                            int64_t nblocks) {
  __shared__ int part[kPlanThreads];
  const int c = blockIdx.x + 1;
  const int t = threadIdx.x;
  int* __restrict__ row = bhist + c * nblocks;
  const int64_t per = (nblocks + kPlanThreads - 1) / kPlanThreads;
  const int64_t lo = min(nblocks, t * per);
  const int64_t hi = min(nblocks, lo + per);
  const bool kept = per <= kScanKeep;
  int v[kScanKeep];
  int sum = 0;
  if (kept) {
#pragma unroll
    for (int q = 0; q < kScanKeep; ++q) v[q] = (lo + q < hi) ? row[lo + q] : 0;
#pragma unroll
    for (int q = 0; q < kScanKeep; ++q) sum += v[q];
  } else {
    for (int64_t b = lo; b < hi; ++b) sum += row[b];
  }
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < kPlanThreads; d <<= 1) {  // inclusive scan of the range sums
    const int add = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int slot = part[t] - sum;
  if (t == kPlanThreads - 1) ws[RaggedWs::COUNTS + c] = part[t];
  if (kept) {
#pragma unroll
    for (int q = 0; q < kScanKeep; ++q) {
      if (lo + q < hi) row[lo + q] = slot;
      slot += v[q];
    }
  } else {
    for (int64_t b = lo; b < hi; ++b) {
      const int cnt = row[b];
      row[b] = slot;
      slot += cnt;
    }
  }
}

// The class table from the class totals: classes 32, 31, ..., 1 in order;
// a descriptor holds MPW(n) * grab(n) matrices, at most kRaggedMaxPer (one
// slot per lane), so grab is clamped to 16 with 16-lane rows and 4 with
// quads.  Every scatter block derives it (32 values); block 0 publishes it
// in the header for the inversion kernel.
__device__ void ragged_class_table(int* __restrict__ ws, int64_t batch, const RaggedTuning& tune, bool publish,
                                   int* first_slot) {
  __shared__ int count[33];
  __shared__ int tiles_of[33];
  __shared__ int lper[33];
  const int c = threadIdx.x;
  if (c >= 1 && c <= 32) {
    const int cnt = ws[RaggedWs::COUNTS + c];
    const int layout = (tune.layout[c] == 1 && c <= kRaggedQuadMaxN) ? 1 : 0;
    const int mpw = ragged_mpw(c, layout);
    const int per_desc = mpw * min(max(1, (int)tune.grab[c]), kRaggedMaxPer / mpw);
    count[c] = cnt;
    tiles_of[c] = (cnt + per_desc - 1) / per_desc;
    lper[c] = layout | (per_desc << 8);
  }
  __syncthreads();
  if (c == 0) {
    int first = 0, tiles = 0;
    for (int n = 32; n >= 1; --n) {
      first_slot[n] = first;
      if (publish) {
        ws[RaggedWs::CLASS_FIRST + n] = first;
        ws[RaggedWs::DESC_FIRST + n] = tiles;
        ws[RaggedWs::LPER + n] = lper[n];
      }
      first += count[n];
      tiles += tiles_of[n];
    }
    if (publish) {
      ws[RaggedWs::DESC_FIRST + 0] = tiles;
      ws[RaggedWs::CLASS_FIRST + 0] = first;
      ws[RaggedWs::LPER + 0] = 0;
      ws[RaggedWs::META + 2] = static_cast<int>(batch - first);  // matrices with n outside 1..32
    }
  }
  __syncthreads();
}

// Slot of each matrix = its class's first slot + its chunk's first slot
// within the class (from the scan) + a rank from an LDS atomic.  The order
// within a chunk is therefore not fixed; results do not depend on it.
__global__ __launch_bounds__(kPlanThreads) void ragged_scatter(const int* __restrict__ ns,
                                                               const int64_t* __restrict__ offs,
                                                               const int64_t* __restrict__ poffs, int64_t batch,
                                                               int* __restrict__ ws, const int* __restrict__ bhist,
                                                               int64_t nblocks, RaggedTuning tune,
                                                               int64_t* __restrict__ soffs,
                                                               int64_t* __restrict__ spoffs) {
  __shared__ int first_slot[33];
  __shared__ int cur[33];
  ragged_class_table(ws, batch, tune, blockIdx.x == 0, first_slot);
  const int c = threadIdx.x;
  if (c >= 1 && c <= 32) cur[c] = first_slot[c] + bhist[c * nblocks + blockIdx.x];
  __syncthreads();
  const int64_t b0 = int64_t(blockIdx.x) * RaggedWs::CHUNK;
#pragma unroll
  for (int k = 0; k < kPlanItems; ++k) {
    const int64_t i = b0 + k * kPlanThreads + threadIdx.x;
    if (i < batch) {
      const int n = ns[i];
      if (n >= 1 && n <= 32) {
        const int pos = atomicAdd(&cur[n], 1);
        soffs[pos] = offs[i];
        if (poffs != nullptr) spoffs[pos] = poffs[i];
      }
    }
  }
}

// The three planning steps in one workgroup, for batches of up to
// kSmallPlanMax matrices (serving batches): an LDS histogram, the class
// table (published), then each matrix's slot from an LDS-atomic rank within
// its class.  One launch instead of three; the header and slot arrays it
// writes have the same meaning (slot order within a class is not fixed, and
// results do not depend on it).
constexpr int kSmallPlanThreads = 1024;
constexpr int64_t kSmallPlanMax = 16384;

__global__ __launch_bounds__(kSmallPlanThreads) void ragged_plan_small(const int* __restrict__ ns,
                                                                       const int64_t* __restrict__ offs,
                                                                       const int64_t* __restrict__ poffs, int64_t batch,
                                                                       int* __restrict__ ws, RaggedTuning tune,
                                                                       int64_t* __restrict__ soffs,
                                                                       int64_t* __restrict__ spoffs) {
  __shared__ int hist[RaggedWs::BSTRIDE];
  __shared__ int first_slot[33];
  __shared__ int cur[33];
  const int t = threadIdx.x;
  if (t < RaggedWs::BSTRIDE) hist[t] = 0;
  __syncthreads();
  for (int64_t i = t; i < batch; i += kSmallPlanThreads) {
    const int n = ns[i];
    atomicAdd(&hist[(n >= 1 && n <= 32) ? n : 0], 1);
  }
  __syncthreads();
  if (t >= 1 && t <= 32) ws[RaggedWs::COUNTS + t] = hist[t];  // read back below by the same thread
  ragged_class_table(ws, batch, tune, true, first_slot);
  if (t >= 1 && t <= 32) cur[t] = first_slot[t];
  __syncthreads();
  for (int64_t i = t; i < batch; i += kSmallPlanThreads) {
    const int n = ns[i];
    if (n >= 1 && n <= 32) {
      const int pos = atomicAdd(&cur[n], 1);
      soffs[pos] = offs[i];
      if (poffs != nullptr) spoffs[pos] = poffs[i];
    }
  }
}

}  // namespace lub

namespace {

constexpr int64_t kMaxRaggedBatch = (int64_t(1) << 31) - 64;

int resident_per_cu(const void* fn) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0) != hipSuccess) return 0;
  return per_cu;
}

int plan_alloc(lub_ragged_plan* p, int64_t capacity) {
  void* ws = nullptr;
  if (hipMalloc(&ws, lub::RaggedWs::bytes(capacity)) != hipSuccess) return LUB_EHIP;
  p->ws = static_cast<char*>(ws);
  p->capacity = capacity;
  return LUB_OK;
}

void plan_release(lub_ragged_plan* p) {
  if (p->ws) (void)hipFree(p->ws);
  p->ws = nullptr;
  p->capacity = 0;
}

unsigned plan_blocks(int64_t batch) { return static_cast<unsigned>(lub::RaggedWs::blocks(batch)); }

// Matrices per descriptor of class n under tuning `rt` (the scan kernel
// computes the same).
int ragged_per(const lub::RaggedTuning& rt, int n) {
  const int layout = (rt.layout[n] == 1 && n <= lub::kRaggedQuadMaxN) ? 1 : 0;
  const int mpw = lub::ragged_mpw(n, layout);
  return mpw * std::min(std::max(1, int(rt.grab[n])), lub::kRaggedMaxPer / mpw);
}

// Grid of the inversion kernel: by default at least the descriptor count
// (ceil(batch / smallest matrices-per-descriptor) + 32 bounds it), so each
// workgroup does one descriptor; LUB_RAGGED_WAVES_PER_CU=<k> launches k
// persistent waves per compute unit instead.
int64_t ragged_grid(int cus, const lub::RaggedTuning& rt, int64_t batch) {
  static const int per_cu = [] {
    const char* e = std::getenv("LUB_RAGGED_WAVES_PER_CU");
    return e != nullptr ? std::atoi(e) : 0;
  }();
  if (per_cu > 0 && cus > 0) return int64_t(cus) * per_cu;
  int min_per = lub::kRaggedMaxPer;
  for (int n = 1; n <= 32; ++n) min_per = std::min(min_per, ragged_per(rt, n));
  return std::min<int64_t>((batch + min_per - 1) / min_per + 32, int64_t(1) << 25);  // grid * 64 within 32 bits
}

int check_ragged_args(float* A, const int64_t* offsets, const int* ns, int* piv, const int64_t* piv_offsets,
                      int64_t batch, int flags) {
  if (batch < 0 || batch > kMaxRaggedBatch) return LUB_EINVAL_BATCH;
  if ((flags & ~(LUB_PIVOT_MASK | LUB_LAYOUT_MASK)) != 0 || (flags & LUB_PIVOT_MASK) > LUB_PIVOT_GEPP ||
      (flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_MASK)
    return LUB_EFLAGS;
  if (batch == 0) return LUB_OK;
  if (A == nullptr || offsets == nullptr || ns == nullptr) return LUB_ENULL;
  if (piv != nullptr && piv_offsets == nullptr) return LUB_ENULL;
  return -1;  // proceed
}

// The four launches of one ragged batch on `stream`, with the planning
// workspace at `ws` (lub::RaggedWs layout for `capacity` matrices, batch <=
// capacity, arguments already checked).  No allocation, no synchronisation.
int enqueue_ragged(char* ws, int64_t capacity, int cus, float* A, const int64_t* offsets, const int* ns, int* piv,
                   const int64_t* piv_offsets, int64_t batch, int flags, hipStream_t s) {
  const int pv = flags & LUB_PIVOT_MASK;
  const void* fn = lub::kernel_table().ragged[pv];
  if (fn == nullptr) return LUB_ENODEVICE;

  lub::Tuning tuning = lub::current_tuning();
  lub::RaggedTuning rt;
  for (int n = 0; n <= 32; ++n) {
    int layout = tuning.layout[n];
    if ((flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_ROW16) layout = 0;
    if ((flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_QUAD4) layout = 1;
    rt.layout[n] = static_cast<int8_t>(layout);
    rt.grab[n] = tuning.grab[n];
  }
  int* hdr = reinterpret_cast<int*>(ws);
  int* bhist = reinterpret_cast<int*>(ws + lub::RaggedWs::bhist_offset());
  int64_t* soffs = reinterpret_cast<int64_t*>(ws + lub::RaggedWs::slots_offset(capacity));
  int64_t* spoffs_all = soffs + capacity;
  if (batch <= lub::kSmallPlanMax) {
    lub::ragged_plan_small<<<1, lub::kSmallPlanThreads, 0, s>>>(ns, offsets, piv ? piv_offsets : nullptr, batch, hdr,
                                                                rt, soffs, spoffs_all);
  } else {
    const unsigned blocks = plan_blocks(batch);
    const int64_t nblocks = blocks;
    lub::ragged_hist<<<blocks, lub::kPlanThreads, 0, s>>>(ns, batch, bhist, nblocks);
    lub::ragged_scan<<<32, lub::kPlanThreads, 0, s>>>(hdr, bhist, nblocks);
    lub::ragged_scatter<<<blocks, lub::kPlanThreads, 0, s>>>(ns, offsets, piv ? piv_offsets : nullptr, batch, hdr,
                                                              bhist, nblocks, rt, soffs, spoffs_all);
  }
  int64_t* spoffs = piv ? spoffs_all : nullptr;
  const int* hdr_arg = hdr;
  void* args[] = {&A, &soffs, &piv, &spoffs, &hdr_arg};
  const int64_t grid = ragged_grid(cus, rt, batch);
  hipError_t e = hipLaunchKernel(fn, dim3(static_cast<unsigned>(grid)), dim3(64), args, 0, s);
  if (e != hipSuccess) return LUB_EHIP;
  return hipGetLastError() == hipSuccess ? LUB_OK : LUB_EHIP;
}

// Scratch of lu_batched_inverse_ragged_noalloc: the planning workspace for
// batches of up to LUB_NOALLOC_MAX_BATCH matrices (about 16.1 MiB), one copy
// per device for the life of the process.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// the device address comes from hipGetSymbolAddress, once per device.
constexpr int64_t kNoAllocMax = LUB_NOALLOC_MAX_BATCH;
constexpr int64_t kNoAllocBytes = lub::RaggedWs::bytes(kNoAllocMax);
__device__ __align__(256) char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
struct NoAllocDevice {
  std::once_flag once;
  char* ptr = nullptr;  // device address of `scratch` on this device
  int cus = 0;
  int status = LUB_ENODEVICE;
};
NoAllocDevice g_noalloc[kMaxDevices];

// Looks up this device's scratch address (and compute-unit count) on the
// first call made with the device current; later calls read the cache.
const NoAllocDevice* noalloc_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  NoAllocDevice& d = g_noalloc[dev];
  std::call_once(d.once, [&d, dev] {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
      d.status = LUB_EHIP;
      return;
    }
    if (hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      d.status = LUB_EHIP;
      return;
    }
    d.ptr = static_cast<char*>(p);
    d.status = LUB_OK;
  });
  return &d;
}

struct ThreadPlan {
  lub_ragged_plan* plan = nullptr;
  ~ThreadPlan() {
    if (plan) {
      plan_release(plan);
      delete plan;
    }
  }
};
thread_local ThreadPlan g_thread_plan;

}  // namespace

extern "C" {

int lu_ragged_plan_create(lub_ragged_plan** plan, int64_t max_batch) {
  if (plan == nullptr) return LUB_ENULL;
  // matrix indices and descriptor starts are int32 on the device
  if (max_batch < 0 || max_batch > kMaxRaggedBatch) return LUB_EINVAL_BATCH;
  auto* p = new lub_ragged_plan;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return LUB_EHIP;
  }
  // Occupancy is queried here, not at enqueue time, so that an enqueue under
  // stream capture makes launches only.
  const lub::KernelTable& kt = lub::kernel_table();
  if (hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
      p->cus <= 0) {
    delete p;
    return LUB_ENODEVICE;
  }
  for (int pv = 0; pv < lub::kNumPivot; ++pv) {
    p->resident[pv] = kt.ragged[pv] ? p->cus * resident_per_cu(kt.ragged[pv]) : 0;
    if (p->resident[pv] <= 0) {
      delete p;
      return LUB_ENODEVICE;
    }
  }
  int rc = plan_alloc(p, std::max<int64_t>(max_batch, 1));
  if (rc != LUB_OK) {
    delete p;
    return rc;
  }
  *plan = p;
  return LUB_OK;
}

int lu_ragged_plan_destroy(lub_ragged_plan* plan) {
  if (plan == nullptr) return LUB_ENULL;
  plan_release(plan);
  delete plan;
  return LUB_OK;
}

int lu_batched_inverse_ragged_plan(lub_ragged_plan* plan, float* A, const int64_t* offsets, const int* ns, int* piv,
                                   const int64_t* piv_offsets, int64_t batch, int flags, void* stream) {
  if (plan == nullptr) return LUB_ENULL;
  const int rc = check_ragged_args(A, offsets, ns, piv, piv_offsets, batch, flags);
  if (rc >= 0) return rc;
  if (batch > plan->capacity) return LUB_EINVAL_ARG;
  return enqueue_ragged(plan->ws, plan->capacity, plan->cus, A, offsets, ns, piv, piv_offsets, batch, flags,
                        static_cast<hipStream_t>(stream));
}

int lu_batched_inverse_ragged(float* A, const int64_t* offsets, const int* ns, int* piv, const int64_t* piv_offsets,
                              int batch, int flags, void* stream) {
  if (batch < 0) return LUB_EINVAL_BATCH;
  if (batch == 0) return LUB_OK;
  ThreadPlan& tp = g_thread_plan;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return LUB_EHIP;
  if (tp.plan == nullptr || tp.plan->device != dev || tp.plan->capacity < batch) {
    if (tp.plan != nullptr) {
      // the previous plan may still be in use by work on some stream
      if (hipDeviceSynchronize() != hipSuccess) return LUB_EHIP;
      plan_release(tp.plan);
      delete tp.plan;
      tp.plan = nullptr;
    }
    lub_ragged_plan* p = nullptr;
    int rc = lu_ragged_plan_create(&p, std::max<int64_t>(batch, 1024));
    if (rc != LUB_OK) return rc;
    tp.plan = p;
  }
  return lu_batched_inverse_ragged_plan(tp.plan, A, offsets, ns, piv, piv_offsets, batch, flags, stream);
}

int lu_batched_inverse_ragged_noalloc(float* A, const int64_t* offsets, const int* ns, int* piv,
                                      const int64_t* piv_offsets, int64_t batch, int flags, void* stream) {
  const int rc = check_ragged_args(A, offsets, ns, piv, piv_offsets, batch, flags);
  if (rc >= 0) return rc;
  if (batch > kNoAllocMax) return LUB_ETOOBIG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return LUB_EHIP;
  const NoAllocDevice* d = noalloc_device(dev);
  if (d == nullptr) return LUB_ENODEVICE;
  if (d->status != LUB_OK) return d->status;
  return enqueue_ragged(d->ptr, kNoAllocMax, d->cus, A, offsets, ns, piv, piv_offsets, batch, flags,
                        static_cast<hipStream_t>(stream));
}

int lu_ragged_invalid_count(const lub_ragged_plan* plan, int* out) {
  if (plan == nullptr || out == nullptr) return LUB_ENULL;
  return hipMemcpy(out, plan->header() + lub::RaggedWs::META + 2, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess
             ? LUB_OK
             : LUB_EHIP;
}

}  // extern "C"
