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

// Planning kernels (one translation unit: they are not templates).

constexpr int kPlanThreads = 256;
constexpr int kPlanItems = RaggedWs::CHUNK / kPlanThreads;  // matrices per thread
constexpr int kScanThreads = 1024;                           // 32 lanes per class
static_assert(kPlanItems * kPlanThreads == RaggedWs::CHUNK, "a planning block covers one chunk");

// Counts per n of one chunk of CHUNK matrices (LDS histogram), written as the
// chunk's row of the block table: [0] matrices with n outside 1..32, [n].
__global__ __launch_bounds__(kPlanThreads) void ragged_hist(const int* __restrict__ ns, int64_t batch,
                                                            int* __restrict__ bhist) {
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
  if (threadIdx.x < RaggedWs::BSTRIDE) bhist[int64_t(blockIdx.x) * RaggedWs::BSTRIDE + threadIdx.x] = hist[threadIdx.x];
}

// One workgroup.  Lanes 32(c-1) .. 32(c-1)+31 own class c and split the
// blocks into 32 consecutive ranges: per range the class count, then the
// class table (classes 32, 31, ..., 1 in order; a descriptor holds MPW(n) *
// grab(n) matrices, at most kRaggedMaxPer -- one slot per lane -- so grab is
// clamped to 16 with 16-lane rows and 4 with quads), then each block's
// count in the block table is replaced by the block's first slot in class c.
__global__ __launch_bounds__(kScanThreads) void ragged_scan(int* __restrict__ ws, int* __restrict__ bhist,
                                                            int64_t nblocks, RaggedTuning tune) {
  __shared__ int part[32][33];  // [class - 1][range] (padded row)
  __shared__ int bad[32];       // per range: matrices with n outside 1..32
  __shared__ int count[33];
  __shared__ int first_slot[33];
  const int c = threadIdx.x / 32 + 1;
  const int j = threadIdx.x % 32;
  const int64_t per = (nblocks + 31) / 32;
  const int64_t lo = min(nblocks, j * per);
  const int64_t hi = min(nblocks, lo + per);
  int sum = 0, sum0 = 0;
#pragma unroll 8
  for (int64_t b = lo; b < hi; ++b) {
    sum += bhist[b * RaggedWs::BSTRIDE + c];
    if (c == 1) sum0 += bhist[b * RaggedWs::BSTRIDE];
  }
  part[c - 1][j] = sum;
  if (c == 1) bad[j] = sum0;
  __syncthreads();
  int excl = 0;
  for (int t = 0; t < j; ++t) excl += part[c - 1][t];
  if (j == 31) count[c] = excl + sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int nbad = 0;
    for (int t = 0; t < 32; ++t) nbad += bad[t];
    ws[RaggedWs::META + 2] = nbad;
    int first = 0, tiles = 0;
    for (int n = 32; n >= 1; --n) {
      const int cnt = count[n];
      const int layout = (tune.layout[n] == 1 && n <= kRaggedQuadMaxN) ? 1 : 0;
      const int mpw = ragged_mpw(n, layout);
      const int per_desc = mpw * min(max(1, (int)tune.grab[n]), kRaggedMaxPer / mpw);
      ws[RaggedWs::COUNTS + n] = cnt;
      ws[RaggedWs::CLASS_FIRST + n] = first;
      ws[RaggedWs::DESC_FIRST + n] = tiles;
      ws[RaggedWs::LPER + n] = layout | (per_desc << 8);
      first_slot[n] = first;
      first += cnt;
      tiles += (cnt + per_desc - 1) / per_desc;
    }
    ws[RaggedWs::DESC_FIRST + 0] = tiles;
    ws[RaggedWs::CLASS_FIRST + 0] = first;
    ws[RaggedWs::LPER + 0] = 0;
  }
  __syncthreads();
  int slot = first_slot[c] + excl;
  for (int64_t b = lo; b < hi; ++b) {
    const int v = bhist[b * RaggedWs::BSTRIDE + c];
    bhist[b * RaggedWs::BSTRIDE + c] = slot;
    slot += v;
  }
}

// Slot of each matrix = its block's first slot in its class (from the scan)
// + a rank from an LDS atomic.  The order within a block is therefore not
// fixed; results do not depend on it.
__global__ __launch_bounds__(kPlanThreads) void ragged_scatter(const int* __restrict__ ns,
                                                               const int64_t* __restrict__ offs,
                                                               const int64_t* __restrict__ poffs, int64_t batch,
                                                               const int* __restrict__ bbase,
                                                               int64_t* __restrict__ soffs,
                                                               int64_t* __restrict__ spoffs) {
  __shared__ int cur[RaggedWs::BSTRIDE];
  if (threadIdx.x < RaggedWs::BSTRIDE) cur[threadIdx.x] = bbase[int64_t(blockIdx.x) * RaggedWs::BSTRIDE + threadIdx.x];
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
int64_t ragged_grid(const lub_ragged_plan* p, const lub::RaggedTuning& rt, int64_t batch) {
  static const int per_cu = [] {
    const char* e = std::getenv("LUB_RAGGED_WAVES_PER_CU");
    return e != nullptr ? std::atoi(e) : 0;
  }();
  if (per_cu > 0) return int64_t(p->cus) * per_cu;
  int min_per = lub::kRaggedMaxPer;
  for (int n = 1; n <= 32; ++n) min_per = std::min(min_per, ragged_per(rt, n));
  return std::min<int64_t>((batch + min_per - 1) / min_per + 32, int64_t(1) << 25);  // grid * 64 within 32 bits
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
  if (batch < 0 || batch > kMaxRaggedBatch) return LUB_EINVAL_BATCH;
  if ((flags & ~(LUB_PIVOT_MASK | LUB_LAYOUT_MASK)) != 0 || (flags & LUB_PIVOT_MASK) > LUB_PIVOT_GEPP ||
      (flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_MASK)
    return LUB_EFLAGS;
  if (batch == 0) return LUB_OK;
  if (A == nullptr || offsets == nullptr || ns == nullptr) return LUB_ENULL;
  if (piv != nullptr && piv_offsets == nullptr) return LUB_ENULL;
  if (batch > plan->capacity) return LUB_EINVAL_ARG;
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
  hipStream_t s = static_cast<hipStream_t>(stream);
  int* hdr = plan->header();
  const unsigned blocks = plan_blocks(batch);
  int* bhist = plan->bhist();
  lub::ragged_hist<<<blocks, lub::kPlanThreads, 0, s>>>(ns, batch, bhist);
  lub::ragged_scan<<<1, lub::kScanThreads, 0, s>>>(hdr, bhist, static_cast<int64_t>(blocks), rt);
  lub::ragged_scatter<<<blocks, lub::kPlanThreads, 0, s>>>(ns, offsets, piv ? piv_offsets : nullptr, batch, bhist,
                                                            plan->soffs(), plan->spoffs());
  int64_t* soffs = plan->soffs();
  int64_t* spoffs = piv ? plan->spoffs() : nullptr;
  const int* hdr_arg = hdr;
  void* args[] = {&A, &soffs, &piv, &spoffs, &hdr_arg};
  const int64_t grid = ragged_grid(plan, rt, batch);
  hipError_t e = hipLaunchKernel(fn, dim3(static_cast<unsigned>(grid)), dim3(64), args, 0, s);
  if (e != hipSuccess) return LUB_EHIP;
  return hipGetLastError() == hipSuccess ? LUB_OK : LUB_EHIP;
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

int lu_ragged_invalid_count(const lub_ragged_plan* plan, int* out) {
  if (plan == nullptr || out == nullptr) return LUB_ENULL;
  return hipMemcpy(out, plan->header() + lub::RaggedWs::META + 2, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess
             ? LUB_OK
             : LUB_EHIP;
}

}  // extern "C"
