// lub_lapack.hip -- LAPACK-style outputs of the plain path (include/lubatch.h).
//
// The library's own pivot output is the permutation perm of P*A (row i of
// P*A is row perm[i] of A, the reference's `pivots` array,
// parallel_pivot/luBatchedInplace.cuh:145-148, which it never exports).
// LAPACK's getrf/getri describe the same P as a sequence of interchanges:
// at step k rows k and ipiv[k] - 1 were swapped (1-based), and info = k + 1
// flags the first exactly zero pivot U(k,k).  The interchanges here are those
// of the library's pivot rule (the reference's first-max on the un-updated
// column), not LAPACK's partial pivoting, so they may differ from getrf's
// choices; applied with laswp they give the P of this factorisation.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lub_internal.h"
#include "lubatch.h"

namespace {

constexpr int kThreads = 64;

// One thread per matrix: replay the permutation as interchanges.  cur[p] is
// the original row at position p, where[r] the position of original row r;
// step k swaps positions k and where[perm[k]].  perm[k] is read before
// ipiv[k] overwrites it, and later steps read only perm[k' > k].
__global__ __launch_bounds__(kThreads) void perm_to_ipiv_kernel(int* __restrict__ piv, int n, int64_t batch) {
  __shared__ int cur_s[kThreads][LUB_MAX_N + 1];
  __shared__ int where_s[kThreads][LUB_MAX_N + 1];
  int* cur = cur_s[threadIdx.x];
  int* where = where_s[threadIdx.x];
  for (int64_t m = int64_t(blockIdx.x) * kThreads + threadIdx.x; m < batch; m += int64_t(gridDim.x) * kThreads) {
    int* p = piv + m * n;
    for (int i = 0; i < n; ++i) {
      cur[i] = i;
      where[i] = i;
    }
    for (int k = 0; k < n; ++k) {
      int r = p[k];
      r = r < 0 ? 0 : (r >= n ? n - 1 : r);  // a permutation by construction; clamp keeps LDS in range regardless
      const int pos = where[r];
      const int rk = cur[k];
      cur[pos] = rk;
      where[rk] = pos;
      cur[k] = r;
      where[r] = k;
      p[k] = pos + 1;
    }
  }
}

int enqueue_convert(int* piv, int n, int64_t batch, hipStream_t s) {
  const int64_t blocks = (batch + kThreads - 1) / kThreads;
  const unsigned grid = static_cast<unsigned>(blocks < (int64_t(1) << 20) ? blocks : (int64_t(1) << 20));
  perm_to_ipiv_kernel<<<grid, kThreads, 0, s>>>(piv, n, batch);
  return hipGetLastError() == hipSuccess ? LUB_OK : LUB_EHIP;
}

}  // namespace

extern "C" {

int lu_perm_to_ipiv(int* piv, int n, int64_t batch, void* stream) {
  if (n < 1 || n > LUB_MAX_N) return LUB_EINVAL_N;
  if (batch < 0) return LUB_EINVAL_BATCH;
  if (batch == 0) return LUB_OK;
  if (piv == nullptr) return LUB_ENULL;
  return enqueue_convert(piv, n, batch, static_cast<hipStream_t>(stream));
}

int lu_batched_inverse_lapack(float* A, int* ipiv, int* info, int n, int64_t batch, int flags, void* stream) {
  if (batch > 0 && (ipiv == nullptr || info == nullptr)) return LUB_ENULL;
  int rc = lub::enqueue_plain(A, ipiv, info, n, batch, flags, stream);
  if (rc != LUB_OK || batch == 0) return rc;
  return enqueue_convert(ipiv, n, batch, static_cast<hipStream_t>(stream));
}

}  // extern "C"
