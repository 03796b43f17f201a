// lub_refine.hip -- one step of iterative refinement of batched inverses,
//   X <- X + X (I - A X),
// with the residual I - A X accumulated in FP64 (products of two FP32 values
// are exact in FP64, so only the n-term sums round) and the correction
// X (I - A X) formed by FP32 MFMA (v_mfma_f32_32x32x2f32) and added to X
// with one rounding.  The reference has no refinement: its inverse is the
// FP32 LU result itself (parallel_pivot/luBatchedInplace.cuh:84-125), and
// its check only counts matrices within 1e-3 of I (verify.hpp:50-103).
//
// One wave per matrix, padded to 32 x 32 with zeros: with A and X zero
// outside their n x n block, I - A X is the identity there and X (I - A X)
// is zero there, so the padding leaves the n x n result unchanged.
//
// LDS per wave: A and X row-major (row stride 36 floats, so rows stay
// 16-byte aligned), R row-major for the MFMA B operand.  The FP64 residual
// block of lane (bi, bj) reads A[4bi + r][k..k+3] and X[k + t][4bj..4bj+3] as
// 16-byte reads, eight per four values of k; each element still sums its n
// products in k order (the padding adds exact zeros).  A full, aligned
// n = 32 matrix arrives as 16-byte global loads and needs no zero fill.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "lub_internal.h"
#include "lubatch.h"

namespace {

constexpr int kPad = 32;
constexpr int kLd = 36;  // row stride in floats: rows stay 16-byte aligned

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64) void refine_kernel(const float* __restrict__ A, float* __restrict__ X, int n,
                                                    int64_t batch) {
  __shared__ __attribute__((aligned(16))) float sA[kPad * kLd];
  __shared__ __attribute__((aligned(16))) float sX[kPad * kLd];
  __shared__ __attribute__((aligned(16))) float sR[kPad * kLd];
  const int lane = threadIdx.x;
  const int nn = n * n;
  const bool whole = n == kPad && ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(X)) & 15) == 0;
  for (int64_t m = blockIdx.x; m < batch; m += gridDim.x) {
    const float* a = A + m * nn;
    float* x = X + m * nn;
    if (whole) {
      // chunk c = lane + 64t: floats 4c..4c+3, row c / 8, columns 4 (c % 8) ..
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int c = lane + 64 * t, i = c >> 3, j = 4 * (c & 7);
        const f32x4 av = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a) + c);
        const f32x4 xv = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(x) + c);
        *reinterpret_cast<f32x4*>(sA + i * kLd + j) = av;
        *reinterpret_cast<f32x4*>(sX + i * kLd + j) = xv;
      }
    } else {
      for (int e = lane; e < kPad * kLd; e += 64) {
        sA[e] = 0.0f;
        sX[e] = 0.0f;
      }
      __syncthreads();
      for (int e = lane; e < nn; e += 64) {
        const int i = e / n, j = e - i * n;
        sA[i * kLd + j] = a[e];
        sX[i * kLd + j] = x[e];
      }
    }
    __syncthreads();

    // R = I - A X in FP64: lane (bi, bj) owns rows 4bi..4bi+3, columns 4bj..4bj+3
    const int bi = lane >> 3, bj = lane & 7;
    double acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
    const int n4 = (n + 3) & ~3;  // columns past n are zero padding
    for (int k0 = 0; k0 < n4; k0 += 4) {
      f32x4 av[4], xv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = *reinterpret_cast<const f32x4*>(sA + (4 * bi + r) * kLd + k0);
#pragma unroll
      for (int t = 0; t < 4; ++t) xv[t] = *reinterpret_cast<const f32x4*>(sX + (k0 + t) * kLd + 4 * bj);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = __fma_rn(double(av[r][t]), double(xv[t][c]), acc[r][c]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      f32x4 rv;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = 4 * bi + r, j = 4 * bj + c;
        rv[c] = static_cast<float>((i == j ? 1.0 : 0.0) - acc[r][c]);
      }
      *reinterpret_cast<f32x4*>(sR + (4 * bi + r) * kLd + 4 * bj) = rv;
    }
    __syncthreads();

    // C = X R on the matrix cores, accumulated from zero so that its
    // roundings are at the scale of the (small) correction; then X + C with
    // one rounding.  32x32x2 layout: A operand lane l holds
    // X[l % 32][2s + l / 32], B operand R[2s + l / 32][l % 32]; accumulator
    // element v of lane l is C[(v / 4) * 8 + (l / 32) * 4 + v % 4][l % 32].
    const int col = lane & 31, half = lane >> 5;
    f32x16 c;
#pragma unroll
    for (int v = 0; v < 16; ++v) c[v] = 0.0f;
#pragma unroll
    for (int s = 0; s < kPad / 2; ++s) {
      const float av = sX[col * kLd + 2 * s + half];
      const float bv = sR[(2 * s + half) * kLd + col];
      c = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, c, 0, 0, 0);
    }
    if (col < n) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int i = (v >> 2) * 8 + half * 4 + (v & 3);
        if (i < n) x[i * n + col] = sX[i * kLd + col] + c[v];
      }
    }
    __syncthreads();  // LDS is reused by the next matrix
  }
}

int check_args(const float* A, float* X, int n, int64_t batch) {
  if (n < 1 || n > LUB_MAX_N) return LUB_EINVAL_N;
  if (batch < 0) return LUB_EINVAL_BATCH;
  if (batch == 0) return LUB_OK;
  if (A == nullptr || X == nullptr) return LUB_ENULL;
  if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(X)) & 3) return LUB_EALIGN;
  return -1;  // proceed
}

int enqueue_refine(const float* A, float* X, int n, int64_t batch, hipStream_t s) {
  const int64_t grid = batch < (int64_t(1) << 22) ? batch : (int64_t(1) << 22);
  refine_kernel<<<dim3(static_cast<unsigned>(grid)), dim3(64), 0, s>>>(A, X, n, batch);
  return hipGetLastError() == hipSuccess ? LUB_OK : LUB_EHIP;
}

}  // namespace

extern "C" {

int lu_batched_refine(const float* A, float* X, int n, int64_t batch, void* stream) {
  const int rc = check_args(A, X, n, batch);
  if (rc >= 0) return rc;
  return enqueue_refine(A, X, n, batch, static_cast<hipStream_t>(stream));
}

// The fused kernel (lu_refined.hpp) serves n = 17..32 with no pivoting or
// the first-max rule in the 16-lane layout; other sizes and rules run the
// three steps: copy, plain inverse in place, refinement pass.
int lu_batched_inverse_refined(const float* A, float* X, int* piv, int n, int64_t batch, int flags, void* stream) {
  const int rc = check_args(A, X, n, batch);
  if (rc >= 0) return rc;
  if (A == X) return LUB_EINVAL_ARG;  // A must survive for the residual
  if ((flags & ~(LUB_PIVOT_MASK | LUB_LAYOUT_MASK)) != 0 || (flags & LUB_PIVOT_MASK) > LUB_PIVOT_GEPP ||
      (flags & LUB_LAYOUT_MASK) == LUB_LAYOUT_MASK)
    return LUB_EFLAGS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int rule = flags & LUB_PIVOT_MASK;
  const bool fused = n >= 17 && (rule == LUB_PIVOT_NONE || rule == LUB_PIVOT_PARALLEL || rule == LUB_PIVOT_SERIAL) &&
                     (flags & LUB_LAYOUT_MASK) != LUB_LAYOUT_QUAD4;
  if (fused) {
    const void* fn = lub::kernel_table().refined[n][rule == LUB_PIVOT_NONE ? 0 : 1];
    if (fn == nullptr) return LUB_ENODEVICE;
    int64_t ntiles = (batch + 3) / 4;
    // one workgroup per tile (several tiles per workgroup measured slower)
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(ntiles, int64_t(1) << 22));
    int vec = ((reinterpret_cast<uintptr_t>(A) & 15u) == 0 ? 1 : 0) | ((reinterpret_cast<uintptr_t>(X) & 15u) == 0 ? 2 : 0);
    void* args[] = {const_cast<float**>(&A), &X, &piv, &batch, &ntiles, &vec};
    if (hipLaunchKernel(fn, dim3(grid), dim3(64), args, 0, s) != hipSuccess) return LUB_EHIP;
    return LUB_OK;
  }
  if (hipMemcpyAsync(X, A, size_t(batch) * n * n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return LUB_EHIP;
  int r = lu_batched_inverse_async64(X, piv, n, batch, flags, stream);
  if (r != LUB_OK) return r;
  return enqueue_refine(A, X, n, batch, s);
}

}  // extern "C"
