// lu_refined.hpp -- inversion and one refinement step in one kernel:
//
//   X = A^-1 (the plain kernel's tile routines), then X <- X + X (I - A X)
//
// with A kept on chip, so the refined inverse moves the plain inverse's bytes
// (A read once, X written once, pivots) instead of a copy, an inversion and
// a separate refinement pass over A and X.  The reference has no refinement:
// its inverse is the FP32 LU result (parallel_pivot/luBatchedInplace.cuh:
// 84-125, stored at :194-197).
//
// Per tile of four matrices (one wave, 16 lanes per matrix, n = 17..32):
//   1. the tile arrives in registers (TileRegs) and is staged into LDS;
//   2. tile_front + tile_solve leave X = A^-1 in the tile's LDS region;
//   3. per matrix, the whole wave: A_m is read a second time (normally from
//      the caches) straight into the FP64 product's operand registers,
//      D = A_m X_m is formed on the FP64 matrix cores (v_mfma_f64_16x16x4f64:
//      products of FP32 values are exact in FP64, the sums round in FP64),
//      E = D - I (= -R) is rounded to FP32 and stays in the registers it was
//      accumulated in, which are exactly the operands of C = X_m E on the
//      FP32 matrix cores (v_mfma_f32_16x16x4f32, accumulated from zero so its
//      roundings stay at the correction's scale), and X_m - C (one rounding)
//      replaces X_m in LDS;
//   4. the tile goes to HBM with the plain kernel's coalesced tile store.
// Nothing but the tile itself lives in LDS, so the kernel keeps the plain
// kernel's occupancy.  Padding: A_m's registers hold zeros outside its n x n
// block and X is read as zero outside its block, so D and C vanish there and
// the padding leaves the n x n result unchanged.
#pragma once

#include "lu_kernel.hpp"

namespace lub {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Sizes whose first second read of A is issued before the elimination (its
// registers fit beside the elimination's); the others issue it after.
template <int N>
constexpr bool kRfEarly = N <= 24 || N == 32;

// Sizes whose next matrix's A is loaded before the current matrix's step
// (registers permitting); the others load it after.
template <int N>
constexpr bool kRfAhead = N <= 28 || N == 32;

// Sizes that fetch the workgroup's next tile into registers while the
// current tile is refined (registers permitting).
template <int N>
constexpr bool kRfNextTile = N == 32;

// A_m in the FP64 product's operand layout: lane (r = lane % 16,
// q = lane / 16) holds A_m[16 I + r][8 q + t], I = 0, 1, t = 0..7, and zero
// outside the n x n block.
struct RefinedA {
  float v[2][8];
};

// Loads A_m (matrix at `a`) with ordinary loads: the tile's first read used
// them too, so this second read is normally served from the caches.  `vec`:
// the batch is 16-byte aligned (n % 4 == 0 then reads 16-byte chunks).
template <int N>
__device__ __forceinline__ void refined_load_a(const float* __restrict__ a, RefinedA& r, bool vec, int lane) {
  const int r16 = lane & 15, q = lane >> 4;
#pragma unroll
  for (int I = 0; I < 2; ++I) {
    const int i = 16 * I + r16;
    if constexpr (N % 4 == 0) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = 8 * q + 4 * h;  // n % 4 == 0: a chunk lies wholly inside or outside the block
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        if (N == 32 || (i < N && j < N)) {
          if (vec) {
            c = *reinterpret_cast<const f32x4*>(a + i * N + j);
          } else {
            c = f32x4{a[i * N + j], a[i * N + j + 1], a[i * N + j + 2], a[i * N + j + 3]};
          }
        }
        r.v[I][4 * h] = c.x; r.v[I][4 * h + 1] = c.y; r.v[I][4 * h + 2] = c.z; r.v[I][4 * h + 3] = c.w;
      }
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = 8 * q + t;
        r.v[I][t] = (i < N && j < N) ? a[i * N + j] : 0.0f;
      }
    }
  }
}

// X_m[i][j] (i, j < 32) from the tile's LDS region, zero outside the n x n
// block.  No branch: the read goes to an address inside the block either way
// and a select supplies the zero, so consecutive reads stay in flight together.
template <int N>
__device__ __forceinline__ float x_at(const float* xm, int i, int j) {
  constexpr int LD = Geo<N, 16>::LD;
  if constexpr (N == 32) {
    return xm[i * LD + j];
  } else {
    const bool in = i < N && j < N;
    const float v = xm[(in ? i : 0) * LD + (in ? j : 0)];
    return in ? v : 0.0f;
  }
}

// One refinement step of matrix m: `am` holds A_m (refined_load_a), xm is
// X_m in the tile's LDS region (row stride Geo::LD), where the refined X_m
// replaces it (every read of X_m precedes the first write).
template <int N>
__device__ __forceinline__ void refined_step(const RefinedA& am, float* xm, int lane) {
  constexpr int LD = Geo<N, 16>::LD;
  const int r16 = lane & 15, q = lane >> 4;
  // D = A_m X_m on the FP64 matrix cores, four 16 x 16 blocks (I, J).  In
  // v_mfma_f64_16x16x4f64 operand lane l holds A[l % 16][kk] and X[kk][l % 16]
  // with kk = l / 16; step t lets lane group q stand for k = 8 q + t, so A_m's
  // registers are used as loaded.  Accumulator element v of lane l is
  // D[16 I + l / 16 + 4 v][16 J + l % 16].
  f64x4 acc[2][2];
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int J = 0; J < 2; ++J) acc[I][J] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int k = 8 * q + t;
    double xv[2];
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      xv[J] = static_cast<double>(x_at<N>(xm, k, 16 * J + r16));
    }
#pragma unroll
    for (int I = 0; I < 2; ++I) {
      const double av = static_cast<double>(am.v[I][t]);
#pragma unroll
      for (int J = 0; J < 2; ++J) acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xv[J], acc[I][J], 0, 0, 0);
    }
  }
  // E = D - I rounded to FP32 (E = -R; rounding to nearest is symmetric, so
  // X - X E below equals X + X R bit for bit): ee[I][J][v] = E[16 I + q + 4 v][16 J + r16].
  // Only the diagonal blocks hold ones to subtract (in FP64, before rounding);
  // on the padding diagonal E is -1, which meets only zero columns of X.
  float ee[2][2][4];
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int J = 0; J < 2; ++J)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (I == J) {
          ee[I][J][v] = static_cast<float>(acc[I][J][v] - (q + 4 * v == r16 ? 1.0 : 0.0));
        } else {
          ee[I][J][v] = static_cast<float>(acc[I][J][v]);
        }
      }
  // C = X_m E on the FP32 matrix cores, four 16 x 16 blocks (Io, J).  In
  // v_mfma_f32_16x16x4f32 operand lane l holds X[l % 16][kk] and E[kk][l % 16]
  // with kk = l / 16; step (I, v) lets lane group q stand for
  // k = 16 I + q + 4 v, the row of E that lane already holds.  Accumulator
  // element v of lane l is C[4 (l / 16) + v][l % 16].
  f32x4 c[2][2];
#pragma unroll
  for (int Io = 0; Io < 2; ++Io)
#pragma unroll
    for (int J = 0; J < 2; ++J) c[Io][J] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = 16 * I + q + 4 * v;
      float xa[2];
#pragma unroll
      for (int Io = 0; Io < 2; ++Io) xa[Io] = x_at<N>(xm, 16 * Io + r16, k);
#pragma unroll
      for (int Io = 0; Io < 2; ++Io)
#pragma unroll
        for (int J = 0; J < 2; ++J) c[Io][J] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[Io], ee[I][J][v], c[Io][J], 0, 0, 0);
    }
#pragma unroll
  for (int Io = 0; Io < 2; ++Io)
#pragma unroll
    for (int J = 0; J < 2; ++J)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * Io + 4 * q + v, j = 16 * J + r16;
        if (N == 32 || (i < N && j < N)) xm[i * LD + j] -= c[Io][J][v];
      }
}

// Out-of-place inverse + one refinement step of `batch` n x n matrices
// (n = 17..32, 16 lanes per matrix): A is only read, X receives the refined
// inverse, piv (nullable) the pivot vectors.  `vec`: bit 0, A is 16-byte
// aligned; bit 1, X is.  One workgroup (one wave) per tile of four matrices,
// striding when the grid is capped; a striding workgroup at n = 32 fetches
// its next tile into registers during the current tile's refinement.  A_m is
// read a second time for its refinement (A_{m+1}'s loads are issued before
// A_m's refinement starts where registers allow); keeping the whole tile in
// registers through the factorisation instead exceeds 256 VGPRs and spills.
template <int N, int PIV>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void lu_refined_kernel(
    const float* __restrict__ A, float* __restrict__ X, int* __restrict__ piv, int64_t batch, int64_t ntiles,
    int vec) {
  static_assert(N >= 17 && N <= 32, "the fused refinement takes the 16-lane layout with two rows per lane");
  using G = Geo<N, 16>;
  constexpr int NN = N * N;
  __shared__ __attribute__((aligned(16))) float lds[G::LDS_FLOATS];
  const int lane = threadIdx.x;
  const int l = lane % 16, g = lane / 16;
  const int64_t stride = gridDim.x;
  auto full = [&](int64_t t) { return (vec & 1) != 0 && (t + 1) * G::MPW <= batch; };
  TileRegs<N, 16> next;
  int64_t t = blockIdx.x;
  bool have = t < ntiles && full(t);
  if (have) tile_fetch<N, 16, false>(A + t * G::MPW * NN, next, lane);
  for (; t < ntiles; t += stride) {
    const int64_t m0 = t * G::MPW;
    const int64_t left = batch - m0;
    const int valid = left < G::MPW ? static_cast<int>(left) : G::MPW;
    if (have) {
      tile_stage<N, 16>(next, lds, lane);
    } else {
      tile_load_elementwise<N, 16>(A + m0 * NN, lds, valid, lane);
    }
    wave_sync();
    int prm[G::R];
    RefinedA am;
    {
      float w[G::R][N];
      tile_front<N, 16, PIV>(w, prm, lds, lane);
      // A_0's second read is in flight during the elimination
      if constexpr (kRfEarly<N>) refined_load_a<N>(A + m0 * NN, am, (vec & 1) != 0, lane);
      tile_solve<N, 16>(w, prm, lds, lane);
    }
    if (piv != nullptr && g < valid) {
#pragma unroll
      for (int h = 0; h < G::R; ++h)
        if (l + 16 * h < N) piv[(m0 + g) * N + l + 16 * h] = prm[h];
    }
    // the next tile's HBM reads overlap this tile's refinement
    const int64_t tn = t + stride;
    have = tn < ntiles && full(tn);
    if constexpr (kRfNextTile<N>) {
      if (have) tile_fetch<N, 16, false>(A + tn * G::MPW * NN, next, lane);
    }
    if constexpr (!kRfEarly<N>) refined_load_a<N>(A + m0 * NN, am, (vec & 1) != 0, lane);
#pragma unroll
    for (int m = 0; m < G::MPW; ++m) {
      if (m < valid) {
        if constexpr (kRfAhead<N>) {
          RefinedA an;
          if (m + 1 < valid) refined_load_a<N>(A + (m0 + m + 1) * NN, an, (vec & 1) != 0, lane);
          refined_step<N>(am, lds + m * G::MS, lane);
          am = an;
        } else {
          refined_step<N>(am, lds + m * G::MS, lane);
          if (m + 1 < valid) refined_load_a<N>(A + (m0 + m + 1) * NN, am, (vec & 1) != 0, lane);
        }
      }
    }
    tile_store<N, 16>(X + m0 * NN, lds, valid, (vec & 2) != 0, lane);
    wave_sync();  // the tile's region is rewritten by the next tile
    if constexpr (!kRfNextTile<N>) {
      if (have) tile_fetch<N, 16, false>(A + tn * G::MPW * NN, next, lane);
    }
  }
}

}  // namespace lub
