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
//   1. the tile arrives in registers (TileRegs) and is staged into LDS; the
//      registers keep A through the factorisation;
//   2. tile_front + tile_solve leave X = A^-1 in the tile's LDS region;
//   3. per matrix, the whole wave: A_m goes from the registers to a 32 x 32
//      scratch (zero padded), D = A_m X_m is formed on the FP64 matrix cores
//      (v_mfma_f64_16x16x4f64: products of FP32 values are exact in FP64, the
//      sums round in FP64), R = I - D is rounded to FP32 into the scratch,
//      C = X_m R comes from FP32 matrix-core products (v_mfma_f32_32x32x2f32,
//      accumulated from zero so its roundings stay at the correction's
//      scale) and X_m + C (one rounding) goes to HBM.
// Padding: the scratch holds zeros outside A_m's n x n block, and X is read
// as zero outside its block, so D and C vanish there and the padding leaves
// the n x n result unchanged.
#pragma once

#include "lu_kernel.hpp"

namespace lub {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kRfPad = 32;  // matrices are padded to 32 x 32 for the products
constexpr int kRfLd = 36;   // row stride of the scratch (floats): rows stay 16-byte aligned

// The two matrix products are unrolled (every step's operands read ahead of
// the matrix-core instructions) where that fits 256 VGPRs beside the rest of
// the kernel: n <= 26 and n = 32; n = 27..31 run them rolled.
template <int N>
constexpr bool kRfUnroll = N <= 26 || N == 32;

// Sizes whose first second read of A is issued before the elimination (its
// registers fit beside the elimination's); the others issue it after.
template <int N>
constexpr bool kRfEarly = N <= 24 || N == 32;

// A_m in registers on its way to the scratch: lane l holds floats
// e = 4 (l + 64 k) .. +3 of the matrix (n % 4 == 0, 16-byte aligned batch)
// or floats e = l + 64 k (otherwise).
template <int N>
struct RefinedA {
  static constexpr bool VEC4 = N % 4 == 0;
  static constexpr int PER = VEC4 ? 4 : 1;
  static constexpr int K = (N * N / PER + kWave - 1) / kWave;
  float v[K][PER];
};

// Loads A_m (matrix at `a`) with ordinary loads: the tile's first read used
// them too, so this second read is normally served from the caches.
template <int N>
__device__ __forceinline__ void refined_load_a(const float* __restrict__ a, RefinedA<N>& r, bool vec, int lane) {
  using R = RefinedA<N>;
  constexpr int NN = N * N;
#pragma unroll
  for (int k = 0; k < R::K; ++k) {
    if constexpr (R::VEC4) {
      const int c = lane + kWave * k;
      if (c < NN / 4) {
        if (vec) {
          const f32x4 t = *reinterpret_cast<const f32x4*>(a + 4 * c);
          r.v[k][0] = t.x; r.v[k][1] = t.y; r.v[k][2] = t.z; r.v[k][3] = t.w;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) r.v[k][q] = a[4 * c + q];
        }
      }
    } else {
      const int e = lane + kWave * k;
      if (e < NN) r.v[k][0] = a[e];
    }
  }
}

// The registers of refined_load_a into the scratch (row stride kRfLd).
template <int N>
__device__ __forceinline__ void refined_stage_a(const RefinedA<N>& r, float* sr, int lane) {
  using R = RefinedA<N>;
  constexpr int NN = N * N;
#pragma unroll
  for (int k = 0; k < R::K; ++k) {
    if constexpr (R::VEC4) {
      const int c = lane + kWave * k;
      if (c < NN / 4) {
        const int i = (4 * c) / N, j = (4 * c) % N;  // a row holds whole chunks
        *reinterpret_cast<f32x4*>(sr + i * kRfLd + j) = f32x4{r.v[k][0], r.v[k][1], r.v[k][2], r.v[k][3]};
      }
    } else {
      const int e = lane + kWave * k;
      if (e < NN) sr[(e / N) * kRfLd + e % N] = r.v[k][0];
    }
  }
}

// One refinement step of matrix m: xm is X_m in the tile's LDS region (row
// stride Geo::LD), sr holds A_m; the refined X_m goes to xout (row-major,
// stride n).  Leaves R = I - A_m X_m in the scratch (zero off its block,
// one on its padding diagonal, which meets only zero rows of X).
template <int N>
__device__ __forceinline__ void refined_step(const float* xm, float* sr, float* __restrict__ xout, int lane) {
  constexpr int LD = Geo<N, 16>::LD;
  // D = A_m X_m on the FP64 matrix cores, four 16 x 16 blocks, k in steps of
  // 4: operand lane l holds A[16I + l%16][k0 + l/16] and X[k0 + l/16][16J + l%16];
  // accumulator element v of lane l is D[16I + l/16 + 4v][16J + l%16].
  const int r16 = lane & 15, q = lane >> 4;
  f64x4 acc[2][2];
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int J = 0; J < 2; ++J) acc[I][J] = f64x4{0.0, 0.0, 0.0, 0.0};
  auto dstep = [&](int k0) {
    const int k = k0 + q;
    double av[2], xv[2];
#pragma unroll
    for (int I = 0; I < 2; ++I) av[I] = static_cast<double>(sr[(16 * I + r16) * kRfLd + k]);
#pragma unroll
    for (int J = 0; J < 2; ++J) {
      const int j = 16 * J + r16;
      xv[J] = (k < N && j < N) ? static_cast<double>(xm[k * LD + j]) : 0.0;
    }
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int J = 0; J < 2; ++J) acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[I], xv[J], acc[I][J], 0, 0, 0);
  };
  if constexpr (kRfUnroll<N>) {
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += 4) dstep(k0);
  } else {
#pragma unroll 1
    for (int k0 = 0; k0 < N; k0 += 4) dstep(k0);
  }
  wave_sync();  // every read of A_m is done before R replaces it
#pragma unroll
  for (int I = 0; I < 2; ++I)
#pragma unroll
    for (int J = 0; J < 2; ++J)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * I + q + 4 * v, j = 16 * J + r16;
        sr[i * kRfLd + j] = static_cast<float>((i == j ? 1.0 : 0.0) - acc[I][J][v]);
      }
  wave_sync();

  // C = X_m R on the FP32 matrix cores (32 x 32 x 2): operand lane l holds
  // X[l%32][2s + l/32] and R[2s + l/32][l%32]; accumulator element v of lane
  // l is C[(v/4)*8 + (l/32)*4 + v%4][l%32].
  const int col = lane & 31, half = lane >> 5;
  f32x16 c;
#pragma unroll
  for (int v = 0; v < 16; ++v) c[v] = 0.0f;
  auto cstep = [&](int s) {
    const int k = 2 * s + half;
    const float xa = (col < N && k < N) ? xm[col * LD + k] : 0.0f;
    const float rb = sr[k * kRfLd + col];
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(xa, rb, c, 0, 0, 0);
  };
  if constexpr (kRfUnroll<N>) {
#pragma unroll
    for (int s = 0; s < (N + 1) / 2; ++s) cstep(s);
  } else {
#pragma unroll 1
    for (int s = 0; s < (N + 1) / 2; ++s) cstep(s);
  }
  if (col < N) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int i = (v >> 2) * 8 + half * 4 + (v & 3);
      if (i < N) __builtin_nontemporal_store(xm[i * LD + col] + c[v], xout + i * N + col);
    }
  }
  wave_sync();  // the scratch is rewritten for the next matrix
}

// Out-of-place inverse + one refinement step of `batch` n x n matrices
// (n = 17..32, 16 lanes per matrix): A is only read, X receives the refined
// inverse, piv (nullable) the pivot vectors.  One workgroup (one wave) per
// tile of four matrices, striding when the grid is capped.  A_m is read a
// second time for its refinement (A_{m+1}'s loads are issued before A_m's
// refinement starts); keeping the whole tile in registers through the
// factorisation instead exceeds 256 VGPRs and spills.
template <int N, int PIV>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void lu_refined_kernel(
    const float* __restrict__ A, float* __restrict__ X, int* __restrict__ piv, int64_t batch, int64_t ntiles,
    int vec) {
  static_assert(N >= 17 && N <= 32, "the fused refinement takes the 16-lane layout with two rows per lane");
  using G = Geo<N, 16>;
  constexpr int NN = N * N;
  __shared__ __attribute__((aligned(16))) float lds[G::LDS_FLOATS];
  __shared__ __attribute__((aligned(16))) float sr[kRfPad * kRfLd];
  const int lane = threadIdx.x;
  const int l = lane % 16, g = lane / 16;
  for (int e = lane; e < kRfPad * kRfLd; e += kWave) sr[e] = 0.0f;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t m0 = t * G::MPW;
    const int64_t left = batch - m0;
    const int valid = left < G::MPW ? static_cast<int>(left) : G::MPW;
    if (vec != 0 && valid == G::MPW) {
      TileRegs<N, 16> a;
      tile_fetch<N, 16, false>(A + m0 * NN, a, lane);
      tile_stage<N, 16>(a, lds, lane);
    } else {
      tile_load_elementwise<N, 16>(A + m0 * NN, lds, valid, lane);
    }
    wave_sync();
    int prm[G::R];
    RefinedA<N> am;
    {
      float w[G::R][N];
      tile_front<N, 16, PIV>(w, prm, lds, lane);
      // A_0's second read is in flight during the elimination
      if constexpr (kRfEarly<N>) refined_load_a<N>(A + m0 * NN, am, vec != 0, lane);
      tile_solve<N, 16>(w, prm, lds, lane);
    }
    if (piv != nullptr && g < valid) {
#pragma unroll
      for (int h = 0; h < G::R; ++h)
        if (l + 16 * h < N) piv[(m0 + g) * N + l + 16 * h] = prm[h];
    }
    if constexpr (!kRfEarly<N>) refined_load_a<N>(A + m0 * NN, am, vec != 0, lane);
#pragma unroll
    for (int m = 0; m < G::MPW; ++m) {
      if (m < valid) {
        refined_stage_a<N>(am, sr, lane);
        wave_sync();
        if (m + 1 < valid) refined_load_a<N>(A + (m0 + m + 1) * NN, am, vec != 0, lane);
        refined_step<N>(lds + m * G::MS, sr, X + (m0 + m) * NN, lane);
      }
    }
  }
}

}  // namespace lub
