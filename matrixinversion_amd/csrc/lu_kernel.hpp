// lu_kernel.hpp -- gfx950 batched in-place inversion of small FP32 matrices.
//
// A wave64 owns a tile of consecutive matrices; each matrix is owned by a
// group of L lanes -- a 16-lane DPP row (L = 16, 4 matrices per wave, any n
// up to 32) or a quad (L = 4, 16 matrices per wave, n up to 16) -- so every
// cross-lane broadcast a step needs is one DPP modifier (row_newbcast or
// quad_perm) and no LDS barrier is ever shared between waves.  Per tile:
//
//   1. tile load     HBM -> LDS with coalesced dwordx4 loads (the tile's
//                    matrices are contiguous)
//   2. row read      lane l of a matrix holds original rows l, l+L, ...
//   3. pivot search  REF rule (parallel_pivot/luBatchedInplace.cuh:12-44 reads
//                    column k of the row-permuted *input*; comp_U/comp_L at
//                    :175-185 have not touched it yet), so the permutation is a
//                    function of the input alone and is found before any
//                    arithmetic:
//                      PARALLEL  DPP max of |a| bit patterns per column; ties
//                                and NaN detected once per tile and resolved
//                                by the SERIAL scan for the wave
//                      SERIAL    serial_pivot/'s first-max scan over positions
//                                (strict >, serial_pivot/luBatchedInplace.cuh:
//                                22-36): the same pivots as PARALLEL, which runs
//                                it as its exact fallback; both flags share
//                                one kernel
//                      TREE      parallel_pivot's shared-memory tree, exactly:
//                                the DPP max over the positions whose tree
//                                slots reach slot 0; ties and NaN rerun the
//                                tree itself for the wave
//   4. permute       rows written back to LDS at their pivot positions: P*A
//   5. column read   lane q holds columns q, q+L, ... of P*A
//   6. Gauss-Jordan  row operations on P*A, pivot-free (the permutation
//                    already fixed the pivots, so the multipliers are those of
//                    the reference's Doolittle LU of P*A), in registers, one
//                    v_fmac_f32_dpp per element update
//   7. unpermute     column p of (P*A)^-1 is column perm[p] of A^-1 (the
//                    reference's comp_inv folds P into the right-hand side the
//                    same way, parallel_pivot/luBatchedInplace.cuh:99); it is
//                    written as an LDS column with one runtime base per lane
//   8. tile store    LDS -> HBM with coalesced dwordx4 stores, pivots as int32
//
// The LU-based inverse of the reference (Doolittle + two triangular solves)
// and this Gauss-Jordan both perform no-pivot elimination on P*A with the
// same pivots and multipliers; they differ in rounding order only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gj_asm.inc"

namespace lub {

constexpr int kWave = 64;



typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

enum Pivot : int { kPivNone = 0, kPivParallel = 1, kPivSerial = 2, kPivTree = 3, kPivGepp = 4 };

// Tile geometry for n = N with L lanes per matrix (all sizes in floats).
// V is a bit set: bit 0 selects the gathered layout (below), bit 1 flips the
// row padding from its default.
// Row padding (n % 4 == 0 only): a padded row is n + 4 floats long, so that
// consecutive rows start in different 16-byte bank groups.  By default rows
// are padded when n % 8 == 0 (otherwise 8 consecutive rows would start in
// the same bank group); the tuning table can flip that per n
// (lu_tuning_set_lds_pad), which changes the LDS addresses only.
// For n % 4 == 0 consecutive matrices start L banks apart so the lane groups
// of one half-wave column access hit disjoint banks.  For n % 4 != 0 a
// matrix is stored contiguously (row stride n): in the plain layout the
// whole tile is contiguous, as in HBM, so the tile's dwordx4 chunks (which
// cross row and matrix boundaries) stage unchanged; in the gathered layout
// (V & 1, the ragged kernel, whose matrices arrive one by one) each matrix
// starts on a 16-byte boundary so that its own chunks stage as aligned
// dwordx4.
template <int N, int L, int V = 0>
struct Geo {
  static_assert(L == 16 || L == 4, "a matrix is owned by a DPP row or a quad");
  static_assert(N >= 1 && N <= (L == 16 ? 32 : 16), "n too large for this lane grouping");
  static constexpr bool GATHER = (V & 1) != 0;
  static constexpr bool PAD = (N % 4 == 0) && ((N % 8 == 0) != ((V & 2) != 0));
  static constexpr int MPW = kWave / L;          // matrices per wave (tile)
  static constexpr int R = (N + L - 1) / L;      // rows (columns) per lane
  static constexpr int LD = PAD ? N + 4 : N;
  static constexpr int MS = (N % 4 == 0) ? ((N * LD + 31) / 32) * 32 + L : (GATHER ? (N * N + 3) / 4 * 4 : N * N);
  static constexpr int TILE = MPW * MS;
  static constexpr int PERM = TILE;               // int32[MPW][32] after the tile
  static constexpr int LDS_FLOATS = TILE + MPW * 32;
};

// The padding variant that flips the default (exists for n % 4 == 0).
constexpr int kFlipPad = 2;

// Per-lane register blocks of a matrix: R columns (or rows) of n floats, and
// one int / float per column.
template <int N, int L>
using ColBlock = float[Geo<N, L>::R][N];
template <int N, int L>
using LaneInts = int[Geo<N, L>::R];
template <int N, int L>
using LaneFloats = float[Geo<N, L>::R];

// ------------------------------------------------------------------ helpers

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xf, 0xf, true);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, true));
}

// DPP control that broadcasts lane KL of each lane group to the whole group.
template <int L, int KL>
constexpr int bcast_ctrl() {
  return L == 16 ? 0x150 + KL   // row_newbcast:KL
                 : KL * 0x55;   // quad_perm:[KL,KL,KL,KL]
}

// max / or over the L lanes of a group (quad xor 1, quad xor 2, then for
// 16-lane rows half mirror and mirror).
template <int L>
__device__ __forceinline__ uint32_t group_max_u32(uint32_t m) {
  m = max(m, dpp_u32<0xB1>(m));  // quad_perm [1,0,3,2]
  m = max(m, dpp_u32<0x4E>(m));  // quad_perm [2,3,0,1]
  if constexpr (L == 16) {
    m = max(m, dpp_u32<0x141>(m));  // row_half_mirror
    m = max(m, dpp_u32<0x140>(m));  // row_mirror
  }
  return m;
}

template <int L>
__device__ __forceinline__ uint32_t group_or_u32(uint32_t m) {
  m |= dpp_u32<0xB1>(m);
  m |= dpp_u32<0x4E>(m);
  if constexpr (L == 16) {
    m |= dpp_u32<0x141>(m);
    m |= dpp_u32<0x140>(m);
  }
  return m;
}

// key = live ? bits(|x|) : 0 in one v_cndmask with the |x| source modifier.
// Written as one instruction so that |x| is not computed ahead for every
// column (it depends on the lane mask of still-unplaced rows of this step).
__device__ __forceinline__ uint32_t abs_key(float x, bool live) {
  uint32_t k;
  asm("v_cndmask_b32_e64 %0, 0, |%1|, %2" : "=v"(k) : "v"(x), "s"(__builtin_amdgcn_ballot_w64(live)));
  return k;
}

// Materialises v in a VGPR at this point of the program: keeps a running
// reduction from being re-associated into a tree over all its inputs (which
// would keep every step's value live until the end).
__device__ __forceinline__ void keep_in_vgpr(uint32_t& v) { asm volatile("" : "+v"(v)); }

// Orders this wave's LDS accesses (the tile region is private to the wave;
// the workgroup is one wave, so the barrier is free).
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

template <int N>
__device__ __forceinline__ void lds_row_read(const float* p, float (&w)[N]) {
  if constexpr (N % 4 == 0) {
#pragma unroll
    for (int j = 0; j < N; j += 4) {
      f32x4 v = *reinterpret_cast<const f32x4*>(p + j);
      w[j] = v.x; w[j + 1] = v.y; w[j + 2] = v.z; w[j + 3] = v.w;
    }
  } else if constexpr (N % 2 == 0) {
#pragma unroll
    for (int j = 0; j < N; j += 2) {
      f32x2 v = *reinterpret_cast<const f32x2*>(p + j);
      w[j] = v.x; w[j + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) w[j] = p[j];
  }
}

template <int N>
__device__ __forceinline__ void lds_row_write(float* p, const float (&w)[N]) {
  if constexpr (N % 4 == 0) {
#pragma unroll
    for (int j = 0; j < N; j += 4)
      *reinterpret_cast<f32x4*>(p + j) = f32x4{w[j], w[j + 1], w[j + 2], w[j + 3]};
  } else if constexpr (N % 2 == 0) {
#pragma unroll
    for (int j = 0; j < N; j += 2) *reinterpret_cast<f32x2*>(p + j) = f32x2{w[j], w[j + 1]};
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) p[j] = w[j];
  }
}

// ------------------------------------------------------------- tile I/O

// HBM -> LDS.  Full tiles of a 16-byte aligned batch move as dwordx4 chunks
// (chunk c = floats 4c..4c+3 of the tile), fetched into registers (TileRegs)
// and staged into LDS.  The last, partial tile and unaligned batches move
// element-wise, and matrices past the end of the batch are filled with the
// identity so every lane computes on finite data.
template <int N, int L>
struct TileRegs {
  static constexpr int CH = Geo<N, L>::MPW * N * N / 4;  // dwordx4 chunks per tile
  static constexpr int IT = (CH + kWave - 1) / kWave;
  f32x4 v[IT];
};

// LDS offset (floats) of dwordx4 chunk c = lane + 64*it of the tile.  When
// one wave iteration (256 floats) covers whole rows and whole matrices (or a
// whole number of iterations covers one matrix), it splits into a per-lane
// base and a compile-time offset per iteration, so the staging loop needs one
// address register instead of one per chunk.
template <int N>
constexpr bool kChunkSplits = (N % 4 == 0) && (256 % N == 0) && ((N * N) % 256 == 0 || 256 % (N * N) == 0);

template <int N, int L, int V = 0>
__device__ __forceinline__ int chunk_lds_base(int lane) {
  using G = Geo<N, L, V>;
  constexpr int NN = N * N;
  if constexpr (kChunkSplits<N>) {
    const int f = 4 * lane;  // within the first iteration
    const int m = f / NN, e = f - m * NN, i = e / N, j = e - i * N;
    return m * G::MS + i * G::LD + j;
  } else {
    return 0;
  }
}

template <int N, int L, int V = 0>
__device__ __forceinline__ int chunk_lds_offset(int base, int lane, int it) {
  using G = Geo<N, L, V>;
  constexpr int NN = N * N;
  if constexpr (kChunkSplits<N>) {
    const int f0 = 256 * it;  // iteration it adds 256 floats: whole rows / matrices
    const int m = f0 / NN, e = f0 - m * NN, i = e / N;
    return base + m * G::MS + i * G::LD;  // compile-time after unrolling
  } else {
    const int f = 4 * (lane + kWave * it);
    const int m = f / NN, e = f - m * NN, i = e / N, j = e - i * N;
    return m * G::MS + i * G::LD + j;
  }
}

// NT: non-temporal loads (the data is read once); ordinary loads otherwise.
template <int N, int L, bool NT = true>
__device__ __forceinline__ void tile_fetch(const float* __restrict__ src, TileRegs<N, L>& r, int lane) {
  using T = TileRegs<N, L>;
#pragma unroll
  for (int it = 0; it < T::IT; ++it) {
    const int c = lane + kWave * it;
    if (T::CH % kWave == 0 || c < T::CH) {
      if constexpr (NT) r.v[it] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(src) + c);
      else r.v[it] = reinterpret_cast<const f32x4*>(src)[c];
    }
  }
}

template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_stage(const TileRegs<N, L>& r, float* lds, int lane) {
  using T = TileRegs<N, L>;
  const int base = chunk_lds_base<N, L, V>(lane);
#pragma unroll
  for (int it = 0; it < T::IT; ++it) {
    const int c = lane + kWave * it;
    if (T::CH % kWave == 0 || c < T::CH)
      *reinterpret_cast<f32x4*>(lds + chunk_lds_offset<N, L, V>(base, lane, it)) = r.v[it];
  }
}

template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_load_elementwise(const float* __restrict__ src, float* lds, int valid, int lane) {
  using G = Geo<N, L, V>;
  constexpr int NN = N * N;
  for (int f = lane; f < G::MPW * NN; f += kWave) {
    const int m = f / NN, e = f - m * NN, i = e / N, j = e - i * N;
    lds[m * G::MS + i * G::LD + j] = (m < valid) ? src[f] : (i == j ? 1.0f : 0.0f);
  }
}

// LDS -> HBM, the mirror of the load; only the `valid` leading matrices are stored.
template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_store(float* __restrict__ dst, const float* lds, int valid, bool vec, int lane) {
  using G = Geo<N, L, V>;
  using T = TileRegs<N, L>;
  constexpr int NN = N * N;
  if (vec && valid == G::MPW) {
    const int base = chunk_lds_base<N, L, V>(lane);
#pragma unroll
    for (int it = 0; it < T::IT; ++it) {
      const int c = lane + kWave * it;
      if (T::CH % kWave == 0 || c < T::CH) {
        f32x4 v = *reinterpret_cast<const f32x4*>(lds + chunk_lds_offset<N, L, V>(base, lane, it));
        __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(dst) + c);
      }
    }
  } else {
    for (int f = lane; f < valid * NN; f += kWave) {
      const int m = f / NN, e = f - m * NN, i = e / N, j = e - i * N;
      dst[f] = lds[m * G::MS + i * G::LD + j];
    }
  }
}

// -------------------------------------------------------- pivot search

// First-max scan over positions, identical in every lane of a matrix: the
// rows still unplaced publish |a[.][k]| at their current position, then each
// lane walks positions k..n-1 with strict '>' (serial_pivot/luBatchedInplace.cuh:
// 25-34) and swaps (:134-139).  A NaN at position k is never beaten; a NaN
// elsewhere never wins -- the reference's comparison semantics.
template <int N, int L>
__device__ __forceinline__ void pivot_serial(const ColBlock<N, L>& w, LaneInts<N, L>& pos, float* col,
                                             int l) {
  using G = Geo<N, L>;
#pragma unroll
  for (int h = 0; h < G::R; ++h) pos[h] = l + L * h;
  static_for<0, N>([&](auto K) {
    constexpr int k = decltype(K)::value;
#pragma unroll
    for (int h = 0; h < G::R; ++h)
      if (l + L * h < N && pos[h] >= k) col[pos[h]] = fabsf(w[h][k]);
    wave_sync();
    int best = k;
    float bv = col[k];
#pragma unroll
    for (int p = k + 1; p < N; ++p) {
      const float v = col[p];
      if (v > bv) { bv = v; best = p; }
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < G::R; ++h) pos[h] = (pos[h] == k) ? best : ((pos[h] == best) ? k : pos[h]);
  });
}

// Slot T of the tree below after levels 0..I (level i has stride n >> (i+1):
// n/2, n/4, ..., 1).  Level i writes only slots t < n >> (i+1) and reads only
// slots below n >> i, all written by level i-1, so each slot after level I is
// the left-biased comparison of two slots after level I-1 -- a binary tree,
// evaluated here depth first so that only about log2(n) slots are live.
template <int N, int K, int T, int I>
__device__ __forceinline__ void tree_slot(const float* col, float vk, float& v, int& p) {
  if constexpr (I < 0) {
    v = vk;
    p = K;
    if constexpr (K + 1 + T < N) {
      const float c = col[K + 1 + T];
      if (c > vk) { v = c; p = K + 1 + T; }
    }
  } else {
    float v2;
    int p2;
    tree_slot<N, K, T, I - 1>(col, vk, v, p);
    tree_slot<N, K, T + (N >> (I + 1)), I - 1>(col, vk, v2, p2);
    if (v < v2) { v = v2; p = p2; }
  }
}

// Number of levels of the tree for n slots (strides n/2, ..., 1).
constexpr int tree_levels(int n) {
  int d = 0;
  while ((n >> (d + 1)) > 0) ++d;
  return d;
}

// parallel_pivot's find_pivot_parallel (luBatchedInplace.cuh:12-44) with
// threadsPerMatrix = n, evaluated redundantly by every lane: slot t starts at
// (|a_kk|, k) and takes position k+1+t when strictly larger; the tree then
// walks stride = n/2, n/4, ..., 1 keeping the left slot unless it is strictly
// smaller.  For n not a power of two some slots never reach slot 0, as there.
// Exact for ties and NaN (the comparisons are the reference's).
template <int N, int L>
__device__ __forceinline__ void pivot_tree(const ColBlock<N, L>& w, LaneInts<N, L>& pos, float* col,
                                           int l) {
  using G = Geo<N, L>;
#pragma unroll
  for (int h = 0; h < G::R; ++h) pos[h] = l + L * h;
  static_for<0, N>([&](auto K) {
    constexpr int k = decltype(K)::value;
#pragma unroll
    for (int h = 0; h < G::R; ++h)
      if (l + L * h < N && pos[h] >= k) col[pos[h]] = fabsf(w[h][k]);
    wave_sync();
    const float vk = col[k];
    float bv;
    int best;
    tree_slot<N, k, 0, tree_levels(N) - 1>(col, vk, bv, best);
    wave_sync();
#pragma unroll
    for (int h = 0; h < G::R; ++h) pos[h] = (pos[h] == k) ? best : ((pos[h] == best) ? k : pos[h]);
  });
}

// Cross-lane argmax: per column, each still-unplaced row offers the bit
// pattern of |a| (monotone in |a| for non-NaN values), a DPP max over the
// lane group finds the largest, and the row(s) matching it are placed at
// position k.  With no exact tie and no NaN the winner is unique and equals
// the first-max winner.  Both failure cases are detected once, after the last
// column: a tie places two rows at one position, so the positions of a
// matrix no longer form a permutation of 0..n-1 (checked with a DPP OR of
// position bits); a NaN candidate is the largest key of its column (its bit
// pattern exceeds +inf).  Returns false after a tie or a NaN; the caller then
// reruns the exact serial scan for the whole wave.
template <int N, int L>
__device__ __forceinline__ bool pivot_parallel(const ColBlock<N, L>& w, LaneInts<N, L>& pos, int l) {
  using G = Geo<N, L>;
  bool rem[G::R];
#pragma unroll
  for (int h = 0; h < G::R; ++h) { rem[h] = (l + L * h) < N; pos[h] = 63; }
  uint32_t top = 0;  // largest key seen: > +inf only if some column's maximum was a NaN
  static_for<0, N>([&](auto K) {
    constexpr int k = decltype(K)::value;
    uint32_t key[G::R];
#pragma unroll
    for (int h = 0; h < G::R; ++h) key[h] = abs_key(w[h][k], rem[h]);
    uint32_t m = key[0];
#pragma unroll
    for (int h = 1; h < G::R; ++h) m = max(m, key[h]);
    m = group_max_u32<L>(m);
    top = max(top, m);
    keep_in_vgpr(top);
#pragma unroll
    for (int h = 0; h < G::R; ++h) {
      const bool win = rem[h] && key[h] == m;
      rem[h] = rem[h] && !win;
      pos[h] = win ? k : pos[h];
    }
  });
  uint32_t bits = 0;
#pragma unroll
  for (int h = 0; h < G::R; ++h)
    bits |= ((l + L * h) < N && (uint32_t)pos[h] < 32u) ? (1u << pos[h]) : 0u;
  bits = group_or_u32<L>(bits);
  constexpr uint32_t full = (N == 32) ? 0xffffffffu : ((1u << N) - 1u);
  const bool ok = (bits == full) && (top <= 0x7f800000u);
  return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

// Slots of find_pivot_parallel's tree (threadsPerMatrix = n) whose value can
// reach slot 0: at stride s = n/2, n/4, ..., 1 slot t < s absorbs slot t + s
// (parallel_pivot/luBatchedInplace.cuh:34-42).  All n slots for a power of
// two; for other n some never do (n = 6: slots 0, 1, 3, 4).
constexpr uint32_t tree_reach(int n) {
  uint32_t from[32] = {};
  for (int t = 0; t < n; ++t) from[t] = 1u << t;
  for (int s = n / 2; s > 0; s >>= 1)
    for (int t = 0; t < s; ++t) from[t] |= from[t + s];
  return from[0];
}

// The tree rule without the tree.  Slot t starts as (|a_kk|, k) and takes
// (|a_p|, p), p = k+1+t, only when strictly larger, and each tree comparison
// keeps the left slot unless it is strictly smaller.  So when the candidates
// -- position k and the positions k+1+t of slots that reach slot 0 -- hold no
// NaN and no exact tie for their largest |a|, that strict maximum survives
// every comparison on its way to slot 0, and the winner is the candidate with
// the largest |a|: the DPP max of pivot_parallel over the candidates only.
// The candidate set depends on where rows sit, so positions follow the
// reference's swap (:165-172): the winner goes to k, the row at k to the
// winner's position.  A power of two makes every remaining row a candidate,
// and this is pivot_parallel.  A tie or a NaN is detected as there (two rows
// placed at one position, or a key above +inf); the caller then reruns the
// exact tree for the wave.
template <int N, int L>
__device__ __forceinline__ bool pivot_tree_dpp(const ColBlock<N, L>& w, LaneInts<N, L>& pos, int l) {
  constexpr uint32_t full = (N == 32) ? 0xffffffffu : ((1u << N) - 1u);
  constexpr uint32_t reach = tree_reach(N);
  if constexpr (reach == full) {
    return pivot_parallel<N, L>(w, pos, l);
  } else {
    using G = Geo<N, L>;
#pragma unroll
    for (int h = 0; h < G::R; ++h) pos[h] = l + L * h;  // rows past n sit at positions >= n: never candidates
    uint32_t top = 0;
    static_for<0, N>([&](auto K) {
      constexpr int k = decltype(K)::value;
      // bit p set: position p is a candidate at step k (placed rows sit below k)
      constexpr uint32_t cand_bits = static_cast<uint32_t>((((uint64_t)reach << 1) | 1u) << k) & full;
      bool cand[G::R];
      uint32_t key[G::R];
#pragma unroll
      for (int h = 0; h < G::R; ++h) {
        cand[h] = ((cand_bits >> pos[h]) & 1u) != 0;
        key[h] = abs_key(w[h][k], cand[h]);
      }
      uint32_t m = key[0];
#pragma unroll
      for (int h = 1; h < G::R; ++h) m = max(m, key[h]);
      m = group_max_u32<L>(m);
      top = max(top, m);
      keep_in_vgpr(top);
      bool win[G::R];
      uint32_t wp = 0;
#pragma unroll
      for (int h = 0; h < G::R; ++h) {
        win[h] = cand[h] && key[h] == m;
        wp = max(wp, win[h] ? static_cast<uint32_t>(pos[h]) : 0u);
      }
      wp = group_max_u32<L>(wp);
#pragma unroll
      for (int h = 0; h < G::R; ++h) pos[h] = win[h] ? k : (pos[h] == k ? static_cast<int>(wp) : pos[h]);
    });
    uint32_t bits = 0;
#pragma unroll
    for (int h = 0; h < G::R; ++h) bits |= ((l + L * h) < N) ? (1u << pos[h]) : 0u;
    bits = group_or_u32<L>(bits);
    const bool ok = (bits == full) && (top <= 0x7f800000u);
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
  }
}

// ------------------------------------------------------ Gauss-Jordan

// In-place Gauss-Jordan with row operations on S = P*A held by columns: lane l
// of the group holds columns l, l+L, ... (w[h][i] = S[i][l + L*h]).  The
// pivots are the diagonal of P*A in order, so the multipliers are those of
// the reference's no-pivot Doolittle factorisation of P*A.  Step k:
//   d = S[k][k] (broadcast from lane k%L of the group), r = 1/d
//   g_c = S[k][c] * r, the lane's own multiplier (0 in the lane owning column
//     k, so column k is left as it is)
//   every row i != k:  S[i][c] -= S[i][k] * g_c, one v_fmac_f32_dpp per
//     element with S[i][k] broadcast (negated) from the lane owning column k
//     (row k's update would be discarded, so it is not computed):
//   row k is set to its normalised values g_c
//   column k turns into right-hand column k.  It is kept unscaled: the right
//     part R[i][k] = -S[i][k] / d is sigma_k * S[i][k] with sigma_k = -1/d,
//     and every later row operation commutes with that column scale; its
//     pivot entry becomes R[k][k] / sigma_k = -1.
// At the end (P*A)^-1 = S * diag(sigma); sigma of the lane's columns is cs[].
// A pivot d = 0 leaves sigma = NaN (rcp(0) = inf, then the Newton step gives
// -0 * inf = NaN), as does a pivot without a finite reciprocal (inf, NaN, or
// below ~3e-39); pivot_failure() reads that back.
template <int N, int L>
__device__ __forceinline__ void gauss_jordan(ColBlock<N, L>& w, LaneFloats<N, L>& cs, int l) {
  using G = Geo<N, L>;
#pragma unroll
  for (int h = 0; h < G::R; ++h) cs[h] = 1.0f;
  static_for<0, N>([&](auto K) {
    constexpr int k = decltype(K)::value;
    constexpr int kl = k % L, kh = k / L;
    const float d = dpp_f32<bcast_ctrl<L, kl>()>(w[kh][k]);
    float r = __builtin_amdgcn_rcpf(d);
    r = __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.0f), r);  // one Newton step
    const bool pl = (l == kl);                              // this lane owns column k (set kh)
    float g[G::R];
#pragma unroll
    for (int h = 0; h < G::R; ++h) g[h] = w[h][k] * r;
    g[kh] = pl ? 0.0f : g[kh];
    cs[kh] = pl ? -r : cs[kh];
    static_for<0, G::R>([&](auto H) {
      constexpr int h = decltype(H)::value;
      // (h in the template argument keeps the call dependent, so it is only
      // instantiated for h != kh; cross blocks exist only where R > 1)
      if constexpr (h != kh) GjBlock<N + 0 * h, L, kl, k>::cross(w[h], w[kh], g[h]);
    });
    GjBlock<N, L, kl, k>::self(w[kh], g[kh]);
#pragma unroll
    for (int h = 0; h < G::R; ++h) w[h][k] = g[h];
    w[kh][k] = pl ? -1.0f : g[kh];
  });
}

// min over the L lanes of a group (the same DPP steps as group_max_u32).
template <int L>
__device__ __forceinline__ uint32_t group_min_u32(uint32_t m) {
  m = min(m, dpp_u32<0xB1>(m));
  m = min(m, dpp_u32<0x4E>(m));
  if constexpr (L == 16) {
    m = min(m, dpp_u32<0x141>(m));
    m = min(m, dpp_u32<0x140>(m));
  }
  return m;
}

// First elimination step whose pivot was zero or had no finite reciprocal
// (sigma of that column is NaN), or N: LAPACK's info - 1 for a zero pivot.
// Lane l of a group holds sigma of columns l, l+L, ...
template <int N, int L>
__device__ __forceinline__ int pivot_failure(const LaneFloats<N, L>& cs, int l) {
  using G = Geo<N, L>;
  uint32_t first = N;
#pragma unroll
  for (int h = G::R - 1; h >= 0; --h)
    if (l + L * h < N && cs[h] != cs[h]) first = l + L * h;
  return static_cast<int>(group_min_u32<L>(first));
}

// ------------------------------------------------------------ the kernel

// Phases 2-5 of a tile already staged in LDS: pivot positions, P*A in LDS,
// then columns q, q+L, ... of P*A into the registers of lane q.
template <int N, int L, int PIV, int V = 0>
__device__ __forceinline__ void tile_front(ColBlock<N, L>& w, LaneInts<N, L>& prm, float* lds,
                                           int lane) {
  using G = Geo<N, L, V>;
  const int l = lane % L;
  const int g = lane / L;
  float* const mat = lds + g * G::MS;
  int* const perm = reinterpret_cast<int*>(lds + G::PERM) + g * 32;

  if constexpr (PIV != kPivNone) {
    // 2. original rows l, l+L, ... (rows past n read row n-1 and stay unused)
#pragma unroll
    for (int h = 0; h < G::R; ++h) lds_row_read<N>(mat + min(l + L * h, N - 1) * G::LD, w[h]);
    // 3. pivot positions of this lane's rows
    int pos[G::R];
    if constexpr (PIV == kPivParallel || PIV == kPivSerial) {  // one rule: first max
      if (!pivot_parallel<N, L>(w, pos, l)) {
        wave_sync();
        pivot_serial<N, L>(w, pos, mat, l);
      }
    } else {
      if (!pivot_tree_dpp<N, L>(w, pos, l)) {
        wave_sync();
        pivot_tree<N, L>(w, pos, mat, l);
      }
    }
    wave_sync();
    // 4. P*A: row rho goes to position pos; perm[pos] = rho
#pragma unroll
    for (int h = 0; h < G::R; ++h) {
      if (l + L * h < N) {
        lds_row_write<N>(mat + pos[h] * G::LD, w[h]);
        perm[pos[h]] = l + L * h;
      }
    }
    wave_sync();
  }

  // 5. column q of P*A -> lane q
#pragma unroll
  for (int h = 0; h < G::R; ++h) {
    const int q = min(l + L * h, N - 1);
#pragma unroll
    for (int p = 0; p < N; ++p) w[h][p] = mat[p * G::LD + q];
    prm[h] = (PIV != kPivNone) ? (perm[q] & 31) : q;  // a permutation of 0..n-1; the mask keeps LDS indexing in range
  }
  wave_sync();
}

// Phases 6-7: (P*A)^-1 in registers, column q written to column perm[q] of
// the matrix's LDS region, which then holds A^-1 (row-major, stride LD).
// With `failure` non-null, *failure receives pivot_failure() of this lane's
// matrix.
template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_solve(ColBlock<N, L>& w, const LaneInts<N, L>& prm, float* lds, int lane,
                                           int* failure = nullptr) {
  using G = Geo<N, L, V>;
  const int l = lane % L;
  const int g = lane / L;
  float* const mat = lds + g * G::MS;

  // 6. (P*A)^-1 = S * diag(cs), column q in lane q
  LaneFloats<N, L> cs;
  gauss_jordan<N, L>(w, cs, l);
  if (failure != nullptr) *failure = pivot_failure<N, L>(cs, l);

  // 7. column q of (P*A)^-1 is column perm[q] of A^-1
#pragma unroll
  for (int h = 0; h < G::R; ++h) {
    if (l + L * h < N) {
      float* const colp = mat + prm[h];
#pragma unroll
      for (int i = 0; i < N; ++i) colp[i * G::LD] = w[h][i] * cs[h];
    }
  }
  wave_sync();
}

// Phase 8 for a contiguous tile whose inverses are in LDS: the inverses, the
// info words and the pivot vectors (prm[h] = perm[l + L*h]) to HBM.
template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_finish(const LaneInts<N, L>& prm, int failure, float* __restrict__ A,
                                            int* __restrict__ piv, int* __restrict__ info, int64_t m0, int valid,
                                            bool vec, float* lds, int lane) {
  using G = Geo<N, L, V>;
  const int l = lane % L;
  const int g = lane / L;
  tile_store<N, L, V>(A + m0 * (N * N), lds, valid, vec, lane);
  if (info != nullptr && l == 0 && g < valid) info[m0 + g] = failure < N ? failure + 1 : 0;
  if (piv != nullptr && g < valid) {
#pragma unroll
    for (int h = 0; h < G::R; ++h)
      if (l + L * h < N) piv[(m0 + g) * N + l + L * h] = prm[h];
  }
  wave_sync();  // the next tile reuses the LDS region
}

// Phases 6-8 of the reference's pivot rules for a contiguous tile.
template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_back(ColBlock<N, L>& w, const LaneInts<N, L>& prm, float* __restrict__ A,
                                          int* __restrict__ piv, int* __restrict__ info, int64_t m0, int valid,
                                          bool vec, float* lds, int lane) {
  int failure = N;
  tile_solve<N, L, V>(w, prm, lds, lane, info != nullptr ? &failure : nullptr);
  tile_finish<N, L, V>(prm, failure, A, piv, info, m0, valid, vec, lds, lane);
}

// --------------------------------------------- partial pivoting (GEPP)

// Gauss-Jordan with LAPACK's partial pivoting for a tile staged in LDS: at
// step k the pivot is the largest |a| of the updated column k among the rows
// not yet pivotal, ties going to the smallest current position, and the
// positions of rows k and p are exchanged as getrf's row swap would.  (The
// reference's rules search the un-updated column; this one has no
// counterpart there.)  Rows stay in their lanes: lane l of the group holds
// rows l, l+L, ... (w[h][c] = S[l + L*h][c]).  Step k, pivot row p,
// d = S[p][k], r = 1/d (rcp + one Newton step):
//   every row i != p:  S[i][c] -= (S[i][k] * r) * S[p][c]   (c != k)
//   S[p][k] = -S[p][k];  rho[p] = r;  sigma[k] = -r
// Row p keeps its unnormalised values and column k its unscaled ones: the
// true matrix is rho[i] S[i][c] sigma[c], and both scales commute with every
// later row operation.  The pivot row's lane stores it in the matrix's perm
// slot (free until the end); lanes 0..ceil(N/4)-1 of the group read it back
// four floats each, and each update takes its pivot-row operand from the
// owning lane by DPP broadcast (GeppBlock), one v_fmac_f32_dpp per element.
// After the last step row perm[k] holds row k of (P A)^-1, and
// (P A)^-1 [k][c] = A^-1 [k][perm[c]]: the rows go to LDS by position, then
// column q to column perm[q].  On return the region holds A^-1 (row-major,
// stride LD), perm follows the tile, prm[h] = perm[l + L*h], and *failure
// (if non-null) the first step whose pivot has no finite reciprocal, or N.
// A NaN in the pivot column is taken as its largest entry.
template <int N, int L, int V = 0>
__device__ __forceinline__ void tile_gepp(LaneInts<N, L>& prm, float* lds, int lane, int* failure) {
  using G = Geo<N, L, V>;
  const int l = lane % L;
  const int g = lane / L;
  float* const mat = lds + g * G::MS;
  int* const perm = reinterpret_cast<int*>(lds + G::PERM) + g * 32;
  ColBlock<N, L> w;
#pragma unroll
  for (int h = 0; h < G::R; ++h) lds_row_read<N>(mat + min(l + L * h, N - 1) * G::LD, w[h]);
  bool rem[G::R];
  int pos[G::R];
  float rho[G::R];
#pragma unroll
  for (int h = 0; h < G::R; ++h) {
    rem[h] = (l + L * h) < N;  // rows past n (copies of row n-1) never pivot and are not stored
    pos[h] = l + L * h;
    rho[h] = 1.0f;
  }
  float sig[N];
  float* const prow = reinterpret_cast<float*>(perm);  // the pivot row of the step
  constexpr int kQ = L == 16 ? 8 : 4;                  // lanes that hold it, 4 floats each
  static_for<0, N>([&](auto K) {
    constexpr int k = decltype(K)::value;
    uint32_t key[G::R];
#pragma unroll
    for (int h = 0; h < G::R; ++h) key[h] = abs_key(w[h][k], rem[h]);
    uint32_t m = key[0];
#pragma unroll
    for (int h = 1; h < G::R; ++h) m = max(m, key[h]);
    m = group_max_u32<L>(m);
    uint32_t first = 0xffu;  // smallest current position holding the maximum
#pragma unroll
    for (int h = 0; h < G::R; ++h) first = (rem[h] && key[h] == m) ? min(first, (uint32_t)pos[h]) : first;
    const int pm = static_cast<int>(group_min_u32<L>(first));
    bool pv[G::R];
#pragma unroll
    for (int h = 0; h < G::R; ++h) {
      pv[h] = rem[h] && pos[h] == pm;
      pos[h] = pv[h] ? k : (pos[h] == k ? pm : pos[h]);
      rem[h] = rem[h] && !pv[h];
    }
    // One store block for whichever of the lane's rows is the pivot: a
    // masked store costs the LDS about as much as a full one, so selecting
    // first (VALU) halves the stores of a two-row lane.
    bool mine = pv[0];
#pragma unroll
    for (int h = 1; h < G::R; ++h) mine = mine || pv[h];
    if (mine) {
      float prw[N];
#pragma unroll
      for (int c = 0; c < N; ++c) {
        float v = w[0][c];
#pragma unroll
        for (int h = 1; h < G::R; ++h) v = pv[h] ? w[h][c] : v;
        prw[c] = v;
      }
      lds_row_write<N>(prow, prw);
    }
    wave_sync();
    float q[4];
    {
      const f32x4 v = *reinterpret_cast<const f32x4*>(prow + 4 * (l % kQ));
      q[0] = v.x; q[1] = v.y; q[2] = v.z; q[3] = v.w;
    }
    const float d = dpp_f32<bcast_ctrl<L, k / 4>()>(q[k % 4]);
    float r = __builtin_amdgcn_rcpf(d);
    r = __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.0f), r);
    sig[k] = -r;
#pragma unroll
    for (int h = 0; h < G::R; ++h) {
      const float mlt = pv[h] ? 0.0f : w[h][k] * r;
      GeppBlock<N, L, k>::upd(w[h], q, mlt);  // w[h][c] -= mlt * P[c], c != k
      w[h][k] = pv[h] ? -w[h][k] : w[h][k];
      rho[h] = pv[h] ? r : rho[h];
    }
  });
  if (failure != nullptr) {
    int f = N;
#pragma unroll
    for (int k = N - 1; k >= 0; --k) f = (sig[k] != sig[k]) ? k : f;
    *failure = f;
  }
  wave_sync();  // the last pivot row has been read before its slot takes perm
#pragma unroll
  for (int h = 0; h < G::R; ++h) {
    if (l + L * h < N) {
#pragma unroll
      for (int c = 0; c < N; ++c) w[h][c] = (w[h][c] * sig[c]) * rho[h];
      lds_row_write<N>(mat + pos[h] * G::LD, w[h]);
      perm[pos[h]] = l + L * h;
    }
  }
  wave_sync();
  // column q of (P A)^-1 -> column perm[q] of A^-1
#pragma unroll
  for (int h = 0; h < G::R; ++h) {
    const int q = min(l + L * h, N - 1);
#pragma unroll
    for (int p = 0; p < N; ++p) w[h][p] = mat[p * G::LD + q];
    prm[h] = perm[q] & 31;  // a permutation of 0..n-1; the mask keeps LDS indexing in range
  }
  wave_sync();
#pragma unroll
  for (int h = 0; h < G::R; ++h) {
    if (l + L * h < N) {
      float* const colp = mat + prm[h];
#pragma unroll
      for (int i = 0; i < N; ++i) colp[i * G::LD] = w[h][i];
    }
  }
  wave_sync();
}

// Waves per SIMD each plain kernel is compiled for (its VGPR budget: 2 ->
// 256, 3 -> 168).  3 where the LDS tile leaves room for more than 8 waves per
// CU (160 KB / LDS_FLOATS) and the kernel fits 168 VGPRs without spilling
// (checked per instantiation with -Rpass-analysis=kernel-resource-usage:
// n = 25..31 with the first-max rules, 25..30 with the tree rule); elsewhere
// 2, the LDS or the register tile being the limit anyway.
template <int N, int L, int PIV, int V = 0>
constexpr int kWavesPerEu =
    (L == 16 && N >= 25 && N <= 31 && PIV != kPivGepp && !(PIV == kPivTree && N == 31) &&
     160 * 1024 / (Geo<N, L, V>::LDS_FLOATS * 4) > 8) ? 3 : 2;

// Each workgroup (one wave) strides over tiles.  The next full tile is
// fetched into registers once the current one has left LDS for registers
// (after the column read), so with several tiles per workgroup its HBM reads
// overlap this tile's Gauss-Jordan, the longest phase.
// amdgpu_waves_per_eu: see kWavesPerEu.
template <int N, int L, int PIV, int V = 0>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<N, L, PIV, V>))) void lu_inverse_kernel(
    float* __restrict__ A, int* __restrict__ piv, int* __restrict__ info, int64_t batch, int64_t ntiles, int vec) {
  using G = Geo<N, L, V>;
  __shared__ __attribute__((aligned(16))) float lds[G::LDS_FLOATS];
  const int lane = threadIdx.x;
  const int64_t stride = gridDim.x;
  auto full = [&](int64_t t) { return vec != 0 && (t + 1) * G::MPW <= batch; };
  TileRegs<N, L> next;
  int64_t t = blockIdx.x;
  bool have = t < ntiles && full(t);
  if (have) tile_fetch<N, L>(A + t * G::MPW * (N * N), next, lane);
  for (; t < ntiles; t += stride) {
    const int64_t m0 = t * G::MPW;
    const int64_t left = batch - m0;
    const int valid = left < G::MPW ? (int)left : G::MPW;
    if (have) tile_stage<N, L, V>(next, lds, lane);
    else tile_load_elementwise<N, L, V>(A + m0 * (N * N), lds, valid, lane);
    wave_sync();
    int prm[G::R];
    if constexpr (PIV == kPivGepp) {
      const int64_t tn = t + stride;
      have = tn < ntiles && full(tn);
      if (have) tile_fetch<N, L>(A + tn * G::MPW * (N * N), next, lane);
      int failure = N;
      tile_gepp<N, L, V>(prm, lds, lane, info != nullptr ? &failure : nullptr);
      tile_finish<N, L, V>(prm, failure, A, piv, info, m0, valid, vec != 0, lds, lane);
    } else {
      float w[G::R][N];
      tile_front<N, L, PIV, V>(w, prm, lds, lane);
      const int64_t tn = t + stride;
      have = tn < ntiles && full(tn);
      if (have) tile_fetch<N, L>(A + tn * G::MPW * (N * N), next, lane);
      tile_back<N, L, V>(w, prm, A, piv, info, m0, valid, vec != 0, lds, lane);
    }
  }
}

}  // namespace lub
