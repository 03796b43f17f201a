// lu_ragged.hpp -- mixed-n batches in one inversion launch.
//
// The reference only knows batches of one n (compile-time MATRIXSIZE,
// parallel_pivot/luBatchedInplace.cu:4); a mixed batch there means one build
// and one launch per n.  Here a packed ragged batch -- matrix i of size
// ns[i] at float offset offsets[i], pivots at piv_offsets[i] -- is inverted by
// a short device-side plan and one inversion kernel, all enqueued
// on one stream without host synchronisation or allocation (so the sequence
// can be captured once in a hipGraph and replayed):
//
//   hist      matrices per n in each block of 1024 (LDS histograms), written
//             as the block's column of a class-major block table; no global
//             atomics
//   scan      one workgroup per class: the class's row of the block table
//             becomes each block's first slot within the class, and the
//             class total goes to the header
//   scatter   every block derives the class table from the 32 totals (block
//             0 publishes it) -- size classes ordered largest n first; per
//             class its first slot, its first descriptor, its lane layout and
//             the matrices per descriptor (MPW(n) * grab(n) <= 64: MPW is 4
//             with 16-lane rows, 16 with quads) -- then writes each matrix's
//             offset and pivot offset to its slot (class first + the block's
//             first slot + a rank from an LDS atomic), so a descriptor's
//             matrices sit in consecutive slots
//   invert    one workgroup (one wave) per descriptor, classes largest first
//             so the tail of the launch is made of the cheapest work.  A
//             descriptor is located from the class table held in lanes 0..32
//             (ballot + readlane; the 1.5 KB table is read by every wave and
//             stays cached), its up to 64 slot offsets arrive with one vector
//             load (one per lane), and its tiles run the plain kernel's tile
//             routines; each matrix is gathered and scattered by its own lane
//             group, and the loads of tile b+1 are issued before the stores
//             of tile b.
//
// A matrix's result does not depend on which tile or slot it lands in: every
// tile routine treats its matrices independently (the serial fallback of the
// parallel pivot search reproduces the same first-max pivots), so results
// are identical to a plain launch of the same n.
#pragma once

#include "lu_kernel.hpp"

namespace lub {

struct RaggedTuning {
  int8_t layout[33];  // per n: 0 = 16-lane rows, 1 = quads (n <= 12 only in the ragged kernel)
  int8_t grab[33];    // per n: wave tiles per descriptor (>= 1; at most 64 matrices per descriptor)
};

// Workspace: an int32 header; per planning block (CHUNK consecutive
// matrices) 33 int32 -- its count of matrices with n outside 1..32 and per
// class its count, which the scan turns into the block's first slot in that
// class; then two int64 slot arrays (offsets and pivot offsets in class
// order).  Every field is written before it is read in each call, so nothing
// needs clearing between calls.
struct RaggedWs {
  static constexpr int COUNTS = 0;        // [n] matrices of size n
  static constexpr int CLASS_FIRST = 128; // [n] first slot of class n
  static constexpr int DESC_FIRST = 192;  // [n] first descriptor of class n; [0] descriptors in all
  static constexpr int META = 256;        // [2] matrices with n outside 1..32
  static constexpr int LPER = 320;        // [n] layout | matrices per descriptor << 8
  static constexpr int HEADER = 384;
  static constexpr int CHUNK = 1024;      // matrices per planning block
  static constexpr int BSTRIDE = 33;      // int32 per planning block
  static constexpr int64_t blocks(int64_t cap) { return (cap + CHUNK - 1) / CHUNK; }
  static constexpr int64_t bhist_offset() { return HEADER * 4; }  // bytes
  static constexpr int64_t slots_offset(int64_t cap) {
    return (bhist_offset() + blocks(cap) * BSTRIDE * 4 + 15) / 16 * 16;
  }
  static constexpr int64_t bytes(int64_t cap) { return slots_offset(cap) + 2 * 8 * cap; }
};

constexpr int kRaggedQuadMaxN = 12;
constexpr int kRaggedMaxPer = 64;  // matrices per descriptor: one slot per lane

__host__ __device__ constexpr int ragged_mpw(int n, int layout) { return (layout == 1 && n <= kRaggedQuadMaxN) ? 16 : 4; }

// ---------------------------------------------------------- inversion

template <int A_, int B_>
constexpr int cmax() { return A_ > B_ ? A_ : B_; }

constexpr int kRaggedLdsFloats = cmax<Geo<32, 16, 1>::LDS_FLOATS, Geo<kRaggedQuadMaxN, 4, 1>::LDS_FLOATS>();

// Float4 of a ragged batch: matrices start at any float offset, so global
// accesses are only 4-byte aligned (gfx950 executes dword-aligned dwordx4).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// Slot offsets of one descriptor, one slot per lane (lanes 0..15; lanes past
// the descriptor's count repeat slot 0): matrix offset and pivot offset.
struct RaggedSlots {
  uint32_t so_lo, so_hi, po_lo, po_hi;
};

// Offset of slot `s` (wave-uniform) as a scalar.
__device__ __forceinline__ int64_t slot_uniform(uint32_t lo, uint32_t hi, int s) {
  const uint32_t l = __builtin_amdgcn_readlane(lo, s);
  const uint32_t h = __builtin_amdgcn_readlane(hi, s);
  return static_cast<int64_t>((uint64_t(h) << 32) | l);
}

// Offset of slot `s` (per lane); every lane of the wave must take part.
__device__ __forceinline__ int64_t slot_lane(uint32_t lo, uint32_t hi, int s) {
  const uint32_t l = __builtin_amdgcn_ds_bpermute(s * 4, static_cast<int>(lo));
  const uint32_t h = __builtin_amdgcn_ds_bpermute(s * 4, static_cast<int>(hi));
  return static_cast<int64_t>((uint64_t(h) << 32) | l);
}

// Registers of one tile's gather.  Each matrix is loaded by its own lane
// group (the lanes that will factor it): lane l of group g holds chunks
// c = l + L*k (k < K; 4 consecutive floats each) of matrix g and, when
// n*n % 4 != 0, float 4*C4 + l (l < REM) -- so one load instruction reads
// MPW matrices at once, whatever n is.
template <int N, int L>
struct RaggedRegs {
  static constexpr int NN = N * N, C4 = NN / 4, REM = NN % 4;
  static constexpr int K = (C4 + L - 1) / L;
  f32x4u v[K > 0 ? K : 1];
  float r;
};

// Gather loads of up to MPW matrices of one n (slots base .. base+valid-1 of
// `sl`) into registers; groups past `valid` load nothing.  Every lane of the
// wave must call this (the slot lookup is a cross-lane permute).
template <int N, int L>
__device__ __forceinline__ void ragged_load(const float* __restrict__ A, const RaggedSlots& sl, int base, int valid,
                                            RaggedRegs<N, L>& x, int lane) {
  using X = RaggedRegs<N, L>;
  const int l = lane % L, g = lane / L;
  const int64_t off = slot_lane(sl.so_lo, sl.so_hi, base + (g < valid ? g : 0));
  if (g < valid) {
    const float* src = A + off;
#pragma unroll
    for (int k = 0; k < X::K; ++k) {
      const int c = l + L * k;
      if (X::C4 % L == 0 || c < X::C4)
        x.v[k] = __builtin_nontemporal_load(reinterpret_cast<const f32x4u*>(src) + c);
    }
    if constexpr (X::REM != 0) x.r = __builtin_nontemporal_load(src + 4 * X::C4 + (l < X::REM ? l : 0));
  }
}

// LDS address (floats, within the matrix) of chunk c in the gathered layout:
// rows of n % 4 == 0 hold whole chunks (row stride LD); other n are stored
// contiguously from a 16-byte boundary.
template <int N, int L>
__device__ __forceinline__ int ragged_chunk_at(int c) {
  using G = Geo<N, L, 1>;
  if constexpr (N % 4 == 0) {
    const int i = (4 * c) / N, j = (4 * c) % N;
    return i * G::LD + j;
  } else {
    return 4 * c;
  }
}

// Registers -> the tile's LDS layout (Geo<N, L, 1>), one aligned dwordx4 per
// chunk; groups past `valid` get the identity.
template <int N, int L>
__device__ __forceinline__ void ragged_stage(const RaggedRegs<N, L>& x, int valid, float* lds, int lane) {
  using G = Geo<N, L, 1>;
  using X = RaggedRegs<N, L>;
  const int l = lane % L, g = lane / L;
  float* const mat = lds + g * G::MS;
  if (valid == G::MPW) {  // wave-uniform: full tiles stage the data only
#pragma unroll
    for (int k = 0; k < X::K; ++k) {
      const int c = l + L * k;
      if (X::C4 % L == 0 || c < X::C4) {
        const f32x4u v = x.v[k];
        *reinterpret_cast<f32x4*>(mat + ragged_chunk_at<N, L>(c)) = f32x4{v[0], v[1], v[2], v[3]};
      }
    }
    if constexpr (X::REM != 0)
      if (l < X::REM) mat[4 * X::C4 + l] = x.r;
  } else {
    // A partial tile: groups past `valid` get the identity, as zeros and then
    // the diagonal (the same wave's LDS writes land in order).  Nothing here
    // is a function of the lane alone, so nothing is hoisted out of the tile
    // loop into registers that would stay live through the factorisation.
    const bool live = g < valid;
#pragma unroll
    for (int k = 0; k < X::K; ++k) {
      const int c = l + L * k;
      if (X::C4 % L == 0 || c < X::C4) {
        const f32x4u v = x.v[k];
        const f32x4 w = live ? f32x4{v[0], v[1], v[2], v[3]} : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        *reinterpret_cast<f32x4*>(mat + ragged_chunk_at<N, L>(c)) = w;
      }
    }
    if constexpr (X::REM != 0)
      if (l < X::REM) mat[4 * X::C4 + l] = live ? x.r : 0.0f;
    if (!live) {
#pragma unroll
      for (int h = 0; h < G::R; ++h) {
        const int i = l + L * h;
        if (i < N) mat[i * G::LD + i] = 1.0f;
      }
    }
  }
}

// The mirror of ragged_load + ragged_stage for the `valid` leading matrices
// of the tile; every lane of the wave must call it.
template <int N, int L>
__device__ __forceinline__ void ragged_scatter_store(float* __restrict__ A, const RaggedSlots& sl, int base,
                                                     int valid, const float* lds, int lane) {
  using G = Geo<N, L, 1>;
  using X = RaggedRegs<N, L>;
  const int l = lane % L, g = lane / L;
  const int64_t off = slot_lane(sl.so_lo, sl.so_hi, base + (g < valid ? g : 0));
  if (g < valid) {
    const float* const mat = lds + g * G::MS;
    float* dst = A + off;
#pragma unroll
    for (int k = 0; k < X::K; ++k) {
      const int c = l + L * k;
      if (X::C4 % L == 0 || c < X::C4) {
        const f32x4 t = *reinterpret_cast<const f32x4*>(mat + ragged_chunk_at<N, L>(c));
        __builtin_nontemporal_store(f32x4u{t[0], t[1], t[2], t[3]}, reinterpret_cast<f32x4u*>(dst) + c);
      }
    }
    if constexpr (X::REM != 0)
      if (l < X::REM) __builtin_nontemporal_store(mat[4 * X::C4 + l], dst + 4 * X::C4 + l);
  }
}

// One descriptor (count <= 64 matrices of one n) in tiles of MPW: gather,
// the plain kernel's tile routines (factor, invert, unpermute into LDS),
// scatter, pivots.  The loads of tile b+1 are issued before the stores of
// tile b: the memory counter retires in order, so loads issued after the
// stores would also wait for the stores to complete.
template <int N, int L, int PIV>
__device__ __forceinline__ void ragged_tile(float* __restrict__ A, int* __restrict__ piv, const RaggedSlots& sl,
                                            int count, float* lds, int lane) {
  using G = Geo<N, L, 1>;
  static_assert(G::LDS_FLOATS <= kRaggedLdsFloats, "tile exceeds the ragged kernel's LDS array");
  const int l = lane % L;
  const int g = lane / L;
  RaggedRegs<N, L> x;
  ragged_load<N, L>(A, sl, 0, min(G::MPW, count), x, lane);
  for (int base = 0;;) {
    const int valid = min(G::MPW, count - base);
    ragged_stage<N, L>(x, valid, lds, lane);
    wave_sync();
    LaneInts<N, L> prm;
    if constexpr (PIV == kPivGepp) {
      tile_gepp<N, L, 1>(prm, lds, lane, nullptr);
    } else {
      ColBlock<N, L> w;
      tile_front<N, L, PIV, 1>(w, prm, lds, lane);
      tile_solve<N, L, 1>(w, prm, lds, lane);
    }
    auto finish = [&](int b0, int nv) {
      ragged_scatter_store<N, L>(A, sl, b0, nv, lds, lane);
      if (piv != nullptr) {
        const int64_t po = slot_lane(sl.po_lo, sl.po_hi, b0 + (g < nv ? g : 0));
        if (g < nv) {
#pragma unroll
          for (int h = 0; h < G::R; ++h)
            if (l + L * h < N) piv[po + l + L * h] = prm[h];
        }
      }
      wave_sync();  // the next tile reuses the LDS region
    };
    // Two exits so that only freshly loaded registers reach the loop head:
    // the staged ones are dead during the factorisation.
    const int nb = base + G::MPW;
    if (nb < count) {
      ragged_load<N, L>(A, sl, nb, min(G::MPW, count - nb), x, lane);
      finish(base, valid);
      base = nb;
    } else {
      finish(base, valid);
      break;
    }
  }
}

struct RaggedDesc {
  int n;
  int layout;  // 0: 16-lane rows, 1: quads
  int first;   // first slot
  int count;   // matrices (<= 16)
};

// Class table of the plan, lane c (1..32) holding class c.
struct RaggedClasses {
  int desc_first, class_first, count, lper;
};

__device__ __forceinline__ RaggedClasses ragged_load_classes(const int* __restrict__ hdr, int lane) {
  RaggedClasses c;
  const bool row = lane <= 32;
  c.desc_first = row ? hdr[RaggedWs::DESC_FIRST + lane] : 0x7fffffff;
  c.class_first = row ? hdr[RaggedWs::CLASS_FIRST + lane] : 0;
  c.count = row ? hdr[RaggedWs::COUNTS + lane] : 0;
  c.lper = row ? hdr[RaggedWs::LPER + lane] : 0;
  return c;
}

// Descriptor t (< the descriptor total): its class is the smallest n whose
// first descriptor is <= t (classes are laid out 32, 31, ..., 1 and an empty
// class starts where the next one does).
__device__ __forceinline__ RaggedDesc ragged_locate(const RaggedClasses& c, int t, int lane) {
  const uint64_t m = __ballot(lane >= 1 && lane <= 32 && c.desc_first <= t);
  const int n = __builtin_ctzll(m);
  const int df = __builtin_amdgcn_readlane(c.desc_first, n);
  const int cf = __builtin_amdgcn_readlane(c.class_first, n);
  const int cnt = __builtin_amdgcn_readlane(c.count, n);
  const int lp = __builtin_amdgcn_readlane(c.lper, n);
  const int per = lp >> 8;
  const int first = cf + (t - df) * per;
  return RaggedDesc{n, lp & 0xff, first, min(per, cf + cnt - first)};
}

__device__ __forceinline__ RaggedSlots ragged_load_slots(const int64_t* __restrict__ soffs,
                                                         const int64_t* __restrict__ spoffs, const RaggedDesc& d,
                                                         int lane) {
  RaggedSlots s{0, 0, 0, 0};
  {
    const int i = d.first + (lane < d.count ? lane : 0);
    const uint64_t so = static_cast<uint64_t>(soffs[i]);
    s.so_lo = static_cast<uint32_t>(so);
    s.so_hi = static_cast<uint32_t>(so >> 32);
    if (spoffs != nullptr) {
      const uint64_t po = static_cast<uint64_t>(spoffs[i]);
      s.po_lo = static_cast<uint32_t>(po);
      s.po_hi = static_cast<uint32_t>(po >> 32);
    }
  }
  return s;
}

template <int PIV, int N>
__device__ __forceinline__ void ragged_dispatch(const RaggedDesc& d, const RaggedSlots& sl, float* __restrict__ A,
                                                int* __restrict__ piv, float* lds, int lane) {
  if constexpr (N <= 32) {
    if (d.n == N) {
      if constexpr (N <= kRaggedQuadMaxN) {
        if (d.layout == 1) {
          ragged_tile<N, 4, PIV>(A, piv, sl, d.count, lds, lane);
          return;
        }
      }
      ragged_tile<N, 16, PIV>(A, piv, sl, d.count, lds, lane);
    } else {
      ragged_dispatch<PIV, N + 1>(d, sl, A, piv, lds, lane);
    }
  }
}

// Workgroup w (one wave) takes descriptors w, w + grid, ... below the
// descriptor total, so every wave reaches the exit.  With a grid of at
// least the descriptor count each workgroup does one descriptor and the
// hardware dispatcher starts a new wave whenever one ends; a resident-size
// grid makes the waves persistent.
template <int PIV>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2))) void lu_ragged_kernel(
    float* __restrict__ A, const int64_t* __restrict__ soffs, int* __restrict__ piv,
    const int64_t* __restrict__ spoffs, const int* __restrict__ hdr) {
  __shared__ __attribute__((aligned(16))) float lds[kRaggedLdsFloats];
  const int lane = threadIdx.x;
  const RaggedClasses cls = ragged_load_classes(hdr, lane);
  const int total = __builtin_amdgcn_readlane(cls.desc_first, 0);
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const RaggedDesc d = ragged_locate(cls, t, lane);
    const RaggedSlots sl = ragged_load_slots(soffs, spoffs, d, lane);
    // Opaque per descriptor: keeps the lane-derived addresses of all the
    // (n, layout) variants from being hoisted out of the loop, where they
    // would all stay live at once.
    int ln = lane;
    asm volatile("" : "+v"(ln));
    ragged_dispatch<PIV, 1>(d, sl, A, piv, lds, ln);
  }
}

}  // namespace lub
