// lub_internal.h -- declarations shared by the kernel translation units and the
// C-ABI layer of liblubatch.so (not part of the public interface).
#pragma once

#include <stdint.h>

namespace lub {

constexpr int kMaxN = 32;
constexpr int kNumPivot = 5;   // NONE, PARALLEL, SERIAL, TREE, GEPP
constexpr int kNumLayout = 2;  // 0: 16 lanes per matrix (n <= 32), 1: 4 lanes per matrix (n <= 16)
constexpr int kLayoutLanes[kNumLayout] = {16, 4};
constexpr int kLayoutMaxN[kNumLayout] = {32, 16};

// Host handles of the launchable kernels, filled by the per-size-range
// translation units (each one instantiates lu_inverse_kernel<N, L, PIV> for
// its range of N and registers the host stubs here).
struct KernelTable {
  const void* plain[kMaxN + 1][kNumPivot][kNumLayout];
  // n % 4 == 0: the same kernels with the row padding flipped from its default
  const void* plain_flip[kMaxN + 1][kNumPivot][kNumLayout];
  const void* ragged[kNumPivot];
  // inverse + one refinement step (n = 17..32): [n][0] no pivoting, [n][1] first-max
  const void* refined[kMaxN + 1][2];
};

// Per-n launch choices shared by the plain and the ragged paths (set by the
// autotuner through lu_tuning_set, defaults in lub_api.hip).
struct Tuning {
  int8_t layout[kMaxN + 1];  // 0: 16-lane rows, 1: quads
  int8_t grab[kMaxN + 1];    // ragged: wave tiles per work item
  int8_t pad[kMaxN + 1];     // plain path, n % 4 == 0: LDS rows padded by 4 floats (1) or not (0)
};

// LDS row padding compiled into a kernel without the flip bit
// (Geo<N, L, V>::PAD for V & 2 == 0): rows of n % 8 == 0.
constexpr int8_t geo_pad(int n) { return (n > 0 && n % 8 == 0) ? 1 : 0; }
// Initial padding table of the plain path: every n % 4 == 0 but n = 28
// (measured on MI355X with the autotuner: padding n = 12 and n = 20 saves
// 4 % and 9 % of the kernel time, at n = 28 it costs 7 %).
constexpr int8_t initial_pad(int n) { return (n > 0 && n % 4 == 0 && n != 28) ? 1 : 0; }
Tuning current_tuning();
const KernelTable& kernel_table();

// The plain launch behind lu_batched_inverse_async64; info (nullable, one int
// per matrix) receives 0 or 1 + the first step with an exactly zero pivot.
int enqueue_plain(float* A, int* piv, int* info, int n, int64_t batch, int flags, void* stream);

}  // namespace lub

// One registration function per kernel translation unit.
#define LUB_REGISTER_DECL(lo, hi) extern "C" void lub_register_##lo##_##hi(lub::KernelTable* t)
LUB_REGISTER_DECL(1, 8);
LUB_REGISTER_DECL(9, 16);
LUB_REGISTER_DECL(17, 20);
LUB_REGISTER_DECL(21, 24);
LUB_REGISTER_DECL(25, 28);
LUB_REGISTER_DECL(29, 32);
extern "C" void lub_register_ragged_0(lub::KernelTable* t);
extern "C" void lub_register_ragged_1(lub::KernelTable* t);
extern "C" void lub_register_ragged_3(lub::KernelTable* t);
extern "C" void lub_register_ragged_4(lub::KernelTable* t);
extern "C" void lub_register_refined_17_24(lub::KernelTable* t);
extern "C" void lub_register_refined_25_32(lub::KernelTable* t);
