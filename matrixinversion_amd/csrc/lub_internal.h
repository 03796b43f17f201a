// lub_internal.h -- declarations shared by the kernel translation units and the
// C-ABI layer of liblubatch.so (not part of the public interface).
#pragma once

#include <stdint.h>

namespace lub {

constexpr int kMaxN = 32;
constexpr int kNumPivot = 4;   // NONE, PARALLEL, SERIAL, TREE
constexpr int kNumLayout = 2;  // 0: 16 lanes per matrix (n <= 32), 1: 4 lanes per matrix (n <= 16)
constexpr int kLayoutLanes[kNumLayout] = {16, 4};
constexpr int kLayoutMaxN[kNumLayout] = {32, 16};

// Host handles of the launchable kernels, filled by the per-size-range
// translation units (each one instantiates lu_inverse_kernel<N, L, PIV> for
// its range of N and registers the host stubs here).
struct KernelTable {
  const void* plain[kMaxN + 1][kNumPivot][kNumLayout];
};

}  // namespace lub

// One registration function per kernel translation unit.
#define LUB_REGISTER_DECL(lo, hi) extern "C" void lub_register_##lo##_##hi(lub::KernelTable* t)
LUB_REGISTER_DECL(1, 8);
LUB_REGISTER_DECL(9, 16);
LUB_REGISTER_DECL(17, 20);
LUB_REGISTER_DECL(21, 24);
LUB_REGISTER_DECL(25, 28);
LUB_REGISTER_DECL(29, 32);
