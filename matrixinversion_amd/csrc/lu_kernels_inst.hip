// lu_kernels_inst.hip -- instantiates lu_inverse_kernel<N, L, PIV> for
// N in [LUB_N_LO, LUB_N_HI] (L = 16 always, L = 4 for N <= 16) and registers
// the host handles.  The build
// compiles this file once per size range so the (large, fully unrolled)
// kernels compile in parallel.
#include "lu_kernel.hpp"
#include "lub_internal.h"

#ifndef LUB_N_LO
#error "LUB_N_LO / LUB_N_HI must be defined"
#endif

#define LUB_CAT2(a, b, c) a##b##_##c
#define LUB_CAT(a, b, c) LUB_CAT2(a, b, c)
#define LUB_REGISTER_NAME LUB_CAT(lub_register_, LUB_N_LO, LUB_N_HI)

namespace {

template <int N, int L, int V>
void register_variant(const void* (&slot)[lub::kMaxN + 1][lub::kNumPivot][lub::kNumLayout], int layout) {
  slot[N][lub::kPivNone][layout] = reinterpret_cast<const void*>(&lub::lu_inverse_kernel<N, L, lub::kPivNone, V>);
  slot[N][lub::kPivParallel][layout] =
      reinterpret_cast<const void*>(&lub::lu_inverse_kernel<N, L, lub::kPivParallel, V>);
  // The serial rule's pivots are the first-max rule's: the same kernel (DPP
  // argmax, with the lane-serial scan on ties and NaN) serves both flags.
  slot[N][lub::kPivSerial][layout] =
      reinterpret_cast<const void*>(&lub::lu_inverse_kernel<N, L, lub::kPivParallel, V>);
  slot[N][lub::kPivTree][layout] = reinterpret_cast<const void*>(&lub::lu_inverse_kernel<N, L, lub::kPivTree, V>);
  slot[N][lub::kPivGepp][layout] = reinterpret_cast<const void*>(&lub::lu_inverse_kernel<N, L, lub::kPivGepp, V>);
}

// The default row padding, and for n % 4 == 0 the flipped one.
template <int N, int L>
void register_layout(lub::KernelTable* t, int layout) {
  register_variant<N, L, 0>(t->plain, layout);
  if constexpr (N % 4 == 0) register_variant<N, L, lub::kFlipPad>(t->plain_flip, layout);
}

template <int N>
void register_size(lub::KernelTable* t) {
  register_layout<N, 16>(t, 0);
  if constexpr (N <= 16) register_layout<N, 4>(t, 1);
}

template <int N>
void register_range(lub::KernelTable* t) {
  if constexpr (N <= LUB_N_HI) {
    register_size<N>(t);
    register_range<N + 1>(t);
  }
}

}  // namespace

extern "C" void LUB_REGISTER_NAME(lub::KernelTable* t) { register_range<LUB_N_LO>(t); }
