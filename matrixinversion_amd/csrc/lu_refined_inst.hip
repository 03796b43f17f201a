// lu_refined_inst.hip -- instantiates lu_refined_kernel<N, PIV> (inverse +
// one refinement step in one kernel) for N in [LUB_RF_LO, LUB_RF_HI] and the
// pivot rules none and first-max, and registers the host handles.  Compiled
// once per size range so the units build in parallel.
#include "lu_refined.hpp"
#include "lub_internal.h"

#ifndef LUB_RF_LO
#error "LUB_RF_LO / LUB_RF_HI must be defined"
#endif

#define LUB_RCAT3(a, b, c) a##b##_##c
#define LUB_RCATX(a, b, c) LUB_RCAT3(a, b, c)
#define LUB_RF_REGISTER_NAME LUB_RCATX(lub_register_refined_, LUB_RF_LO, LUB_RF_HI)

namespace {

template <int N>
void register_range(lub::KernelTable* t) {
  if constexpr (N <= LUB_RF_HI) {
    t->refined[N][0] = reinterpret_cast<const void*>(&lub::lu_refined_kernel<N, lub::kPivNone>);
    t->refined[N][1] = reinterpret_cast<const void*>(&lub::lu_refined_kernel<N, lub::kPivParallel>);
    register_range<N + 1>(t);
  }
}

}  // namespace

extern "C" void LUB_RF_REGISTER_NAME(lub::KernelTable* t) { register_range<LUB_RF_LO>(t); }
