// lu_ragged_inst.hip -- instantiates the ragged persistent kernel for the
// pivot rule LUB_RAGGED_PIV (compiled once per rule so they build in
// parallel) and registers its host handle.
#include "lu_ragged.hpp"
#include "lub_internal.h"

#ifndef LUB_RAGGED_PIV
#error "LUB_RAGGED_PIV must be defined"
#endif

#define LUB_RCAT2(a, b) a##b
#define LUB_RCAT(a, b) LUB_RCAT2(a, b)

extern "C" void LUB_RCAT(lub_register_ragged_, LUB_RAGGED_PIV)(lub::KernelTable* t) {
  t->ragged[LUB_RAGGED_PIV] = reinterpret_cast<const void*>(&lub::lu_ragged_kernel<LUB_RAGGED_PIV>);
}
