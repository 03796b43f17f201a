// cmp_rocsolver.cpp -- the rocSOLVER comparator (include/lubatch_compare.h).
// Host code only; built into its own library so that liblubatch.so does not
// load rocSOLVER / rocBLAS.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "lubatch.h"
#include "lubatch_compare.h"

extern "C" {

int lubref_rocsolver_create(void** handle, void* stream) {
  if (handle == nullptr) return LUB_ENULL;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return LUB_EHIP;
  if (rocblas_set_stream(h, static_cast<hipStream_t>(stream)) != rocblas_status_success) {
    rocblas_destroy_handle(h);
    return LUB_EHIP;
  }
  *handle = h;
  return LUB_OK;
}

int lubref_rocsolver_destroy(void* handle) {
  if (handle == nullptr) return LUB_ENULL;
  return rocblas_destroy_handle(static_cast<rocblas_handle>(handle)) == rocblas_status_success ? LUB_OK : LUB_EHIP;
}

int lubref_rocsolver_inverse(void* handle, float* A, int* ipiv, int* info, int n, int batch) {
  if (handle == nullptr || A == nullptr || ipiv == nullptr || info == nullptr) return LUB_ENULL;
  if (n < 1) return LUB_EINVAL_N;
  if (batch < 0) return LUB_EINVAL_BATCH;
  if (batch == 0) return LUB_OK;
  rocblas_handle h = static_cast<rocblas_handle>(handle);
  const rocblas_stride sa = rocblas_stride(n) * n, sp = n;
  if (rocsolver_sgetrf_strided_batched(h, n, n, A, n, sa, ipiv, sp, info, batch) != rocblas_status_success)
    return LUB_EHIP;
  if (rocsolver_sgetri_strided_batched(h, n, A, n, sa, ipiv, sp, info, batch) != rocblas_status_success)
    return LUB_EHIP;
  return LUB_OK;
}

}  // extern "C"
