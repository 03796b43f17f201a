/*
 * lubatch_compare.h -- benchmark-only comparator, exported by
 * liblubref_rocsolver.so (not by liblubatch.so, which does not depend on
 * rocSOLVER).  The reference compares against cuBLAS's batched getrf/getri
 * in a benchmark that is not part of its sources (README.md:63-69); on
 * MI355X the vendor counterpart is rocSOLVER's strided-batched
 * getrf + getri (LAPACK partial pivoting).
 *
 * Matrices are row-major n x n, contiguous, device memory, inverted in place.
 * rocSOLVER is column-major: it sees A^T and returns (A^T)^-1 = (A^-1)^T in
 * column-major order, which is A^-1 in row-major order, so no transposition
 * is needed.  ipiv / info are rocSOLVER's (pivots of A^T, 1-based).
 */
#ifndef LUBATCH_COMPARE_H
#define LUBATCH_COMPARE_H

#ifdef __cplusplus
extern "C" {
#endif

/* A rocBLAS handle bound to `stream` (NULL: the default stream). */
int lubref_rocsolver_create(void** handle, void* stream);
int lubref_rocsolver_destroy(void* handle);

/* getrf_strided_batched + getri_strided_batched, asynchronous on the
 * handle's stream.  ipiv: batch x n ints, info: batch ints (device).
 * rocSOLVER reads column-major: on a row-major batch it factors A^T, so the
 * result, read row-major, is still A^-1 ((A^T)^-1 = (A^-1)^T) while ipiv and
 * info describe A^T.  Pass A^T to get A's own getrf interchanges. */
int lubref_rocsolver_inverse(void* handle, float* A, int* ipiv, int* info, int n, int batch);

#ifdef __cplusplus
}
#endif

#endif /* LUBATCH_COMPARE_H */
