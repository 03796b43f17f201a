/*
 * lu_oracle.c -- CPU oracle for batched small-matrix LU inversion.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and bench.py's cpu_baseline leg load it; the product path
 * (matrixinversion_amd/csrc) never links, calls or falls back to it.
 *
 * It restates, in plain C, the arithmetic of the reference's six CUDA
 * variants of batched_lu_subwarp (sumukhashridhar/matrixInversion):
 *
 *   orc_pivots_firstmax   serial_pivot/luBatchedInplace.cuh:22-36 (find_pivot,
 *                         strict '>' so the first maximum |a| wins) with the
 *                         perm swap of serial_pivot/luBatchedInplace.cuh:132-141;
 *                         the same rule as pivotedA, parallel_pivot/verify.hpp:106-155.
 *   orc_pivots_tree       parallel_pivot/luBatchedInplace.cuh:12-44 (per-lane
 *                         candidate, then shared-memory tree with integer-halved
 *                         strides n/2..1, left slot kept unless strictly less)
 *                         and the swap at :161-172, with threadsPerMatrix = n
 *                         (parallel_pivot/luBatchedInplace.cu:8).
 *   orc_lu_factor_f32     Doolittle comp_U / comp_L (parallel_pivot/
 *                         luBatchedInplace.cuh:60-82, the same in every variant)
 *   orc_lu_inverse_f32    the factorisation, then comp_inv
 *                         (:84-125): column c solves L y = P e_c
 *                         with (P e_c)_i = [perm[i] == c], then U x = y; the
 *                         no-pivot RHS is e_c (templated/luBatchedInplace.cuh:56).
 *   orc_verify_inv        verifyInv, parallel_pivot/verify.hpp:50-103: an
 *                         inversion counts as correct when every entry of A*X
 *                         is within 1e-3 of the identity (FP32 accumulation).
 *   orc_cond1             kappa_1 = ||A||_1 ||A^-1||_1 (the quantity of
 *                         calc_cond_num, verify.hpp:245-338), here in FP64.
 *   orc_inverse_f64       an independent FP64 Gauss elimination with partial
 *                         pivoting on the updated column (LAPACK getrf rule),
 *                         the accuracy reference for FP32 results.
 *   orc_gepp_inverse_f32  FP32 Gauss-Jordan with that rule (no counterpart in
 *                         the reference, whose "partial pivoting" searches
 *                         the un-updated column): the restatement the GPU's
 *                         LUB_PIVOT_GEPP mode is checked against.
 *   orc_mt19937_*         MT19937 (Matsumoto & Nishimura 1998, init_genrand
 *                         seeding) and the pinned float mapping
 *                         a = 2 * ((x >> 8) * 2^-24) - 1, exact in FP32.
 *
 * Pivot vectors use the reference's `pivots` meaning: perm[p] is the original
 * row that sits at position p after factorisation, so P A = L U with
 * (P A)[p][:] = A[perm[p]][:].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXN 64

enum { ORC_PIV_NONE = 0, ORC_PIV_FIRSTMAX = 1, ORC_PIV_TREE = 3, ORC_PIV_GEPP = 4 };

/* ---------------------------------------------------------------- MT19937 */

typedef struct {
    uint32_t mt[624];
    int idx;
} orc_mt19937;

static void mt_seed(orc_mt19937* s, uint32_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->idx = 624;
}

static void mt_twist(orc_mt19937* s) {
    for (int i = 0; i < 624; ++i) {
        uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = s->mt[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        s->mt[i] = v;
    }
    s->idx = 0;
}

static uint32_t mt_next(orc_mt19937* s) {
    if (s->idx >= 624) mt_twist(s);
    uint32_t y = s->mt[s->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* Raw 32-bit outputs of std::mt19937(seed). */
void orc_mt19937_raw(uint32_t seed, uint32_t* out, size_t count) {
    orc_mt19937 s;
    mt_seed(&s, seed);
    for (size_t i = 0; i < count; ++i) out[i] = mt_next(&s);
}

/* U[-1,1) floats: a = 2*((x>>8)*2^-24) - 1 (exact in FP32). */
void orc_mt19937_uniform(uint32_t seed, float* out, size_t count) {
    orc_mt19937 s;
    mt_seed(&s, seed);
    const float scale = 1.0f / 16777216.0f;
    for (size_t i = 0; i < count; ++i) {
        float u = (float)(mt_next(&s) >> 8) * scale;
        out[i] = 2.0f * u - 1.0f;
    }
}

/* --------------------------------------------------------------- pivots */

/* First-max rule on the un-updated column: serial_pivot/luBatchedInplace.cuh:22-36.
 * At step k the column-k entries of the rows at positions >= k have not been
 * touched by comp_U / comp_L yet, so the search reads original values. */
void orc_pivots_firstmax(const float* A, int n, int* perm) {
    for (int p = 0; p < n; ++p) perm[p] = p;
    for (int k = 0; k < n; ++k) {
        int best = k;
        float bv = fabsf(A[(size_t)perm[k] * n + k]);
        for (int p = k + 1; p < n; ++p) {
            float v = fabsf(A[(size_t)perm[p] * n + k]);
            if (v > bv) { bv = v; best = p; }
        }
        if (best != k) { int t = perm[k]; perm[k] = perm[best]; perm[best] = t; }
    }
}

/* Tree rule of parallel_pivot/luBatchedInplace.cuh:12-44 with threadsPerMatrix = n.
 * Slot t starts at (|a_kk|, k) and takes row k+1+t when strictly larger; the
 * reduction walks stride = n/2, n/4, ..., 1 (integer halving), so for n not a
 * power of two some slots are never merged into slot 0. */
void orc_pivots_tree(const float* A, int n, int* perm) {
    float val[ORC_MAXN];
    int idx[ORC_MAXN];
    for (int p = 0; p < n; ++p) perm[p] = p;
    for (int k = 0; k < n; ++k) {
        float vk = fabsf(A[(size_t)perm[k] * n + k]);
        for (int t = 0; t < n; ++t) {
            val[t] = vk;
            idx[t] = k;
            for (int i = k + 1 + t; i < n; i += n) {
                float v = fabsf(A[(size_t)perm[i] * n + k]);
                if (v > val[t]) { val[t] = v; idx[t] = i; }
            }
        }
        for (int stride = n / 2; stride > 0; stride >>= 1)
            for (int t = 0; t < stride; ++t)
                if (val[t] < val[t + stride]) { val[t] = val[t + stride]; idx[t] = idx[t + stride]; }
        int best = idx[0];
        if (best != k) { int t = perm[k]; perm[k] = perm[best]; perm[best] = t; }
    }
}

/* Pivot vector by rule; GEPP pivots come from orc_inverse_f64. */
int orc_pivots(const float* A, int n, int rule, int* perm) {
    if (n < 1 || n > ORC_MAXN) return -1;
    switch (rule) {
    case ORC_PIV_NONE:
        for (int p = 0; p < n; ++p) perm[p] = p;
        return 0;
    case ORC_PIV_FIRSTMAX: orc_pivots_firstmax(A, n, perm); return 0;
    case ORC_PIV_TREE: orc_pivots_tree(A, n, perm); return 0;
    default: return -1;
    }
}

/* ------------------------------------------------------- FP32 inversion */

/* FP32 Doolittle factorisation of P*A in the reference's order: rows swapped
 * by `rule`, then for each k comp_U (row k of U, parallel_pivot/
 * luBatchedInplace.cuh:60-70) and comp_L (column k of L, :72-82); LU receives
 * L (unit diagonal implied) below and U on/above the diagonal, as the
 * reference leaves them in shared memory before comp_inv.  perm (may be NULL)
 * receives the pivot vector.  Returns 0. */
int orc_lu_factor_f32(const float* A, float* LU, int n, int rule, int* perm_out) {
    int perm[ORC_MAXN];
    if (n < 1 || n > ORC_MAXN) return -1;
    if (orc_pivots(A, n, rule, perm) != 0) return -1;
    float* M = LU;
    for (int p = 0; p < n; ++p)
        memcpy(&M[p * n], &A[(size_t)perm[p] * n], sizeof(float) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        for (int j = k; j < n; ++j) { /* comp_U, row k */
            float sum = 0.0f;
            for (int l = 0; l < k; ++l) sum += M[k * n + l] * M[l * n + j];
            M[k * n + j] = M[k * n + j] - sum;
        }
        for (int i = k + 1; i < n; ++i) { /* comp_L, column k */
            float sum = 0.0f;
            for (int l = 0; l < k; ++l) sum += M[i * n + l] * M[l * n + k];
            M[i * n + k] = (M[i * n + k] - sum) / M[k * n + k];
        }
    }
    if (perm_out) memcpy(perm_out, perm, sizeof(int) * (size_t)n);
    return 0;
}

/* In-place FP32 inverse of one n x n row-major matrix following the reference's
 * arithmetic order: orc_lu_factor_f32, then one comp_inv per column
 * (parallel_pivot/luBatchedInplace.cuh:84-125) whose results are written back
 * only after every column has been solved (the reference's lanes finish their
 * solves before the store loop).  perm (may be NULL) receives the pivot
 * vector.  Returns 0. */
int orc_lu_inverse_f32(float* A, int n, int rule, int* perm_out) {
    int perm[ORC_MAXN];
    float M[ORC_MAXN * ORC_MAXN];
    float X[ORC_MAXN * ORC_MAXN];
    if (orc_lu_factor_f32(A, M, n, rule, perm) != 0) return -1;
    for (int c = 0; c < n; ++c) { /* comp_inv for column c */
        float y[ORC_MAXN], x[ORC_MAXN];
        for (int i = 0; i < n; ++i) {
            float b = (perm[i] == c) ? 1.0f : 0.0f;
            float sum = 0.0f;
            for (int j = 0; j < i; ++j) sum += M[i * n + j] * y[j];
            y[i] = b - sum;
        }
        for (int i = n - 1; i >= 0; --i) {
            float sum = 0.0f;
            for (int j = i + 1; j < n; ++j) sum += M[i * n + j] * x[j];
            x[i] = (y[i] - sum) / M[i * n + i];
        }
        for (int i = 0; i < n; ++i) X[i * n + c] = x[i];
    }
    memcpy(A, X, sizeof(float) * (size_t)n * (size_t)n);
    if (perm_out) memcpy(perm_out, perm, sizeof(int) * (size_t)n);
    return 0;
}

/* Batched form over contiguous matrices (stride n*n); piv may be NULL.
 * threads <= 0 uses the OpenMP default. */
int orc_batch_inverse_f32(float* A, int* piv, int n, long long batch, int rule, int threads) {
    if (n < 1 || n > ORC_MAXN || batch < 0) return -1;
    int err = 0;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static) reduction(| : err)
#endif
    for (long long b = 0; b < batch; ++b)
        err |= orc_lu_inverse_f32(A + (size_t)b * n * n, n, rule, piv ? piv + (size_t)b * n : NULL);
    (void)threads;
    return err ? -1 : 0;
}

int orc_openmp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------- FP64 reference */

/* FP64 inverse by Gauss elimination with partial pivoting on the updated column
 * (the LAPACK getrf rule, first maximum on ties).  X may alias nothing.  gepp_perm
 * (may be NULL) receives the getrf-style permutation in the perm[p] = source-row
 * convention.  Returns 0, or k+1 when the k-th pivot is exactly zero. */
int orc_inverse_f64(const float* A, double* X, int n, int* gepp_perm) {
    double M[ORC_MAXN * ORC_MAXN];
    int perm[ORC_MAXN];
    int info = 0;
    if (n < 1 || n > ORC_MAXN) return -1;
    for (int i = 0; i < n * n; ++i) M[i] = (double)A[i];
    for (int p = 0; p < n; ++p) perm[p] = p;
    for (int k = 0; k < n; ++k) {
        int best = k;
        double bv = fabs(M[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(M[i * n + k]) > bv) { bv = fabs(M[i * n + k]); best = i; }
        if (best != k) {
            for (int j = 0; j < n; ++j) { double t = M[k * n + j]; M[k * n + j] = M[best * n + j]; M[best * n + j] = t; }
            int t = perm[k]; perm[k] = perm[best]; perm[best] = t;
        }
        if (M[k * n + k] == 0.0) { if (!info) info = k + 1; continue; }
        for (int i = k + 1; i < n; ++i) {
            double l = M[i * n + k] / M[k * n + k];
            M[i * n + k] = l;
            for (int j = k + 1; j < n; ++j) M[i * n + j] -= l * M[k * n + j];
        }
    }
    if (gepp_perm) memcpy(gepp_perm, perm, sizeof(int) * (size_t)n);
    if (info) return info;
    for (int c = 0; c < n; ++c) {
        double y[ORC_MAXN];
        for (int i = 0; i < n; ++i) {
            double s = (perm[i] == c) ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) s -= M[i * n + j] * y[j];
            y[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int j = i + 1; j < n; ++j) s -= M[i * n + j] * X[j * n + c];
            X[i * n + c] = s / M[i * n + i];
        }
    }
    return 0;
}

/* FP32 Gauss-Jordan inverse with partial pivoting on the updated column
 * (LAPACK getrf's rule: largest |a| among the rows not yet pivotal, ties to
 * the smallest current position, where step k exchanges the positions of
 * rows k and p as getrf's row swap would).  Rows are not moved:
 *   step k, pivot row p, d = S[p][k], r = 1/d:
 *     every row i != p:  S[i][c] = fma(-(S[i][k] * r), S[p][c], S[i][c]), c != k
 *     S[p][k] = -S[p][k];  row scale rho[p] = r;  column scale sigma[k] = -r
 * Row p keeps its unnormalised values and column k its unscaled ones; the
 * scales commute with every later row operation, so the true matrix is
 * rho[i] * S[i][c] * sigma[c].  After the last step row i = perm[k] holds row
 * k of (P A)^-1, and (P A)^-1 [k][c] = A^-1 [k][perm[c]].  The kernel forms r
 * as rcp + one Newton step; here it is a division (equal but for a rare last
 * bit).  A NaN in the pivot column is taken as its largest entry.  perm (may
 * be NULL) receives perm[p] = original row at position p.  Returns 0, or
 * k + 1 for the first pivot without a finite reciprocal. */
static uint32_t orc_abs_bits(float x) {
    uint32_t u;
    memcpy(&u, &x, sizeof(u));
    return u & 0x7fffffffu;
}

int orc_gepp_inverse_f32(float* A, int n, int* perm_out) {
    float S[ORC_MAXN * ORC_MAXN], rho[ORC_MAXN], sig[ORC_MAXN];
    int pos[ORC_MAXN], rem[ORC_MAXN], perm[ORC_MAXN];
    int info = 0;
    if (n < 1 || n > ORC_MAXN) return -1;
    memcpy(S, A, sizeof(float) * (size_t)n * (size_t)n);
    for (int i = 0; i < n; ++i) { pos[i] = i; rem[i] = 1; rho[i] = 1.0f; sig[i] = 1.0f; }
    for (int k = 0; k < n; ++k) {
        uint32_t m = 0;
        for (int i = 0; i < n; ++i)
            if (rem[i] && orc_abs_bits(S[i * n + k]) > m) m = orc_abs_bits(S[i * n + k]);
        int p = -1;
        for (int i = 0; i < n; ++i)
            if (rem[i] && orc_abs_bits(S[i * n + k]) == m && (p < 0 || pos[i] < pos[p])) p = i;
        for (int i = 0; i < n; ++i)
            if (pos[i] == k && i != p) pos[i] = pos[p];
        pos[p] = k;
        rem[p] = 0;
        const float d = S[p * n + k];
        const float r = 1.0f / d;
        if (!isfinite(r) && !info) info = k + 1;
        for (int i = 0; i < n; ++i) {
            if (i == p) continue;
            const float mlt = S[i * n + k] * r;
            for (int c = 0; c < n; ++c)
                if (c != k) S[i * n + c] = fmaf(-mlt, S[p * n + c], S[i * n + c]);
        }
        S[p * n + k] = -d;
        rho[p] = r;
        sig[k] = -r;
    }
    for (int i = 0; i < n; ++i) perm[pos[i]] = i;
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < n; ++c) A[(size_t)pos[i] * n + perm[c]] = (S[i * n + c] * sig[c]) * rho[i];
    if (perm_out) memcpy(perm_out, perm, sizeof(int) * (size_t)n);
    return info;
}

/* Batched form of orc_gepp_inverse_f32 (stride n*n); piv and info may be NULL. */
int orc_batch_gepp_f32(float* A, int* piv, int* info, int n, long long batch, int threads) {
    if (n < 1 || n > ORC_MAXN || batch < 0) return -1;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (long long b = 0; b < batch; ++b) {
        int f = orc_gepp_inverse_f32(A + (size_t)b * n * n, n, piv ? piv + (size_t)b * n : NULL);
        if (info) info[b] = f;
    }
    (void)threads;
    return 0;
}

/* kappa_1 in FP64; returns +inf for an exactly singular matrix. */
double orc_cond1(const float* A, int n) {
    double X[ORC_MAXN * ORC_MAXN];
    if (n < 1 || n > ORC_MAXN) return -1.0;
    if (orc_inverse_f64(A, X, n, NULL) != 0) return INFINITY;
    double na = 0.0, nx = 0.0;
    for (int j = 0; j < n; ++j) {
        double sa = 0.0, sx = 0.0;
        for (int i = 0; i < n; ++i) { sa += fabs((double)A[i * n + j]); sx += fabs(X[i * n + j]); }
        if (sa > na) na = sa;
        if (sx > nx) nx = sx;
    }
    return na * nx;
}

/* ----------------------------------------------------------- checking */

/* verifyInv semantics (parallel_pivot/verify.hpp:50-103): for every matrix and
 * every (i, j), r = sum_l A[j][l] * X[l][i] accumulated in FP32 in l order; the
 * matrix counts as correct when all diagonal r are within thr of 1 and all
 * off-diagonal r within thr of 0.  Returns the number of correct inversions;
 * max_abs (may be NULL) receives max |A X - I| over the batch computed in FP64. */
long long orc_verify_inv_threads(const float* A, const float* X, int n, long long batch, float thr,
                                 double* max_abs, int threads);

long long orc_verify_inv(const float* A, const float* X, int n, long long batch, float thr, double* max_abs) {
    return orc_verify_inv_threads(A, X, n, batch, thr, max_abs, 1);
}

/* The same count over `threads` OpenMP threads (matrices split between them;
 * each matrix is still checked in the reference's order).  threads <= 0: all. */
long long orc_verify_inv_threads(const float* A, const float* X, int n, long long batch, float thr,
                                 double* max_abs, int threads) {
    long long correct = 0;
    double worst = 0.0;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static) reduction(+ : correct) reduction(max : worst)
#endif
    for (long long b = 0; b < batch; ++b) {
        const float* a = A + (size_t)b * n * n;
        const float* x = X + (size_t)b * n * n;
        int ok = 1;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                float r = 0.0f;
                double rd = 0.0;
                for (int l = 0; l < n; ++l) {
                    r += a[j * n + l] * x[l * n + i];
                    rd += (double)a[j * n + l] * (double)x[l * n + i];
                }
                float target = (i == j) ? 1.0f : 0.0f;
                if (!(fabsf(r - target) < thr)) ok = 0;
                double e = fabs(rd - (double)target);
                if (isnan(e)) e = INFINITY; /* a NaN entry is the worst error */
                if (e > worst) worst = e;
            }
        correct += ok;
    }
    if (max_abs) *max_abs = worst;
    (void)threads;
    return correct;
}
