/*
 * san_driver.c -- runs every function of lu_oracle.c once per n = 1..32 on
 * seeded random, tied, NaN and zero inputs, for the sanitizer build of the
 * oracle (make -C oracle sanitize: AddressSanitizer + UndefinedBehavior-
 * Sanitizer, host code only).  Test infrastructure; exits 0 when every call
 * returned its documented status.  The checks of the values themselves live
 * in tests/test_oracle.py; this driver only drives the code under the
 * sanitizers, including the OpenMP batch forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_mt19937_raw(uint32_t seed, uint32_t* out, size_t count);
void orc_mt19937_uniform(uint32_t seed, float* out, size_t count);
int orc_pivots(const float* A, int n, int rule, int* perm);
int orc_lu_factor_f32(const float* A, float* LU, int n, int rule, int* perm_out);
int orc_lu_inverse_f32(float* A, int n, int rule, int* perm_out);
int orc_batch_inverse_f32(float* A, int* piv, int n, long long batch, int rule, int threads);
int orc_inverse_f64(const float* A, double* X, int n, int* gepp_perm);
int orc_gepp_inverse_f32(float* A, int n, int* perm_out);
int orc_batch_gepp_f32(float* A, int* piv, int* info, int n, long long batch, int threads);
double orc_cond1(const float* A, int n);
long long orc_verify_inv(const float* A, const float* X, int n, long long batch, float thr, double* max_abs);
long long orc_verify_inv_threads(const float* A, const float* X, int n, long long batch, float thr,
                                 double* max_abs, int threads);

#define BATCH 9

static int fail(const char* what, int n) {
    fprintf(stderr, "san_driver: %s failed at n=%d\n", what, n);
    return 1;
}

int main(void) {
    static float a[BATCH * 32 * 32], w[BATCH * 32 * 32], lu[32 * 32];
    static double x64[32 * 32];
    static int piv[BATCH * 32], info[BATCH], perm[32];
    uint32_t raw[8];
    orc_mt19937_raw(42, raw, 8);
    if (raw[0] != 1608637542u) return fail("mt19937 seed 42", 0);
    for (int n = 1; n <= 32; ++n) {
        const int nn = n * n;
        orc_mt19937_uniform(100 + n, a, (size_t)BATCH * nn);
        for (int i = 0; i < nn; ++i) a[1 * nn + i] = (float)((i * 7) % 3);  /* ties */
        a[2 * nn] = NAN;                                                     /* NaN at (0, 0) */
        if (n > 1) a[3 * nn + n] = NAN;                                      /* NaN below the diagonal */
        memset(a + 4 * nn, 0, sizeof(float) * nn);                          /* singular */
        for (int rule = 0; rule <= 4; ++rule) {
            if (rule == 2) continue;
            for (int b = 0; b < BATCH; ++b) {
                if (orc_pivots(a + b * nn, n, rule, perm) != (rule == 4 ? -1 : 0)) return fail("orc_pivots", n);
                if (rule != 4 && orc_lu_factor_f32(a + b * nn, lu, n, rule, perm) != 0) return fail("factor", n);
            }
            if (rule == 4) {
                memcpy(w, a, sizeof(float) * BATCH * nn);
                if (orc_batch_gepp_f32(w, piv, info, n, BATCH, 0) != 0) return fail("batch gepp", n);
                memcpy(w, a, sizeof(float) * nn);
                (void)orc_gepp_inverse_f32(w, n, perm);
            } else {
                memcpy(w, a, sizeof(float) * BATCH * nn);
                if (orc_batch_inverse_f32(w, piv, n, BATCH, rule, 0) != 0) return fail("batch inverse", n);
                memcpy(w, a, sizeof(float) * nn);
                if (orc_lu_inverse_f32(w, n, rule, perm) != 0) return fail("inverse", n);
            }
        }
        for (int b = 0; b < BATCH; ++b) {
            (void)orc_inverse_f64(a + b * nn, x64, n, perm);
            (void)orc_cond1(a + b * nn, n);
        }
        memcpy(w, a, sizeof(float) * BATCH * nn);
        (void)orc_batch_inverse_f32(w, piv, n, BATCH, 1, 0);
        double m1 = 0.0, m2 = 0.0;
        long long c1 = orc_verify_inv(a, w, n, BATCH, 1e-3f, &m1);
        long long c2 = orc_verify_inv_threads(a, w, n, BATCH, 1e-3f, &m2, 0);
        if (c1 != c2) return fail("verify counts", n);
    }
    printf("san_driver: ok\n");
    return 0;
}
