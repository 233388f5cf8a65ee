/*
 * spgemm_oracle.c — CPU oracle for the CSR SpGEMM C = A * B (FP32).
 *
 * TEST INFRASTRUCTURE ONLY. This file restates the reference's scalar
 * algorithm so tests, the smoke check and the benchmark's CPU-baseline leg
 * can check and time the GPU library against it. Nothing in the product
 * library links to or calls it.
 *
 * Built with gcc -std=c11 -ffp-contract=off (oracle/Makefile): like the
 * reference's -std=c11 build, the accumulate is a rounded multiply followed
 * by a rounded add, never a fused multiply-add.
 *
 * What each function restates:
 *   oracle_symbolic_count  src/kernels/spgemm/rvsp_symbolic.c:13-88
 *                          (distinct columns of the union of the B rows a
 *                          row selects, Op_i = sum of |B_k|, exclusive scan,
 *                          INT32_MAX guard -> -6 = RVSP_ERROR_ALLOCATION_FAILED)
 *   oracle_symbolic_fill   src/kernels/spgemm/rvsp_symbolic.c:91-138 and
 *                          rvsp_sort.h:105-124 (columns ascending per row;
 *                          the sort method is not observable)
 *   oracle_numeric_f32     src/kernels/spgemm/gustavson_core_f32.inc:18-64
 *                          with accum_scalar_f32.c:19-71 (acc = acc + a*b,
 *                          A entries ascending, +0.0f start, cancelled
 *                          entries kept, drain in C's column order)
 *   oracle_numeric_f64     the same sums in double precision (textbook
 *                          FP64 reference, no counterpart in rv-sparse)
 *
 * The implementation is this project's own: a per-row stamp array instead
 * of the reference's mark/touched pair, and an in-place heap sort.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_OVERFLOW (-6)
#define ORACLE_NOMEM (-7)

/* ---- per-row column set via stamps: col is in the current row's set iff
 * stamp[col] == row + 1, so nothing is cleared between rows. */

static void sift_down(int32_t *x, int32_t n, int32_t i)
{
    for (;;) {
        int32_t l = 2 * i + 1, big = i;
        if (l < n && x[l] > x[big]) big = l;
        if (l + 1 < n && x[l + 1] > x[big]) big = l + 1;
        if (big == i) return;
        int32_t t = x[i];
        x[i] = x[big];
        x[big] = t;
        i = big;
    }
}

static void heap_sort(int32_t *x, int32_t n)
{
    for (int32_t i = n / 2 - 1; i >= 0; --i) sift_down(x, n, i);
    for (int32_t end = n - 1; end > 0; --end) {
        int32_t t = x[0];
        x[0] = x[end];
        x[end] = t;
        sift_down(x, end, 0);
    }
}

/*
 * Rows [r0, r1) of A. c_rp receives r1 - r0 + 1 entries starting at 0 (a
 * block-local exclusive scan). op_counts (optional) receives r1 - r0
 * entries. Returns 0, or ORACLE_OVERFLOW when the block's nnz exceeds
 * INT32_MAX.
 */
int oracle_symbolic_count(int32_t r0, int32_t r1, int32_t b_cols,
                          const int32_t *a_rp, const int32_t *a_ci,
                          const int32_t *b_rp, const int32_t *b_ci,
                          int32_t *c_rp, int64_t *op_counts, int64_t *total_out)
{
    int32_t *stamp = (int32_t *)calloc((size_t)(b_cols > 0 ? b_cols : 1), sizeof(int32_t));
    if (!stamp) return ORACLE_NOMEM;
    int64_t total = 0;
    c_rp[0] = 0;
    for (int32_t row = r0; row < r1; ++row) {
        int64_t ops = 0;
        int32_t distinct = 0;
        const int32_t tag = row + 1;
        for (int32_t ap = a_rp[row]; ap < a_rp[row + 1]; ++ap) {
            const int32_t k = a_ci[ap];
            ops += b_rp[k + 1] - b_rp[k];
            for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; ++bp) {
                const int32_t col = b_ci[bp];
                if (stamp[col] != tag) {
                    stamp[col] = tag;
                    ++distinct;
                }
            }
        }
        if (op_counts) op_counts[row - r0] = ops;
        total += distinct;
        if (total > INT32_MAX) {
            free(stamp);
            return ORACLE_OVERFLOW;
        }
        c_rp[row - r0 + 1] = (int32_t)total;
    }
    free(stamp);
    *total_out = total;
    return 0;
}

/* Columns of C's rows [r0, r1), ascending per row, at c_rp (block-local). */
int oracle_symbolic_fill(int32_t r0, int32_t r1, int32_t b_cols,
                         const int32_t *a_rp, const int32_t *a_ci,
                         const int32_t *b_rp, const int32_t *b_ci,
                         const int32_t *c_rp, int32_t *c_ci)
{
    int32_t *stamp = (int32_t *)calloc((size_t)(b_cols > 0 ? b_cols : 1), sizeof(int32_t));
    if (!stamp) return ORACLE_NOMEM;
    for (int32_t row = r0; row < r1; ++row) {
        int32_t *out = c_ci + c_rp[row - r0];
        int32_t n = 0;
        const int32_t tag = row + 1;
        for (int32_t ap = a_rp[row]; ap < a_rp[row + 1]; ++ap) {
            const int32_t k = a_ci[ap];
            for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; ++bp) {
                const int32_t col = b_ci[bp];
                if (stamp[col] != tag) {
                    stamp[col] = tag;
                    out[n++] = col;
                }
            }
        }
        heap_sort(out, n);
    }
    free(stamp);
    return 0;
}

/*
 * Values of C's rows [r0, r1). `variant` exists so tests can plant a
 * mistake in the oracle and show the checks catch it:
 *   0  the reference order and rounding (the oracle proper)
 *   1  fused multiply-add instead of multiply then add
 *   2  A entries of each row taken in descending order
 */
int oracle_numeric_f32_variant(int32_t r0, int32_t r1, int32_t b_cols,
                               const int32_t *a_rp, const int32_t *a_ci,
                               const float *a_val, const int32_t *b_rp,
                               const int32_t *b_ci, const float *b_val,
                               const int32_t *c_rp, const int32_t *c_ci,
                               float *c_val, int variant)
{
    float *acc = (float *)calloc((size_t)(b_cols > 0 ? b_cols : 1), sizeof(float));
    if (!acc) return ORACLE_NOMEM;
    for (int32_t row = r0; row < r1; ++row) {
        const int32_t s = a_rp[row], e = a_rp[row + 1];
        for (int32_t t = 0; t < e - s; ++t) {
            const int32_t ap = (variant == 2) ? (e - 1 - t) : (s + t);
            const int32_t k = a_ci[ap];
            const float a = a_val[ap];
            for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; ++bp) {
                float *slot = &acc[b_ci[bp]];
                if (variant == 1) {
                    *slot = fmaf(a, b_val[bp], *slot);
                } else {
                    const float prod = a * b_val[bp];
                    *slot = *slot + prod;
                }
            }
        }
        for (int32_t p = c_rp[row - r0]; p < c_rp[row - r0 + 1]; ++p) {
            c_val[p] = acc[c_ci[p]];
            acc[c_ci[p]] = 0.0f;
        }
    }
    free(acc);
    return 0;
}

int oracle_numeric_f32(int32_t r0, int32_t r1, int32_t b_cols,
                       const int32_t *a_rp, const int32_t *a_ci, const float *a_val,
                       const int32_t *b_rp, const int32_t *b_ci, const float *b_val,
                       const int32_t *c_rp, const int32_t *c_ci, float *c_val)
{
    return oracle_numeric_f32_variant(r0, r1, b_cols, a_rp, a_ci, a_val, b_rp, b_ci,
                                      b_val, c_rp, c_ci, c_val, 0);
}

/*
 * FP64 oracle for the same rows: exact products of the FP32 inputs summed in
 * double. abs_sum (optional) receives sum_k |a_ik * b_kj| per entry, the
 * scale of a condition-aware relative error.
 */
int oracle_numeric_f64(int32_t r0, int32_t r1, int32_t b_cols,
                       const int32_t *a_rp, const int32_t *a_ci, const float *a_val,
                       const int32_t *b_rp, const int32_t *b_ci, const float *b_val,
                       const int32_t *c_rp, const int32_t *c_ci, double *c_val,
                       double *abs_sum)
{
    const size_t n = (size_t)(b_cols > 0 ? b_cols : 1);
    double *acc = (double *)calloc(n, sizeof(double));
    double *mag = (double *)calloc(n, sizeof(double));
    if (!acc || !mag) {
        free(acc);
        free(mag);
        return ORACLE_NOMEM;
    }
    for (int32_t row = r0; row < r1; ++row) {
        for (int32_t ap = a_rp[row]; ap < a_rp[row + 1]; ++ap) {
            const int32_t k = a_ci[ap];
            const double a = (double)a_val[ap];
            for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; ++bp) {
                const double p = a * (double)b_val[bp];
                acc[b_ci[bp]] += p;
                mag[b_ci[bp]] += fabs(p);
            }
        }
        for (int32_t p = c_rp[row - r0]; p < c_rp[row - r0 + 1]; ++p) {
            const int32_t col = c_ci[p];
            c_val[p] = acc[col];
            if (abs_sum) abs_sum[p] = mag[col];
            acc[col] = 0.0;
            mag[col] = 0.0;
        }
    }
    free(acc);
    free(mag);
    return 0;
}

/*
 * Values of the legacy dtypes on the structure (c_rp, c_ci) of rows
 * [r0, r1), zeros included (the caller drops them): the reference's legacy
 * kernels, src/kernels/spgemm/legacy/csr_scalar_f64.c:214-236 (FP64
 * product, then FP64 sum) and csr_scalar_i8.c:140-161 (int32 product and
 * sum; kept here in unsigned arithmetic so a wrap is defined), both over A's
 * entries in order and each B row's entries in order. variant 1 takes A's
 * entries in descending order (a planted mistake for the tests).
 */
int oracle_numeric_legacy(int32_t r0, int32_t r1, int32_t b_cols, int i8,
                          const int32_t *a_rp, const int32_t *a_ci, const void *a_val,
                          const int32_t *b_rp, const int32_t *b_ci, const void *b_val,
                          const int32_t *c_rp, const int32_t *c_ci, void *c_val,
                          int variant)
{
    const size_t n = (size_t)(b_cols > 0 ? b_cols : 1);
    double *accd = i8 ? NULL : (double *)calloc(n, sizeof(double));
    uint32_t *acci = i8 ? (uint32_t *)calloc(n, sizeof(uint32_t)) : NULL;
    if (!accd && !acci) return ORACLE_NOMEM;
    for (int32_t row = r0; row < r1; ++row) {
        const int32_t s = a_rp[row], e = a_rp[row + 1];
        for (int32_t t = 0; t < e - s; ++t) {
            const int32_t ap = variant == 1 ? e - 1 - t : s + t;
            const int32_t k = a_ci[ap];
            for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; ++bp) {
                const int32_t col = b_ci[bp];
                if (i8) {
                    const int32_t prod = (int32_t)((const int8_t *)a_val)[ap] *
                                         (int32_t)((const int8_t *)b_val)[bp];
                    acci[col] += (uint32_t)prod;
                } else {
                    const double prod = ((const double *)a_val)[ap] * ((const double *)b_val)[bp];
                    accd[col] = accd[col] + prod;
                }
            }
        }
        for (int32_t p = c_rp[row - r0]; p < c_rp[row - r0 + 1]; ++p) {
            const int32_t col = c_ci[p];
            if (i8) {
                ((int32_t *)c_val)[p] = (int32_t)acci[col];
                acci[col] = 0;
            } else {
                ((double *)c_val)[p] = accd[col];
                accd[col] = 0.0;
            }
        }
    }
    free(accd);
    free(acci);
    return 0;
}
