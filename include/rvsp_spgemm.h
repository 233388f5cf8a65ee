/*
 * rvsp_spgemm.h — harness-level exports that the reference keeps in its
 * internal header src/kernels/spgemm/rvsp_spgemm.h:34-53 and that its bench
 * (bench/bench.c:166-196, 771) and CSR checker (bench/csr_check.c:93) call.
 * They are exported here so that harness links unchanged.
 */
#ifndef RVSP_SPGEMM_H
#define RVSP_SPGEMM_H

#include <stddef.h>
#include <stdint.h>

#include "rv_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Host workspace size for the raw one-shot kernel, a function of b_cols only
 * (reference: src/kernels/spgemm/rvsp_support.c:11-29 and the four 64-byte
 * aligned regions of rvsp_common.h:48-60). The HIP kernels keep their scratch
 * on the device, but the size is reported as the reference reports it so a
 * harness that sizes its buffer from it behaves the same.
 */
rvsp_status_t rvsp_spgemm_buffer_size(int32_t b_cols, size_t *bytes_out);

/* Canonical-CSR verdicts (reference: src/kernels/spgemm/rvsp_spgemm.h:38-45). */
typedef enum
{
    RVSP_CSR_OK = 0,
    RVSP_CSR_BAD_ROW_PTR,
    RVSP_CSR_COL_OUT_OF_RANGE,
    RVSP_CSR_UNSORTED,
    RVSP_CSR_DUPLICATE
} rvsp_csr_status_t;

/* Canonical check on host arrays, O(nnz), first failing row in
 * *bad_row_out (-1 when not row specific). Reference:
 * src/kernels/spgemm/rvsp_support.c:36-126. */
rvsp_csr_status_t rvsp_csr_check(int32_t rows, int32_t cols,
                                 const int32_t *row_ptr,
                                 const int32_t *col_idx,
                                 int32_t *bad_row_out);

/* Reference: src/kernels/spgemm/rvsp_support.c:128-150. */
const char *rvsp_csr_status_string(rvsp_csr_status_t status);

/*
 * Raw one-shot kernel with the parameter list of the reference's
 * RVSP_SPGEMM_PARAMS (src/kernels/spgemm/rvsp_spgemm.h:68-77; scalar instance
 * gustavson_core_f32.inc:66-194): symbolic count, malloc of C, fill and
 * numeric, all on the GPU. Outputs are malloc'd host arrays owned by the
 * caller; op_counts_out is optional (a_rows entries). workspace must be
 * non-NULL and at least rvsp_spgemm_buffer_size(b_cols) bytes, as in the
 * reference, although the GPU kernels do not touch it.
 */
rvsp_status_t rvsp_spgemm_hip_f32(int32_t a_rows, int32_t a_cols,
                                  int32_t b_cols,
                                  const int32_t *a_row_ptr,
                                  const int32_t *a_col_idx,
                                  const float *a_values,
                                  const int32_t *b_row_ptr,
                                  const int32_t *b_col_idx,
                                  const float *b_values,
                                  void *workspace, size_t workspace_bytes,
                                  int32_t **c_row_ptr_out,
                                  int32_t **c_col_idx_out,
                                  float **c_values_out, int32_t *c_nnz_out,
                                  int64_t *op_counts_out);

/*
 * The reference's scalar raw kernel (src/kernels/spgemm/rvsp_spgemm.h:79),
 * which its bench harness calls (bench/bench.c:166): here the same GPU path
 * as rvsp_spgemm_hip_f32, as the scalar backend of rvsp_spgemm_csr is.
 */
rvsp_status_t rvsp_spgemm_scalar_f32(int32_t a_rows, int32_t a_cols,
                                     int32_t b_cols,
                                     const int32_t *a_row_ptr,
                                     const int32_t *a_col_idx,
                                     const float *a_values,
                                     const int32_t *b_row_ptr,
                                     const int32_t *b_col_idx,
                                     const float *b_values,
                                     void *workspace, size_t workspace_bytes,
                                     int32_t **c_row_ptr_out,
                                     int32_t **c_col_idx_out,
                                     float **c_values_out, int32_t *c_nnz_out,
                                     int64_t *op_counts_out);

#ifdef __cplusplus
}
#endif

#endif /* RVSP_SPGEMM_H */
