/*
 * rv_sparse.h — public C API of rv-sparse, served by the MI355X HIP library
 * librvsparse_hip.so.
 *
 * Every prototype below has the name, signature and error contract of the
 * reference's include/rv_sparse.h:27-142, so existing callers (the
 * reference's examples/, tests/ and bench/) link against this library
 * without source changes. The work behind each call runs on the GPU; there
 * is no CPU compute path. Where the reference's behaviour is observable
 * (check order, status codes, ownership), this library reproduces it; each
 * function's definition cites the reference line it follows.
 *
 * Operands passed through this header are HOST arrays, exactly as in the
 * reference. Device-resident operands, explicit streams and the batched entry
 * point are declared in rv_sparse_hip.h.
 */
#ifndef RV_SPARSE_H
#define RV_SPARSE_H

#include <stddef.h>

#include "rv_sparse_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RVSP_VERSION_MAJOR 0
#define RVSP_VERSION_MINOR 1
#define RVSP_VERSION_PATCH 0

/* Reference: src/core/error.c:17-20 (returns "0.1.0"). */
const char *rvsp_get_version(void);

/* Reference: src/core/error.c:22-43. RVSP_ERROR_DEVICE maps to
 * "device error". */
const char *rvsp_status_to_string(rvsp_status_t status);

/*
 * CSR lifecycle (reference: src/core/matrix.c:18-115).
 * rvsp_csr_create wraps caller arrays without copying or owning them and
 * runs rvsp_csr_validate. rvsp_csr_destroy frees with free() only when
 * owns_data is set, then resets the struct; NULL is accepted.
 */
rvsp_status_t rvsp_csr_create(rvsp_csr_matrix_t *A,
                              int32_t rows,
                              int32_t cols,
                              int32_t nnz,
                              int32_t *row_ptr,
                              int32_t *col_idx,
                              void *values,
                              rvsp_dtype_t dtype);

rvsp_status_t rvsp_csr_validate(const rvsp_csr_matrix_t *A);

void rvsp_csr_destroy(rvsp_csr_matrix_t *A);

/*
 * One-shot C = A * B (reference: src/core/spgemm.c:135-211).
 * FP32 x FP32 -> FP32 with backend SCALAR, SCALAR_UNROLL4 or HIP runs the HIP
 * descriptor path with RVSP_SPGEMM_ALGO_DEFAULT. C's arrays are malloc'd host
 * memory with owns_data = 1; release them with rvsp_csr_destroy.
 */
rvsp_status_t rvsp_spgemm_csr(const rvsp_csr_matrix_t *A,
                              const rvsp_csr_matrix_t *B,
                              rvsp_csr_matrix_t *C,
                              const rvsp_spgemm_options_t *options);

/*
 * Two-phase descriptor API (reference: src/core/spgemm_descr.c:66-380).
 * A and B must be canonical CSR. work_estimation runs the symbolic phase on
 * the GPU and reports the exact c_nnz; compute fills C's columns when the
 * descriptor has not yet filled this C.col_idx buffer, then runs the numeric
 * phase. Repeated computes with the same C.col_idx run the numeric phase
 * only.
 */
rvsp_status_t rvsp_spgemm_descr_create(rvsp_spgemm_descr_t *descr);

void rvsp_spgemm_descr_destroy(rvsp_spgemm_descr_t descr);

rvsp_status_t rvsp_spgemm_set_algo(rvsp_spgemm_descr_t descr,
                                   rvsp_spgemm_algo_t algo);

rvsp_status_t rvsp_spgemm_work_estimation(rvsp_spgemm_descr_t descr,
                                          const rvsp_csr_matrix_t *A,
                                          const rvsp_csr_matrix_t *B,
                                          size_t *workspace_bytes,
                                          int32_t *c_nnz_out);

rvsp_status_t rvsp_spgemm_compute(rvsp_spgemm_descr_t descr,
                                  const rvsp_csr_matrix_t *A,
                                  const rvsp_csr_matrix_t *B,
                                  rvsp_csr_matrix_t *C,
                                  void *workspace);

rvsp_status_t rvsp_spgemm_get_op_counts(rvsp_spgemm_descr_t descr,
                                        int64_t *op_counts_out,
                                        int32_t n);

void rvsp_spgemm_invalidate_structure(rvsp_spgemm_descr_t descr);

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_H */
