/*
 * rv_sparse_types.h — types of the rv-sparse C API, MI355X (gfx950) HIP build.
 *
 * The enumerators, their numeric values and the struct layouts match
 * include/rv_sparse_types.h:21-96 of the rv-sparse reference, so a binary
 * compiled against that header links against this library unchanged.
 * Values added for the HIP backend are appended after the reference's last
 * value of each enum; no existing value moves.
 */
#ifndef RV_SPARSE_TYPES_H
#define RV_SPARSE_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference: include/rv_sparse_types.h:21-30. */
typedef enum
{
    RVSP_SUCCESS = 0,
    RVSP_ERROR_NULL_POINTER = -1,
    RVSP_ERROR_INVALID_ARGUMENT = -2,
    RVSP_ERROR_INVALID_CSR = -3,
    RVSP_ERROR_UNSUPPORTED_DTYPE = -4,
    RVSP_ERROR_UNSUPPORTED_BACKEND = -5,
    RVSP_ERROR_ALLOCATION_FAILED = -6,
    /* Appended: a HIP runtime call failed for a reason other than running out
     * of device memory, or no usable gfx950 device is present. */
    RVSP_ERROR_DEVICE = -7
} rvsp_status_t;

/* Reference: include/rv_sparse_types.h:32-39. */
typedef enum
{
    RVSP_DTYPE_INT8,
    RVSP_DTYPE_INT32,
    RVSP_DTYPE_BF16,
    RVSP_DTYPE_FP32,
    RVSP_DTYPE_FP64
} rvsp_dtype_t;

/* Reference: include/rv_sparse_types.h:41-44. */
typedef enum
{
    RVSP_FORMAT_CSR
} rvsp_format_t;

/* Reference: include/rv_sparse_types.h:46-51. */
typedef enum
{
    RVSP_BACKEND_SCALAR,
    RVSP_BACKEND_SCALAR_UNROLL4,
    RVSP_BACKEND_RVV_INTRINSICS,
    /* Appended: the gfx950 HIP implementation. In this library the scalar
     * backends resolve to it as well; its output is bit-identical to the
     * reference's scalar kernel. */
    RVSP_BACKEND_HIP
} rvsp_backend_t;

/* Reference: include/rv_sparse_types.h:53-67. */
typedef struct
{
    int32_t rows;
    int32_t cols;
    int32_t nnz;

    int32_t *row_ptr;
    int32_t *col_idx;
    void *values;

    rvsp_dtype_t dtype;
    rvsp_format_t format;

    int owns_data;
} rvsp_csr_matrix_t;

/* Reference: include/rv_sparse_types.h:69-74. */
typedef struct
{
    rvsp_backend_t backend;
    rvsp_dtype_t input_dtype;
    rvsp_dtype_t output_dtype;
} rvsp_spgemm_options_t;

/* Opaque two-phase SpGEMM descriptor (reference: rv_sparse_types.h:82). */
typedef struct rvsp_spgemm_descr *rvsp_spgemm_descr_t;

/*
 * Accumulation strategy (reference: include/rv_sparse_types.h:90-96).
 *
 * DEFAULT is the scalar strategy of the reference; here it runs on the GPU
 * with the fixed per-bin plan and yields the reference scalar kernel's exact
 * output bits. RVV, CONTIG and ADAPTIVE are RISC-V vector strategies and are
 * reported as RVSP_ERROR_UNSUPPORTED_BACKEND, as on a non-vector reference
 * build.
 *
 * The appended HIP strategies differ only in how rows are assigned to GPU
 * accumulator kernels; every one of them produces the same output bits:
 *   HIP_FIXED    fixed bin thresholds and workgroup shapes (plain HIP)
 *   HIP_TUNED    bin shapes from the autotuner's table
 *   HIP_LEARNED  per-bin accumulator kind and table size chosen by the
 *                MLP selector; the exact symbolic count vetoes any choice
 *                that does not fit and routes it to global dense
 */
typedef enum
{
    RVSP_SPGEMM_ALGO_DEFAULT = 0,
    RVSP_SPGEMM_ALGO_RVV,
    RVSP_SPGEMM_ALGO_CONTIG,
    RVSP_SPGEMM_ALGO_ADAPTIVE,
    RVSP_SPGEMM_ALGO_HIP_FIXED,
    RVSP_SPGEMM_ALGO_HIP_TUNED,
    RVSP_SPGEMM_ALGO_HIP_LEARNED
} rvsp_spgemm_algo_t;

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_TYPES_H */
