// matrix.cpp — host-side parts of the API that never touch the GPU: version
// and status strings, the non-owning CSR wrapper, basic and canonical CSR
// validation, and the reference's workspace-size query.
#include <stdlib.h>

#include "rv_sparse.h"
#include "rvsp_spgemm.h"

extern "C" {

// Reference: src/core/error.c:17-20.
const char* rvsp_get_version(void) { return "0.1.0"; }

// Reference: src/core/error.c:22-43, plus the appended device status.
const char* rvsp_status_to_string(rvsp_status_t status) {
  switch (status) {
    case RVSP_SUCCESS: return "success";
    case RVSP_ERROR_NULL_POINTER: return "null pointer";
    case RVSP_ERROR_INVALID_ARGUMENT: return "invalid argument";
    case RVSP_ERROR_INVALID_CSR: return "invalid CSR matrix";
    case RVSP_ERROR_UNSUPPORTED_DTYPE: return "unsupported data type";
    case RVSP_ERROR_UNSUPPORTED_BACKEND: return "unsupported backend";
    case RVSP_ERROR_ALLOCATION_FAILED: return "allocation failed";
    case RVSP_ERROR_DEVICE: return "device error";
  }
  return "unknown error";
}

// Reference: src/core/matrix.c:51-90. Structural checks only: no sortedness
// or duplicate check (that is rvsp_csr_check).
rvsp_status_t rvsp_csr_validate(const rvsp_csr_matrix_t* A) {
  if (A == NULL || A->row_ptr == NULL || A->col_idx == NULL || A->values == NULL)
    return RVSP_ERROR_NULL_POINTER;
  if (A->rows <= 0 || A->cols <= 0 || A->nnz < 0) return RVSP_ERROR_INVALID_ARGUMENT;
  const int32_t* rp = A->row_ptr;
  if (rp[0] != 0 || rp[A->rows] != A->nnz) return RVSP_ERROR_INVALID_CSR;
  for (int32_t r = 0; r < A->rows; ++r)
    if (rp[r + 1] < rp[r]) return RVSP_ERROR_INVALID_CSR;
  for (int32_t p = 0; p < A->nnz; ++p) {
    const int32_t c = A->col_idx[p];
    if (c < 0 || c >= A->cols) return RVSP_ERROR_INVALID_CSR;
  }
  return RVSP_SUCCESS;
}

// Reference: src/core/matrix.c:18-49. Wraps without copying; owns_data = 0.
rvsp_status_t rvsp_csr_create(rvsp_csr_matrix_t* A, int32_t rows, int32_t cols,
                              int32_t nnz, int32_t* row_ptr, int32_t* col_idx,
                              void* values, rvsp_dtype_t dtype) {
  if (A == NULL || row_ptr == NULL || col_idx == NULL || values == NULL)
    return RVSP_ERROR_NULL_POINTER;
  if (rows <= 0 || cols <= 0 || nnz < 0) return RVSP_ERROR_INVALID_ARGUMENT;
  A->rows = rows;
  A->cols = cols;
  A->nnz = nnz;
  A->row_ptr = row_ptr;
  A->col_idx = col_idx;
  A->values = values;
  A->dtype = dtype;
  A->format = RVSP_FORMAT_CSR;
  A->owns_data = 0;
  return rvsp_csr_validate(A);
}

// Reference: src/core/matrix.c:92-115. Frees only library-allocated arrays
// (rvsp_spgemm_csr hands back malloc'd host memory), then resets the struct.
void rvsp_csr_destroy(rvsp_csr_matrix_t* A) {
  if (A == NULL) return;
  if (A->owns_data) {
    free(A->row_ptr);
    free(A->col_idx);
    free(A->values);
  }
  A->rows = 0;
  A->cols = 0;
  A->nnz = 0;
  A->row_ptr = NULL;
  A->col_idx = NULL;
  A->values = NULL;
  A->dtype = RVSP_DTYPE_FP32;
  A->format = RVSP_FORMAT_CSR;
  A->owns_data = 0;
}

// Reference: src/kernels/spgemm/rvsp_support.c:36-126. Within a row a
// repeated column is reported as DUPLICATE and a descent as UNSORTED, first
// failure wins, scanning rows in order.
rvsp_csr_status_t rvsp_csr_check(int32_t rows, int32_t cols,
                                 const int32_t* row_ptr, const int32_t* col_idx,
                                 int32_t* bad_row_out) {
  if (bad_row_out) *bad_row_out = -1;
  if (rows < 0 || cols < 0 || row_ptr == NULL || row_ptr[0] != 0)
    return RVSP_CSR_BAD_ROW_PTR;
  for (int32_t r = 0; r < rows; ++r) {
    const int32_t s = row_ptr[r], e = row_ptr[r + 1];
    if (e < s) {
      if (bad_row_out) *bad_row_out = r;
      return RVSP_CSR_BAD_ROW_PTR;
    }
    if (e == s) continue;
    if (col_idx == NULL) return RVSP_CSR_BAD_ROW_PTR;
    int32_t prev = -1;
    for (int32_t p = s; p < e; ++p) {
      const int32_t c = col_idx[p];
      rvsp_csr_status_t bad = RVSP_CSR_OK;
      if (c < 0 || c >= cols)
        bad = RVSP_CSR_COL_OUT_OF_RANGE;
      else if (c == prev)
        bad = RVSP_CSR_DUPLICATE;
      else if (c < prev)
        bad = RVSP_CSR_UNSORTED;
      if (bad != RVSP_CSR_OK) {
        if (bad_row_out) *bad_row_out = r;
        return bad;
      }
      prev = c;
    }
  }
  return RVSP_CSR_OK;
}

// Reference: src/kernels/spgemm/rvsp_support.c:128-150.
const char* rvsp_csr_status_string(rvsp_csr_status_t status) {
  switch (status) {
    case RVSP_CSR_OK: return "canonical CSR";
    case RVSP_CSR_BAD_ROW_PTR: return "malformed row pointer";
    case RVSP_CSR_COL_OUT_OF_RANGE: return "column index out of range";
    case RVSP_CSR_UNSORTED: return "column indices not ascending";
    case RVSP_CSR_DUPLICATE: return "duplicate column in row";
  }
  return "unknown";
}

// Reference: src/kernels/spgemm/rvsp_support.c:11-29 with the layout of
// rvsp_common.h:48-60: four 64-byte aligned regions (float acc, byte mark,
// int32 touched, int32 scratch) of b_cols entries each.
rvsp_status_t rvsp_spgemm_buffer_size(int32_t b_cols, size_t* bytes_out) {
  if (bytes_out == NULL) return RVSP_ERROR_NULL_POINTER;
  if (b_cols <= 0) return RVSP_ERROR_INVALID_ARGUMENT;
  const size_t n = (size_t)b_cols;
  auto up64 = [](size_t x) { return (x + 63) & ~(size_t)63; };
  *bytes_out = up64(n * 4) + up64(n) + up64(n * 4) + up64(n * 4);
  return RVSP_SUCCESS;
}

}  // extern "C"
