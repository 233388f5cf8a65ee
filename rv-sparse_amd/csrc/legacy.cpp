// legacy.cpp — rvsp_spgemm_csr for the reference's two other dtype pairs,
// FP64 x FP64 -> FP64 and INT8 x INT8 -> INT32 (reference dispatch
// src/core/spgemm.c:182-196 to the scalar legacy kernels
// src/kernels/spgemm/legacy/csr_scalar_f64.c and csr_scalar_i8.c through
// csr_spgemm_wrappers.c:19-90 and :352-390), on the GPU.
//
// Semantics kept from those kernels: C(i,j) sums a * b over A's entries in
// order and B's in order (FP64: product and sum rounded separately; INT8:
// int32 products and sums), rows come out in ascending column order, and
// entries whose sum is exactly zero are not stored (unlike the FP32 path,
// which keeps them); C's arrays are malloc'ed with room for the symbolic
// count, C.nnz is the count after dropping. The argument checks and status
// codes are each wrapper's own (see below). One difference: the legacy
// kernels also accept rows with unsorted or repeated columns; this path
// needs canonical columns (the structure comes from the FP32 path's
// symbolic phase) and reports RVSP_ERROR_INVALID_CSR for such rows.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include "common.hpp"
#include "descr_impl.hpp"
#include "kernels.hpp"

namespace rvsp {

namespace {

struct Staged {
  DevBuf rp, ci, val;
};

rvsp_status_t stage(const rvsp_csr_matrix_t* M, size_t esize, Staged& out, hipStream_t s) {
  const size_t n = (size_t)M->nnz;
  RVSP_TRY(ensure(out.rp, ((size_t)M->rows + 1) * 4));
  RVSP_TRY(ensure(out.ci, n ? n * 4 : 4));
  RVSP_TRY(ensure(out.val, n ? n * esize : esize));
  RVSP_HIP_TRY(hipMemcpyAsync(out.rp.p, M->row_ptr, ((size_t)M->rows + 1) * 4,
                              hipMemcpyHostToDevice, s));
  if (n) {
    RVSP_HIP_TRY(hipMemcpyAsync(out.ci.p, M->col_idx, n * 4, hipMemcpyHostToDevice, s));
    RVSP_HIP_TRY(hipMemcpyAsync(out.val.p, M->values, n * esize, hipMemcpyHostToDevice, s));
  }
  return RVSP_SUCCESS;
}

void release_staged(Staged& m) {
  release(m.rp);
  release(m.ci);
  release(m.val);
}

// Host view of a staged operand for the FP32 path's symbolic phase, which
// reads the structure only.
rvsp_csr_matrix_t structure_view(const rvsp_csr_matrix_t* M, const Staged& st) {
  rvsp_csr_matrix_t v;
  memset(&v, 0, sizeof v);
  v.rows = M->rows;
  v.cols = M->cols;
  v.nnz = M->nnz;
  v.row_ptr = (int32_t*)st.rp.p;
  v.col_idx = (int32_t*)st.ci.p;
  v.values = st.val.p;
  v.dtype = RVSP_DTYPE_FP32;
  v.format = RVSP_FORMAT_CSR;
  return v;
}

// C with `rows` empty rows (FP64 operands with no rows or no columns).
rvsp_status_t empty_result(int rows, int cols, size_t esize, rvsp_dtype_t dtype,
                           rvsp_csr_matrix_t* C) {
  int32_t* rp = (int32_t*)calloc((size_t)rows + 1, sizeof(int32_t));
  int32_t* ci = (int32_t*)malloc(sizeof(int32_t));
  void* val = malloc(esize);
  if (!rp || !ci || !val) {
    free(rp); free(ci); free(val);
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
  C->rows = rows;
  C->cols = cols;
  C->nnz = 0;
  C->row_ptr = rp;
  C->col_idx = ci;
  C->values = val;
  C->dtype = dtype;
  C->format = RVSP_FORMAT_CSR;
  C->owns_data = 1;
  return RVSP_SUCCESS;
}

// The product of two staged, checked operands.
rvsp_status_t multiply(const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B, const Staged& a,
                       const Staged& b, bool i8, rvsp_spgemm_descr* d, hipStream_t s,
                       rvsp_csr_matrix_t* C) {
  const size_t out_size = i8 ? 4 : 8;
  const rvsp_csr_matrix_t va = structure_view(A, a), vb = structure_view(B, b);
  size_t ws = 0;
  int32_t sym_nnz = 0;
  RVSP_TRY(rvsp_spgemm_work_estimation(d, &va, &vb, &ws, &sym_nnz));
  const int rows = A->rows;
  const size_t cap = sym_nnz > 0 ? (size_t)sym_nnz : 1;
  DevBuf c_ci, acc, nonzero, out_rp, tiles, total, out_ci, out_val;
  rvsp_status_t st = RVSP_SUCCESS;
  long long kept = 0;
  int32_t* h_rp = (int32_t*)malloc(((size_t)rows + 1) * sizeof(int32_t));
  int32_t* h_ci = (int32_t*)malloc(cap * sizeof(int32_t));
  void* h_val = malloc(cap * out_size);
  if (!h_rp || !h_ci || !h_val) st = RVSP_ERROR_ALLOCATION_FAILED;
  for (DevBuf* x : {&c_ci, &out_ci}) if (st == RVSP_SUCCESS) st = ensure(*x, cap * 4);
  for (DevBuf* x : {&acc, &out_val}) if (st == RVSP_SUCCESS) st = ensure(*x, cap * out_size);
  if (st == RVSP_SUCCESS) st = ensure(nonzero, (size_t)rows * 4);
  if (st == RVSP_SUCCESS) st = ensure(out_rp, ((size_t)rows + 1) * 4);
  if (st == RVSP_SUCCESS) st = ensure(tiles, scan_tile_count(rows) * 8);
  if (st == RVSP_SUCCESS) st = ensure(total, 8);
  DevCsr da, db;
  da.rows = A->rows; da.cols = A->cols; da.nnz = A->nnz;
  da.rp = (const int*)a.rp.p; da.ci = (const int*)a.ci.p;
  db.rows = B->rows; db.cols = B->cols; db.nnz = B->nnz;
  db.rp = (const int*)b.rp.p; db.ci = (const int*)b.ci.p;
  if (st == RVSP_SUCCESS) st = run_fill(d, da, db, (int*)c_ci.p, s);
  hipError_t e = hipSuccess;
  if (st == RVSP_SUCCESS) {
    const int* c_rp = (const int*)d->c_rp.p;
    e = launch_numeric_legacy(i8, rows, da.rp, da.ci, a.val.p, db.rp, db.ci, b.val.p, c_rp,
                              (const int*)c_ci.p, acc.p, (int*)nonzero.p, s);
    if (e == hipSuccess)
      e = launch_scan((const int*)nonzero.p, rows, (int*)out_rp.p, (long long*)total.p,
                      (long long*)tiles.p, s);
    if (e == hipSuccess)
      e = launch_compact_legacy(i8, rows, c_rp, (const int*)c_ci.p, acc.p,
                                (const int*)out_rp.p, (int*)out_ci.p, out_val.p, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&kept, total.p, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_rp, out_rp.p, ((size_t)rows + 1) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && kept > 0)
      e = hipMemcpyAsync(h_ci, out_ci.p, (size_t)kept * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && kept > 0)
      e = hipMemcpyAsync(h_val, out_val.p, (size_t)kept * out_size, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = from_hip(e);
  }
  (void)hipStreamSynchronize(s);
  for (DevBuf* x : {&c_ci, &acc, &nonzero, &out_rp, &tiles, &total, &out_ci, &out_val})
    release(*x);
  if (st != RVSP_SUCCESS) {
    free(h_rp); free(h_ci); free(h_val);
    return st;
  }
  C->rows = rows;
  C->cols = B->cols;
  C->nnz = (int32_t)kept;
  C->row_ptr = h_rp;
  C->col_idx = h_ci;
  C->values = h_val;
  C->dtype = i8 ? RVSP_DTYPE_INT32 : RVSP_DTYPE_FP64;
  C->format = RVSP_FORMAT_CSR;
  C->owns_data = 1;
  return RVSP_SUCCESS;
}

// Basic checks of rvsp_csr_validate (src/core/matrix.c:51-90) that need no
// scan of the columns; the scan runs on the device afterwards.
rvsp_status_t validate_head(const rvsp_csr_matrix_t* M) {
  if (!M->row_ptr || !M->col_idx || !M->values) return RVSP_ERROR_NULL_POINTER;
  if (M->rows <= 0 || M->cols <= 0 || M->nnz < 0) return RVSP_ERROR_INVALID_ARGUMENT;
  if (M->row_ptr[0] != 0 || M->row_ptr[M->rows] != M->nnz) return RVSP_ERROR_INVALID_CSR;
  return RVSP_SUCCESS;
}

}  // namespace

rvsp_status_t oneshot_legacy(const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B,
                             rvsp_csr_matrix_t* C, bool i8) {
  const size_t esize = i8 ? 1 : 8;
  if (i8) {
    // csr_spgemm_wrappers.c:26-51: validate A, validate B, dtypes, shapes.
    RVSP_TRY(validate_head(A));
  } else {
    // csr_spgemm_wrappers.c:357-370, csr_scalar_f64.c:55-62: dtypes, shapes,
    // then sizes and arrays, all INVALID_ARGUMENT.
    if (A->dtype != RVSP_DTYPE_FP64 || B->dtype != RVSP_DTYPE_FP64)
      return RVSP_ERROR_INVALID_ARGUMENT;
    if (A->cols != B->rows) return RVSP_ERROR_INVALID_ARGUMENT;
    if (A->rows < 0 || A->cols < 0 || B->cols < 0 || A->nnz < 0 || B->nnz < 0 ||
        !A->row_ptr || !A->col_idx || !A->values || !B->row_ptr || !B->col_idx ||
        !B->values)
      return RVSP_ERROR_INVALID_ARGUMENT;
  }
  RVSP_TRY(device_ready());
  rvsp_spgemm_descr_t d = NULL;
  RVSP_TRY(rvsp_spgemm_descr_create(&d));
  rvsp_status_t st = rvsp_spgemm_set_algo(d, RVSP_SPGEMM_ALGO_HIP_FIXED);
  if (st == RVSP_SUCCESS) st = rvsp_spgemm_set_memory(d, RVSP_MEMORY_DEVICE);
  hipStream_t s = nullptr;
  if (st == RVSP_SUCCESS) st = descr_stream(d, &s);
  Staged a, b;
  const bool same = A->row_ptr == B->row_ptr && A->col_idx == B->col_idx &&
                    A->values == B->values && A->rows == B->rows && A->nnz == B->nnz;
  // Structure verdicts map to each wrapper's codes: INT8 validates with
  // rvsp_csr_validate (INVALID_CSR for every structural fault), the FP64
  // kernel reports malformed row pointers and columns out of range as
  // INVALID_ARGUMENT; unsorted or repeated columns are this path's
  // INVALID_CSR.
  auto check = [&](const rvsp_csr_matrix_t* M, const Staged& m) -> rvsp_status_t {
    rvsp_csr_status_t v = RVSP_CSR_OK;
    int bad = -1;
    RVSP_TRY(csr_check_device(M->rows, M->cols, M->nnz, (const int*)m.rp.p,
                              (const int*)m.ci.p, d->check_key, s, &v, &bad));
    if (v == RVSP_CSR_OK) return RVSP_SUCCESS;
    if (!i8 && (v == RVSP_CSR_BAD_ROW_PTR || v == RVSP_CSR_COL_OUT_OF_RANGE))
      return RVSP_ERROR_INVALID_ARGUMENT;
    return RVSP_ERROR_INVALID_CSR;
  };
  if (st == RVSP_SUCCESS) st = stage(A, esize, a, s);
  if (st == RVSP_SUCCESS) st = check(A, a);
  if (st == RVSP_SUCCESS && i8 && !same) st = validate_head(B);
  if (st == RVSP_SUCCESS && !same) st = stage(B, esize, b, s);
  if (st == RVSP_SUCCESS && !same) st = check(B, b);
  if (st == RVSP_SUCCESS && i8) {
    if (A->dtype != RVSP_DTYPE_INT8 || B->dtype != RVSP_DTYPE_INT8)
      st = RVSP_ERROR_UNSUPPORTED_DTYPE;
    else if (A->cols != B->rows)
      st = RVSP_ERROR_INVALID_ARGUMENT;
  }
  if (st == RVSP_SUCCESS) {
    if (A->rows == 0 || A->cols == 0 || B->cols == 0)
      st = empty_result(A->rows, B->cols, 8, RVSP_DTYPE_FP64, C);  // FP64 only
    else
      st = multiply(A, B, a, same ? a : b, i8, d, s, C);
  }
  if (s) (void)hipStreamSynchronize(s);
  release_staged(a);
  release_staged(b);
  rvsp_spgemm_descr_destroy(d);
  return st;
}

}  // namespace rvsp
