/*
 * rv_sparse_mtx.h — Matrix Market input and CSR canonicalisation.
 *
 * The reference converts Matrix Market files with its bench tool's
 * assemble_csr_matrix (tools/src/mtx_to_csr_formatter.c:74-325): banner
 * (pattern / symmetric), 1-based coordinates, symmetric storage expanded by
 * mirroring every off-diagonal entry, then every row sorted by column with
 * repeated (row, column) entries summed. These calls give the same matrix
 * through the library's own CSR type, so a caller of the reference's bench
 * path can load its SuiteSparse inputs here and hand them straight to
 * rvsp_spgemm_csr or the descriptor API.
 */
#ifndef RV_SPARSE_MTX_H
#define RV_SPARSE_MTX_H

#include <stdint.h>

#include "rv_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Reads the Matrix Market coordinate file at `path` into *out: an FP32 CSR
 * matrix in host memory with sorted, distinct columns per row, allocated
 * with malloc and owned by *out (owns_data = 1; release with
 * rvsp_csr_destroy).
 *
 * Same matrix as the reference's assemble_csr_matrix for a well-formed
 * file: "pattern" entries get the value 1.0f; "symmetric" storage is
 * expanded; values are parsed as FP32 from their decimal text (strtof);
 * entries land in file order (stored entries first, then the mirrored
 * ones), and repeated entries of a row are summed in that order in FP32,
 * first value plus the next, and so on.
 *
 * Deliberate differences: where the reference reads past its buffers or
 * silently stops, this call reports RVSP_ERROR_INVALID_ARGUMENT: a missing
 * or malformed banner or size line, an array (non-coordinate) or complex
 * file, fewer entries than the size line states, a coordinate outside
 * [1, rows] x [1, cols], or sizes beyond int32. RVSP_ERROR_NULL_POINTER for
 * NULL arguments, RVSP_ERROR_ALLOCATION_FAILED when memory runs out. On
 * failure *out is left empty. The text is parsed on up to `threads` host
 * threads (0: the machine's, at most 16).
 */
rvsp_status_t rvsp_csr_read_mtx(const char *path, int32_t threads,
                                rvsp_csr_matrix_t *out);

/*
 * Canonicalises a host FP32 CSR matrix in place, the reference's
 * canonicalise_rows (tools/src/mtx_to_csr_formatter.c:74-151): every row
 * sorted by column (stable, so repeated entries keep their order) and
 * repeated columns folded into one entry, their values summed in order.
 * row_ptr is rewritten and A->nnz shrinks by the folded entries; the
 * arrays keep their allocation. The row pointers must be non-decreasing
 * from 0 and columns within [0, cols) (RVSP_ERROR_INVALID_ARGUMENT
 * otherwise, matrix unchanged).
 */
rvsp_status_t rvsp_csr_canonicalize(rvsp_csr_matrix_t *A);

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_MTX_H */
