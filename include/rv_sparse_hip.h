/*
 * rv_sparse_hip.h — MI355X (gfx950) extensions to the rv-sparse C API.
 *
 * None of these has a counterpart in the reference, whose operands are always
 * host arrays and whose work always runs on the calling thread. They exist
 * because a GPU caller needs three things the reference API cannot express:
 * operands already resident in device memory, a stream to order the work on,
 * and a way to see which accumulator kernels the analysis chose. hip types
 * are passed as void * so this header needs no HIP include.
 */
#ifndef RV_SPARSE_HIP_H
#define RV_SPARSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "rv_sparse.h"
#include "rvsp_spgemm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Where the descriptor's operand arrays live. */
typedef enum
{
    /* A, B and C arrays are host memory (reference semantics). Inputs are
     * copied to the device by work_estimation/compute; C is copied back. */
    RVSP_MEMORY_HOST = 0,
    /* A, B and C arrays are device pointers of the current HIP device. No
     * host copies are made; compute writes C in place. */
    RVSP_MEMORY_DEVICE = 1
} rvsp_memory_t;

/* Selects operand memory for this descriptor. Must precede work_estimation
 * (RVSP_ERROR_INVALID_ARGUMENT afterwards, as for set_algo). */
rvsp_status_t rvsp_spgemm_set_memory(rvsp_spgemm_descr_t descr,
                                     rvsp_memory_t memory);

/* Orders all of the descriptor's device work on `hip_stream` (a
 * hipStream_t); NULL returns to the descriptor's own stream. */
rvsp_status_t rvsp_spgemm_set_stream(rvsp_spgemm_descr_t descr,
                                     void *hip_stream);

/* The hipStream_t the descriptor enqueues on (creating its own stream on
 * first use). */
rvsp_status_t rvsp_spgemm_get_stream(rvsp_spgemm_descr_t descr,
                                     void **hip_stream_out);

/* async != 0: in RVSP_MEMORY_DEVICE mode, compute returns once its kernels
 * are enqueued instead of waiting for them (the caller synchronises the
 * stream). work_estimation always waits, because it reports c_nnz. Host
 * mode always waits. */
rvsp_status_t rvsp_spgemm_set_async(rvsp_spgemm_descr_t descr, int async);

/* on != 0 (default): from the first numeric-only compute on (C.col_idx is
 * the array filled before), the descriptor's searching numeric bins switch to
 * the position map (RVSP_KIND_MAP), built once on that compute; that compute
 * synchronises. Tuned plans keep their measured shapes. Output bits are the
 * same either way. */
rvsp_status_t rvsp_spgemm_set_position_map(rvsp_spgemm_descr_t descr, int on);

/* Phase ids used by the plan queries below. */
typedef enum
{
    RVSP_PHASE_COUNT = 0,   /* symbolic nnz count */
    RVSP_PHASE_FILL = 1,    /* sorted column indices */
    RVSP_PHASE_NUMERIC = 2  /* values */
} rvsp_phase_t;

/* Accumulator kinds a bin can be served by. */
typedef enum
{
    RVSP_KIND_NONE = 0,    /* empty rows, nothing to do */
    RVSP_KIND_DIRECT = 1,  /* one contributing B row: scaled copy */
    RVSP_KIND_HASH = 2,    /* LDS hash table, sub-wave group per row */
    RVSP_KIND_SEARCH = 3,  /* LDS sorted columns + binary search */
    RVSP_KIND_BITMAP = 4,  /* LDS bitmap over the row's column span */
    RVSP_KIND_BLOCK = 5,   /* long rows: workgroups over runs of output
                              positions, tiles of up to `cap` positions with
                              LDS accumulators and a column index */
    RVSP_KIND_GLOBAL = 6,  /* workgroup per row, dense global scratch */
    RVSP_KIND_MAP = 7      /* numeric via a cached position per product: the
                              descriptor records, on the first compute, where
                              in C's row every product lands (1 or 2 bytes
                              per product) and later numeric-only computes
                              read that instead of searching */
} rvsp_kind_t;

/* One bin of the analysed plan. */
typedef struct
{
    int32_t phase;     /* rvsp_phase_t */
    int32_t bin;
    int32_t kind;      /* rvsp_kind_t */
    int32_t group;     /* lanes per row */
    int32_t cap;       /* table slots / columns / bitmap bits */
    int32_t rows;      /* rows assigned to the bin */
    int32_t lds_bytes; /* dynamic LDS per workgroup */
    int32_t bound;     /* largest size metric the bin admits */
} rvsp_bin_info_t;

/* Copies up to `max` bins of the plan built by the last work_estimation
 * (count, fill and numeric phases, empty bins included) and the number
 * available in *n_out. */
rvsp_status_t rvsp_spgemm_get_plan(rvsp_spgemm_descr_t descr,
                                   rvsp_bin_info_t *out, int32_t max,
                                   int32_t *n_out);

/* A per-bin override applied by the next work_estimation. kind/group/cap of
 * 0 keep the table's value. Whatever is asked, the exact symbolic count has
 * the last word: a choice whose table cannot hold the bin's largest row, or
 * whose shape cannot launch, turns the bin into RVSP_KIND_GLOBAL. Output bits
 * never depend on these choices. */
typedef struct
{
    int32_t phase;
    int32_t bin;
    int32_t kind;
    int32_t group;
    int32_t cap;
    int32_t variant; /* kernel variant of `kind` (RVSP_KIND_MAP: 1 stages A's
                        metadata in LDS); taken together with a non-zero kind */
} rvsp_bin_choice_t;

/* Replaces the descriptor's override list (n == 0 clears it). Must precede
 * work_estimation. */
rvsp_status_t rvsp_spgemm_set_bin_choices(rvsp_spgemm_descr_t descr,
                                          const rvsp_bin_choice_t *choices,
                                          int32_t n);

/*
 * Autotuner. Does everything rvsp_spgemm_compute does (same arguments and
 * checks) and, in addition, times every launchable kernel shape of every
 * non-empty bin of the descriptor's plan on these operands (`reps` timed
 * launches each), keeps the fastest in the descriptor, and records it in the
 * process-wide tuning table under (phase, bin, mean B-row length class).
 * Descriptors with RVSP_SPGEMM_ALGO_HIP_TUNED look their bins up in that
 * table at work_estimation. C holds the product on return.
 */
rvsp_status_t rvsp_spgemm_autotune(rvsp_spgemm_descr_t descr,
                                   const rvsp_csr_matrix_t *A,
                                   const rvsp_csr_matrix_t *B,
                                   rvsp_csr_matrix_t *C, void *workspace,
                                   int32_t reps);

/* Tuning table persistence. path NULL means $RVSP_TUNING_TABLE, else
 * tuning_table.txt next to the library (also read automatically on the first
 * HIP_TUNED estimation). save/load return the entry count or -1. */
int32_t rvsp_tuning_table_save(const char *path);
int32_t rvsp_tuning_table_load(const char *path);
void rvsp_tuning_table_clear(void);
int32_t rvsp_tuning_table_size(void);

/*
 * Batched products: C[p] = A[p] * B[p] for p < count, all pairs served by
 * one set of launches per phase instead of one per pair. A, B and C are
 * arrays of `count` matrices; RVSP_MEMORY_DEVICE descriptors only
 * (RVSP_ERROR_INVALID_ARGUMENT otherwise). Every pair's C is bit-identical to
 * what rvsp_spgemm_compute gives for that pair alone, and every check of the
 * single-product calls applies per pair.
 *
 * The pairs are stacked into one block-diagonal product, cut into groups of
 * consecutive pairs whose product count fits int32 row pointers. Pairs
 * passing the same B arrays share one stacked copy of B. When the C arrays
 * of the pairs of a group lie back to back in memory (C[p+1].col_idx ==
 * C[p].col_idx + c_nnz[p], likewise values), they are written in place;
 * otherwise through one segmented copy.
 *
 * Repeated batch computes with the same arrays are numeric-only (like
 * rvsp_spgemm_compute's reuse of C.col_idx); from the second such compute on
 * they replay one captured HIP graph.
 */
rvsp_status_t rvsp_spgemm_batch_work_estimation(rvsp_spgemm_descr_t descr,
                                                int32_t count,
                                                const rvsp_csr_matrix_t *A,
                                                const rvsp_csr_matrix_t *B,
                                                int32_t *c_nnz);
rvsp_status_t rvsp_spgemm_batch_compute(rvsp_spgemm_descr_t descr,
                                        int32_t count,
                                        const rvsp_csr_matrix_t *A,
                                        const rvsp_csr_matrix_t *B,
                                        rvsp_csr_matrix_t *C);
/* Groups formed by the last batch estimation (-1 before one). */
int32_t rvsp_spgemm_batch_groups(rvsp_spgemm_descr_t descr);
/* Batch computes served by a graph launch since the last batch estimation. */
int32_t rvsp_spgemm_batch_graph_launches(rvsp_spgemm_descr_t descr);
/* on == 0: repeated batch computes launch eagerly instead of from a graph. */
rvsp_status_t rvsp_spgemm_batch_set_graph(rvsp_spgemm_descr_t descr, int on);
/* on != 0: the groups' long-row numeric bins (workgroup tiles) run as one
 * fused launch per tile size after every group's other work, so the long
 * rows of all groups share one grid; 0 (default) launches them per group
 * and bin, overlapped with the other groups on the batch's lanes. Output
 * bits are the same either way. Fusing saves launches on many small groups
 * but gives up the overlap. Must precede the batch estimation. */
rvsp_status_t rvsp_spgemm_batch_set_fuse(rvsp_spgemm_descr_t descr, int on);
/* How batch computes run the groups' numeric bins of rows up to 4096 entries
 * long (and of rows fed by a single B row): as one persistent launch over a
 * product-balanced work queue of (group, bin, row block) items, captured with
 * the rest of a numeric-only compute in the batch's graph; the first compute
 * finds each product's position by a search of its C row, numeric-only
 * computes read a position map (16 bits per product) built by the first of
 * them. mode 1 (default): the first numeric-only compute also times the
 * groups' per-bin launches and keeps the faster path; 2: always the
 * persistent launch; 0: per-bin launches. Output bits are the same in every
 * mode. No reference counterpart (batching is not part of the reference). */
rvsp_status_t rvsp_spgemm_batch_set_persistent(rvsp_spgemm_descr_t descr, int mode);
/* Segments (group, bin) served by the persistent launch; 0 before the
 * first numeric-only compute or when it is off. */
int32_t rvsp_spgemm_batch_stream_segments(rvsp_spgemm_descr_t descr);
/* The numeric path of the batch's computes: 1 the persistent launch, 2 the
 * per-bin launches after measuring them faster, 0 neither (mode 0, or not
 * estimated). The times the first numeric-only compute measured (0 when it
 * made no choice) go to *persistent_ms and *per_bin_ms when not NULL. */
int32_t rvsp_spgemm_batch_numeric_path(rvsp_spgemm_descr_t descr, float* persistent_ms,
                                       float* per_bin_ms);

/* Caps the product count of a group (default and maximum INT32_MAX; 0
 * restores the default). Must precede the batch estimation. */
rvsp_status_t rvsp_spgemm_batch_set_group_limit(rvsp_spgemm_descr_t descr,
                                                int64_t max_products);

/* Device memory helpers for callers without a HIP binding of their own
 * (for example ctypes). kind: 1 host->device, 2 device->host,
 * 3 device->device. */
rvsp_status_t rvsp_hip_malloc(void **ptr, size_t bytes);
rvsp_status_t rvsp_hip_free(void *ptr);
rvsp_status_t rvsp_hip_memcpy(void *dst, const void *src, size_t bytes,
                              int kind);
rvsp_status_t rvsp_hip_synchronize(void);

/* Descriptors take their device state from a cache of released buffers, so
 * repeated estimations do not pay for fresh multi-GB device allocations
 * (no counterpart in the reference, which allocates host memory per call).
 * The cache holds at most half of the device's memory (environment
 * variable RVSP_POOL_CAP_MB overrides) and is emptied when an allocation
 * runs out of memory. rvsp_hip_pool_release returns every cached buffer to
 * the device; rvsp_hip_pool_cached_bytes reports how much is held. */
rvsp_status_t rvsp_hip_pool_release(void);
size_t rvsp_hip_pool_cached_bytes(void);

/* Canonical-CSR check of device-resident arrays: the verdict and failing
 * row that rvsp_csr_check (src/kernels/spgemm/rvsp_support.c:36-126) gives
 * for the same arrays on the host, computed on the GPU. Rows are judged in
 * order and the first failure wins; within a row a decreasing row pointer
 * comes first, then per entry out of range, repeated (DUPLICATE),
 * descending (UNSORTED). col_idx holds nnz entries: with a column array, a
 * row pointer beyond nnz is reported as RVSP_CSR_BAD_ROW_PTR at that row
 * (the host check has no nnz and would read past the array). *bad_row_out (may be NULL) gets the
 * failing row, or -1. The rvsp_spgemm_csr path runs this check on its
 * staged operands instead of scanning them on the host. Synchronous; the
 * return value reports HIP failures only. */
rvsp_status_t rvsp_csr_check_device(int32_t rows, int32_t cols, int32_t nnz,
                                    const int32_t *row_ptr, const int32_t *col_idx,
                                    rvsp_csr_status_t *status_out,
                                    int32_t *bad_row_out);

/* rvsp_csr_check_device without any device allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * verdicts and return values as rvsp_csr_check_device. Its scratch (the
 * check's 8-byte key) is a 256-byte static device array per device, held
 * until the process ends; no matrix size exceeds it. Calls that overlap on
 * one device share that array and race. */
rvsp_status_t rvsp_csr_check_device_noalloc(int32_t rows, int32_t cols, int32_t nnz,
                                            const int32_t *row_ptr,
                                            const int32_t *col_idx,
                                            rvsp_csr_status_t *status_out,
                                            int32_t *bad_row_out);

/* Number of usable gfx950 devices (0 when none; never an error for "no
 * GPU"). */
int32_t rvsp_hip_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_HIP_H */
