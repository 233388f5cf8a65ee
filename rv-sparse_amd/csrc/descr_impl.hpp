// descr_impl.hpp — layout of the SpGEMM descriptor and the internal calls
// shared by descr.cpp, the tuner, the learned selector and the batch path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "kernels.hpp"
#include "plan.hpp"
#include "rv_sparse.h"
#include "rv_sparse_hip.h"

namespace rvsp {

struct BatchState;

// A long-row numeric bin whose launch the caller (a batch) fuses with other
// descriptors' (defer_block): its rows and the arrays the kernel needs.
struct DeferredBin {
  const int* rows = nullptr;
  int nrows = 0;
  int max_row = 0;
  BinSpec spec;
  const int* a_rp = nullptr;
  const int* c_rp = nullptr;
  const int* c_ci = nullptr;
  float* c_val = nullptr;
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Device buffers come from the library's cache (devpool.cpp). ensure grows
// b to at least `bytes` (contents are not kept); release hands b back as
// possibly still in use by queued work; release_idle when the caller knows
// the GPU is done with it.
rvsp_status_t ensure(DevBuf& b, size_t bytes);
void release(DevBuf& b);
void release_idle(DevBuf& b);
hipError_t pool_alloc(void** out, size_t* bytes_io);
void pool_free(void* ptr, size_t bytes, bool idle);
size_t pool_idle_bytes();
// Free device memory including the library's cached buffers.
size_t device_free_bytes();
rvsp_status_t device_ready();

// One phase's bin table plus the rows assigned to each bin.
struct PhasePlan {
  PhaseTable table;
  std::vector<int> offsets;  // host copy, nbins + 1 entries
  DevBuf rows;               // row ids grouped by bin, ascending within a bin
  // Per bin: r when the bin's rows are exactly r, r + 1, ... (kernels that
  // can use it then skip reading the row list), else -1.
  std::vector<int> row_base;
};

}  // namespace rvsp

// The reference's descriptor (src/core/spgemm_descr.c:19-37: algo,
// estimated flag, c_row_ptr, c_nnz, op counts, the filled_into reuse key,
// captured dimensions, workspace size) plus the device state of the HIP
// path. op counts live on the device and are copied out on request.
struct rvsp_spgemm_descr {
  rvsp_spgemm_algo_t algo = RVSP_SPGEMM_ALGO_DEFAULT;
  int estimated = 0;
  int32_t* c_row_ptr = nullptr;  // host copy (host memory mode only)
  int32_t c_nnz = 0;
  const int32_t* filled_into = nullptr;
  const int32_t* rowptr_into = nullptr;  // device mode: C.row_ptr already written
  bool dev_cols_valid = false;           // host mode: c_ci_dev holds the columns
  int32_t a_rows = 0, a_cols = 0, a_nnz = 0;
  int32_t b_rows = 0, b_cols = 0, b_nnz = 0;
  size_t workspace_bytes = 0;

  rvsp_memory_t memory = RVSP_MEMORY_HOST;
  hipStream_t own_stream = nullptr;
  hipStream_t user_stream = nullptr;
  int async = 0;
  std::vector<rvsp_bin_choice_t> choices;

  // Host-mode staging of the operands.
  rvsp::DevBuf a_rp, a_ci, a_val, b_rp, b_ci, b_val;
  bool b_struct_alias = false, b_val_alias = false;

  // Analysis and symbolic state.
  rvsp::DevBuf ops, lo, hi, maxb, long_list, n_long;
  rvsp::DevBuf cnt, c_rp, scan_tiles, total;
  rvsp::DevBuf bin_of, tile_hist, tile_base, bin_offsets;
  int tile_rows = 0, ntiles = 0;
  rvsp::PhasePlan count, fill, numeric;

  // Host-mode device copy of C.
  rvsp::DevBuf c_ci_dev, c_val_dev;

  // Global scratch of GLOBAL bins.
  rvsp::DevBuf gbits, gdense;
  rvsp::GlobalSlots gs_bits, gs_dense;

  // Per-bin feature scratch (tuned and learned selectors); learn_feat holds
  // the selector network's parameters (of model version learn_version; 0 =
  // none) and one prediction's inputs and outputs.
  rvsp::DevBuf learn_feat, bin_stats;
  unsigned long long learn_version = 0;

  // Position map of the MAP numeric bins (structure cache for repeated
  // numeric-only computes): offsets, entries, coverage and validity.
  rvsp::DevBuf map, poff, map_weights;
  int map_width = 0;        // bytes per entry; 0 = no map planned
  unsigned map_bins = 0;    // numeric bins whose rows have map entries
  bool map_valid = false;   // entries built for the current structure
  bool map_promoted = false;  // searching bins already moved to the map
  bool position_map = true;   // promote on the first numeric-only compute
  // B-row cache of the map kernels (per A entry: B-row start, 16-bit
  // length), built with the map entries; brow_ok: usable by the map bins.
  rvsp::DevBuf brow;
  bool brow_ok = false;

  // Split table of B's long rows for the tiled numeric kernel's narrowing
  // searches (per-row counts, offsets, table).
  rvsp::DevBuf split_cnt, split_off, split_sp;
  rvsp::SplitTable split;
  bool split_valid = false;
  // A descriptor over the same B whose table this one uses (the groups of a
  // batch share one stacked B); not owned.
  rvsp_spgemm_descr* split_owner = nullptr;

  // Batched products (batch.cpp): stacked operands and one inner
  // descriptor per group of pairs; settings apply at the next estimation.
  rvsp::BatchState* batch = nullptr;
  bool batch_graph = true;
  bool batch_fuse = false;    // fuse the groups' long-row numeric launches
  long long batch_limit = 0;  // products per group; 0 = INT32_MAX

  // Set on a batch's inner descriptors whose groups run on two streams at
  // once: fill bins planned as a wide LDS bitmap (one workgroup per CU, so
  // it waits for a whole CU to drain of the other stream's workgroups) run
  // their bitmaps in global scratch instead and share CUs.
  bool shared_cu = false;

  // Set on a batch's inner descriptors: numeric bins on workgroup tiles
  // (KIND_BLOCK variant 0) are not launched by run_numeric but appended to
  // `deferred`, and the batch launches all groups' such bins at once.
  bool defer_block = false;
  std::vector<rvsp::DeferredBin> deferred;

  // Set on a batch's inner descriptors: numeric bins served by the batch's
  // persistent row-stream launch (run_numeric skips them, and they never move
  // to the position map).
  unsigned stream_bins = 0;
  // Batch setting (rvsp_spgemm_batch_set_persistent): 0 per-bin launches,
  // 1 the persistent row-stream pass unless the first numeric-only compute
  // measures the per-bin launches faster, 2 the persistent pass always.
  int batch_persistent = 1;

  // Set by the one-shot path: work_estimation checks the staged operands
  // for canonical CSR on the device before any other kernel reads them.
  bool check_csr = false;
  rvsp::DevBuf check_key;

  // Sorted columns of the count phase's LDS-bitmap rows, kept for the fill
  // phase (offsets koff: exclusive scan of min(Op_i, span_i) over them).
  rvsp::DevBuf keep, koff, keep_w;
  unsigned keep_bins = 0;  // count bins whose rows keep their columns
  bool keep_ok = false;

  // Captured numeric-only graph.
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const void* graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace rvsp {

rvsp_status_t descr_stream(rvsp_spgemm_descr* d, hipStream_t* out);
RowStats row_stats(rvsp_spgemm_descr* d);
rvsp_status_t run_fill(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                       int* c_ci, hipStream_t s);
rvsp_status_t run_numeric(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                          const int* c_ci, float* c_val, hipStream_t s);

// Plans the position map for the numeric bins of kind MAP (offsets and
// storage); bins that do not fit the memory budget fall back to SEARCH.
rvsp_status_t plan_map(rvsp_spgemm_descr* d, hipStream_t s);
// Builds the map entries if they are not valid for the current structure.
rvsp_status_t ensure_split(rvsp_spgemm_descr* d, const DevCsr& B, hipStream_t s);
rvsp_status_t ensure_map_built(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                               const int* c_ci, hipStream_t s);
// Drops the position map: its bins search C's columns again, and the next
// numeric-only compute may promote them anew.
rvsp_status_t release_position_map(rvsp_spgemm_descr* d);
// Moves the non-empty searching numeric bins to the position map (chunked
// variant, same lanes and capacity) and builds the map; bins over the memory
// budget keep searching. Synchronous.
rvsp_status_t promote_to_map(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                             const int* c_ci, hipStream_t s);

// rvsp_csr_check's verdict (and failing row, or -1) for device arrays,
// computed on the GPU; key_buf is 8 bytes of scratch. Synchronous on s.
rvsp_status_t csr_check_device(int rows, int cols, int nnz, const int* rp, const int* ci,
                               DevBuf& key_buf, hipStream_t s, rvsp_csr_status_t* verdict,
                               int* bad_row);
// The same check with the caller's 8-byte device key (a 64-bit atomic
// minimum, set before each pass); allocates nothing.
rvsp_status_t csr_check_device_key(int rows, int cols, int nnz, const int* rp, const int* ci,
                                   unsigned long long* dkey, hipStream_t s,
                                   rvsp_csr_status_t* verdict, int* bad_row);
// rvsp_spgemm_csr for FP64 x FP64 -> FP64 (i8 false) and INT8 x INT8 ->
// INT32 (i8 true) operands (legacy.cpp).
rvsp_status_t oneshot_legacy(const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B,
                             rvsp_csr_matrix_t* C, bool i8);

// Re-bins one phase of an estimated descriptor under table t (classification,
// partition, offsets, GLOBAL scratch); t becomes the phase's table.
rvsp_status_t repartition_phase(rvsp_spgemm_descr* d, int phase, const DevCsr& A,
                                const DevCsr& B, const PhaseTable& t, hipStream_t s);
// Key of the tuned bin thresholds: floor(log2(B's mean row length)) in [0, 10].
int rule_len_class(int b_rows, int b_nnz);
// The tuned rule parameter of a phase for that key (apply_rule), or -1.
int tuned_rule(int phase, int key);

// Rows of bin b of a partitioned phase.
inline int bin_rows(const PhasePlan& p, int b) { return p.offsets[b + 1] - p.offsets[b]; }

// The descriptor's position map (and B-row cache, when built) as the map
// kernels take it; row_base left at -1.
MapView map_view(const rvsp_spgemm_descr* d);

// row_base of the numeric bin whose row list starts at `rows` (-1 if none).
int numeric_row_base(const rvsp_spgemm_descr* d, const int* rows);

void destroy_graph(rvsp_spgemm_descr* d);
void destroy_batch(rvsp_spgemm_descr* d);
void invalidate_batch(rvsp_spgemm_descr* d);

}  // namespace rvsp
