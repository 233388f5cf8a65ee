// descr.cpp — the two-phase SpGEMM descriptor, the one-shot entry points and
// the HIP extension calls. This is the drop-in boundary: the check order and
// status codes follow the reference's src/core/spgemm_descr.c and
// src/core/spgemm.c line by line (cited per function); the work behind them
// runs on the GPU.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "descr_impl.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "rv_sparse.h"
#include "rv_sparse_hip.h"
#include "rvsp_spgemm.h"
#include "selector.hpp"

using namespace rvsp;

namespace rvsp {

rvsp_status_t ensure(DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 4;
  if (b.bytes >= bytes) return RVSP_SUCCESS;
  release(b);
  size_t got = bytes;
  hipError_t e = pool_alloc(&b.p, &got);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return from_hip(e);
  }
  b.bytes = got;
  return RVSP_SUCCESS;
}

void release(DevBuf& b) {
  pool_free(b.p, b.bytes, false);
  b.p = nullptr;
  b.bytes = 0;
}

void release_idle(DevBuf& b) {
  pool_free(b.p, b.bytes, true);
  b.p = nullptr;
  b.bytes = 0;
}

size_t device_free_bytes() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    free_b = 0;
  }
  return free_b + pool_idle_bytes();
}

rvsp_status_t device_ready() {
  static std::once_flag once;
  static rvsp_status_t status = RVSP_ERROR_DEVICE;
  std::call_once(once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
      (void)hipGetLastError();
      return;
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&prop, dev) != hipSuccess)
      return;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return;
    // Dynamic LDS above the default needs an opt-in per kernel; the limit
    // the plans may use is what the device grants a workgroup, less the
    // kernels' static LDS.
    const int limit = (int)prop.sharedMemPerBlock - kStaticLdsReserve;
    if (init_kernel_attributes(limit) != hipSuccess) {
      (void)hipGetLastError();
      set_lds_limit(64 * 1024 - kStaticLdsReserve);
    } else {
      set_lds_limit(limit);
    }
    status = RVSP_SUCCESS;
  });
  return status;
}

rvsp_status_t descr_stream(rvsp_spgemm_descr* d, hipStream_t* out) {
  if (d->user_stream) {
    *out = d->user_stream;
    return RVSP_SUCCESS;
  }
  if (!d->own_stream) {
    RVSP_TRY(device_ready());
    RVSP_HIP_TRY(hipStreamCreateWithFlags(&d->own_stream, hipStreamNonBlocking));
  }
  *out = d->own_stream;
  return RVSP_SUCCESS;
}

RowStats row_stats(rvsp_spgemm_descr* d) {
  RowStats st;
  st.ops = (long long*)d->ops.p;
  st.lo = (int*)d->lo.p;
  st.hi = (int*)d->hi.p;
  st.maxb = (int*)d->maxb.p;
  return st;
}

// rvsp_csr_check's verdict for device arrays, found on the GPU: a row pass
// (row pointers; a pointer past nnz counts as malformed, since no more
// entries exist), then an entry pass over the rows before the first
// malformed pointer, each proposing its failure as (row << 32) | code; the
// smallest is where the reference's in-order scan stops. The kind of an
// entry failure is then read off that entry and its predecessor.
rvsp_status_t csr_check_device(int rows, int cols, int nnz, const int* rp, const int* ci,
                               DevBuf& key_buf, hipStream_t s, rvsp_csr_status_t* verdict,
                               int* bad_row) {
  if (rows > 0 && rp != NULL) RVSP_TRY(ensure(key_buf, 8));
  return csr_check_device_key(rows, cols, nnz, rp, ci, (unsigned long long*)key_buf.p, s,
                              verdict, bad_row);
}

rvsp_status_t csr_check_device_key(int rows, int cols, int nnz, const int* rp, const int* ci,
                                   unsigned long long* dkey, hipStream_t s,
                                   rvsp_csr_status_t* verdict, int* bad_row) {
  *verdict = RVSP_CSR_OK;
  *bad_row = -1;
  if (rows < 0 || cols < 0 || nnz < 0 || rp == NULL) {
    *verdict = RVSP_CSR_BAD_ROW_PTR;
    return RVSP_SUCCESS;
  }
  int head = 0;
  RVSP_HIP_TRY(hipMemcpyAsync(&head, rp, 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  if (head != 0) {
    *verdict = RVSP_CSR_BAD_ROW_PTR;
    return RVSP_SUCCESS;
  }
  if (rows == 0) return RVSP_SUCCESS;
  unsigned long long key = ~0ull;
  RVSP_HIP_TRY(hipMemcpyAsync(dkey, &key, 8, hipMemcpyHostToDevice, s));
  RVSP_HIP_TRY(launch_csr_check_rows(rows, nnz, rp, ci != NULL, dkey, s));
  RVSP_HIP_TRY(hipMemcpyAsync(&key, dkey, 8, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  const int limit = key == ~0ull ? rows : (int)(key >> 32);
  if (ci != NULL && limit > 0) {
    RVSP_HIP_TRY(launch_csr_check_cols(limit, cols, nnz, rp, ci, dkey, s));
    RVSP_HIP_TRY(hipMemcpyAsync(&key, dkey, 8, hipMemcpyDeviceToHost, s));
    RVSP_HIP_TRY(hipStreamSynchronize(s));
  }
  if (key == ~0ull) return RVSP_SUCCESS;
  const int row = (int)(key >> 32);
  const unsigned code = (unsigned)(key & 0xffffffffu);
  *verdict = RVSP_CSR_BAD_ROW_PTR;
  if (code == 0) *bad_row = row;
  if (code < 2) return RVSP_SUCCESS;  // code 1: no column array (no row named)
  // Entry j of the row failed with every earlier entry passing.
  const int j = (int)(code - 2);
  int start = 0, pair[2] = {-1, -1};
  RVSP_HIP_TRY(hipMemcpyAsync(&start, rp + row, 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  const int p = start + j;
  if (j > 0)
    RVSP_HIP_TRY(hipMemcpyAsync(pair, ci + p - 1, 8, hipMemcpyDeviceToHost, s));
  else
    RVSP_HIP_TRY(hipMemcpyAsync(pair + 1, ci + p, 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  const int prev = pair[0], c = pair[1];
  *bad_row = row;
  if (c < 0 || c >= cols)
    *verdict = RVSP_CSR_COL_OUT_OF_RANGE;
  else if (c == prev)
    *verdict = RVSP_CSR_DUPLICATE;
  else
    *verdict = RVSP_CSR_UNSORTED;
  return RVSP_SUCCESS;
}

}  // namespace rvsp

namespace {

rvsp_status_t check_operand(const rvsp_csr_matrix_t* M) {
  // Reference: src/core/spgemm_descr.c:137-160.
  if (M == NULL) return RVSP_ERROR_NULL_POINTER;
  if (M->row_ptr == NULL || M->col_idx == NULL || M->values == NULL)
    return RVSP_ERROR_NULL_POINTER;
  if (M->dtype != RVSP_DTYPE_FP32) return RVSP_ERROR_UNSUPPORTED_DTYPE;
  if (M->rows <= 0 || M->cols <= 0 || M->nnz < 0) return RVSP_ERROR_INVALID_ARGUMENT;
  return RVSP_SUCCESS;
}

bool algo_available(rvsp_spgemm_algo_t algo) {
  switch (algo) {
    case RVSP_SPGEMM_ALGO_DEFAULT:
    case RVSP_SPGEMM_ALGO_HIP_FIXED:
    case RVSP_SPGEMM_ALGO_HIP_TUNED:
    case RVSP_SPGEMM_ALGO_HIP_LEARNED:
      return true;
    default:
      // RVV, CONTIG and ADAPTIVE are RISC-V vector strategies.
      return false;
  }
}

void drop_structure(rvsp_spgemm_descr* d) {
  free(d->c_row_ptr);
  d->c_row_ptr = NULL;
  d->filled_into = NULL;
  d->rowptr_into = NULL;
  d->dev_cols_valid = false;
  d->map_valid = false;
  d->map_bins = 0;
  d->map_width = 0;
  d->map_promoted = false;
  d->brow_ok = false;
  d->split_valid = false;
  d->keep_ok = false;
  d->keep_bins = 0;
  d->estimated = 0;
  d->c_nnz = 0;
}

// Uploads one host array into a descriptor-owned device buffer.
rvsp_status_t upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
  RVSP_TRY(ensure(b, bytes));
  if (bytes) RVSP_HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
  return RVSP_SUCCESS;
}

// Device views of A and B for this call. Host mode stages the structure (at
// estimation) and the values (at compute) into descriptor buffers; when B is
// the same host matrix as A (C = A * A) the buffers are shared.
rvsp_status_t operand_views(rvsp_spgemm_descr* d, const rvsp_csr_matrix_t* A,
                            const rvsp_csr_matrix_t* B, bool structure,
                            bool values, DevCsr* dA, DevCsr* dB, hipStream_t s) {
  dA->rows = A->rows; dA->cols = A->cols; dA->nnz = A->nnz;
  dB->rows = B->rows; dB->cols = B->cols; dB->nnz = B->nnz;
  if (d->memory == RVSP_MEMORY_DEVICE) {
    dA->rp = A->row_ptr; dA->ci = A->col_idx; dA->val = (const float*)A->values;
    dB->rp = B->row_ptr; dB->ci = B->col_idx; dB->val = (const float*)B->values;
    return RVSP_SUCCESS;
  }
  const bool same_struct = A->row_ptr == B->row_ptr && A->col_idx == B->col_idx &&
                           A->rows == B->rows && A->nnz == B->nnz;
  const bool same_vals = same_struct && A->values == B->values;
  if (structure) {
    RVSP_TRY(upload(d->a_rp, A->row_ptr, ((size_t)A->rows + 1) * 4, s));
    RVSP_TRY(upload(d->a_ci, A->col_idx, (size_t)A->nnz * 4, s));
    if (!same_struct) {
      RVSP_TRY(upload(d->b_rp, B->row_ptr, ((size_t)B->rows + 1) * 4, s));
      RVSP_TRY(upload(d->b_ci, B->col_idx, (size_t)B->nnz * 4, s));
    }
    d->b_struct_alias = same_struct;
  }
  if (values) {
    RVSP_TRY(upload(d->a_val, A->values, (size_t)A->nnz * 4, s));
    if (!same_vals) RVSP_TRY(upload(d->b_val, B->values, (size_t)B->nnz * 4, s));
    d->b_val_alias = same_vals;
  }
  dA->rp = (const int*)d->a_rp.p; dA->ci = (const int*)d->a_ci.p;
  dA->val = (const float*)d->a_val.p;
  DevBuf& brp = d->b_struct_alias ? d->a_rp : d->b_rp;
  DevBuf& bci = d->b_struct_alias ? d->a_ci : d->b_ci;
  DevBuf& bval = d->b_val_alias ? d->a_val : d->b_val;
  dB->rp = (const int*)brp.p; dB->ci = (const int*)bci.p;
  dB->val = (const float*)bval.p;
  return RVSP_SUCCESS;
}

rvsp_status_t partition_phase(rvsp_spgemm_descr* d, PhasePlan& plan, int M,
                              int offsets_slot, hipStream_t s) {
  RVSP_TRY(ensure(plan.rows, (size_t)M * 4));
  PartitionScratch ps;
  ps.bin_of = (unsigned char*)d->bin_of.p;
  ps.tile_hist = (int*)d->tile_hist.p;
  ps.tile_base = (int*)d->tile_base.p;
  ps.bin_offsets = (int*)d->bin_offsets.p + offsets_slot * (kMaxBins + 1);
  ps.tile_rows = d->tile_rows;
  ps.ntiles = d->ntiles;
  RVSP_HIP_TRY(launch_partition(M, plan.table.rule.nbins, ps, (int*)plan.rows.p, s));
  return RVSP_SUCCESS;
}

rvsp_status_t fetch_offsets(rvsp_spgemm_descr* d, PhasePlan* plans[], int n,
                            int first_slot, hipStream_t s) {
  std::vector<int> host((size_t)n * (kMaxBins + 1));
  RVSP_HIP_TRY(hipMemcpyAsync(host.data(),
                              (const int*)d->bin_offsets.p + first_slot * (kMaxBins + 1),
                              host.size() * 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  for (int i = 0; i < n; ++i) {
    const int nb = plans[i]->table.rule.nbins;
    plans[i]->offsets.assign(host.begin() + (size_t)i * (kMaxBins + 1),
                             host.begin() + (size_t)i * (kMaxBins + 1) + nb + 1);
  }
  return RVSP_SUCCESS;
}


// Finds the bins whose rows are one run of consecutive rows (row lists are
// ascending and distinct, so first and last decide). Synchronous.
rvsp_status_t fetch_row_bases(PhasePlan& p, hipStream_t s) {
  const int nb = p.table.rule.nbins;
  p.row_base.assign(nb, -1);
  std::vector<int> ends(2 * (size_t)nb, 0);
  bool any = false;
  for (int b = 1; b < nb; ++b) {
    const int n = bin_rows(p, b);
    if (n == 0) continue;
    const int* r = (const int*)p.rows.p + p.offsets[b];
    RVSP_HIP_TRY(hipMemcpyAsync(&ends[2 * b], r, 4, hipMemcpyDeviceToHost, s));
    RVSP_HIP_TRY(hipMemcpyAsync(&ends[2 * b + 1], r + n - 1, 4, hipMemcpyDeviceToHost, s));
    any = true;
  }
  if (!any) return RVSP_SUCCESS;
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  for (int b = 1; b < nb; ++b) {
    const int n = bin_rows(p, b);
    if (n > 0 && ends[2 * b + 1] - ends[2 * b] == n - 1) p.row_base[b] = ends[2 * b];
  }
  return RVSP_SUCCESS;
}

// Sizes the global scratch slots a phase's GLOBAL bins need. Bitmaps are
// cleared once at allocation; the kernels leave them cleared.
rvsp_status_t ensure_slots(rvsp_spgemm_descr* d, const PhasePlan& p, int b_cols,
                           bool dense, hipStream_t s) {
  int rows = 0;
  for (int b = 2; b < p.table.rule.nbins; ++b)
    if (p.table.spec[b].kind == KIND_GLOBAL) rows += bin_rows(p, b);
  if (rows == 0) return RVSP_SUCCESS;
  if (b_cols <= 0) return RVSP_ERROR_INVALID_ARGUMENT;
  const size_t n = (size_t)b_cols;
  const size_t slot = dense ? ((n * 4 + 255) & ~(size_t)255)
                            : ((((n + 31) / 32) * 4 + 255) & ~(size_t)255);
  const size_t budget = (size_t)16 << 30;
  int want = rows < 1024 ? rows : 1024;
  size_t fit = budget / slot;
  if (fit < 1) fit = 1;
  if ((size_t)want > fit) want = (int)fit;
  GlobalSlots& gs = dense ? d->gs_dense : d->gs_bits;
  DevBuf& buf = dense ? d->gdense : d->gbits;
  if (gs.nslots >= want && gs.slot_bytes == slot) return RVSP_SUCCESS;
  release(buf);
  RVSP_TRY(ensure(buf, slot * want));
  if (!dense) RVSP_HIP_TRY(hipMemsetAsync(buf.p, 0, slot * want, s));
  gs.base = buf.p;
  gs.slot_bytes = slot;
  gs.nslots = want;
  return RVSP_SUCCESS;
}

// Bins whose LDS the device cannot grant run as GLOBAL bins.
void clamp_to_device(PhaseTable& t) {
  for (int b = 2; b < t.rule.nbins; ++b)
    if (t.spec[b].lds_bytes > lds_limit()) {
      t.spec[b].kind = KIND_GLOBAL;
      t.spec[b].group = kBlock;
      t.spec[b].cap = 0;
      t.spec[b].lds_bytes = 0;
    }
}

// Per-bin shape selection of one partitioned phase: the tuned table or the
// learned selector (by algo), then the caller's explicit overrides, then the
// device's LDS limit. Every step goes through apply_choice, which vetoes a
// shape that cannot hold the bin's rows.
rvsp_status_t select_phase(rvsp_spgemm_descr* d, PhasePlan& p, int phase, const DevCsr& A,
                           const int* c_rp, hipStream_t s) {
  const bool tuned = d->algo == RVSP_SPGEMM_ALGO_HIP_TUNED;
  const bool learned = d->algo == RVSP_SPGEMM_ALGO_HIP_LEARNED;
  if (tuned || learned) {
    std::vector<BinFeatures> f;
    RVSP_TRY(bin_features(d, p, phase, A.rp, c_rp, s, &f));
    if (tuned) select_tuned(p.table, phase, f);
    if (learned) RVSP_TRY(select_learned(d, p.table, phase, f, s));
  }
  for (const rvsp_bin_choice_t& c : d->choices) {
    if (c.phase != phase || c.bin < 1 || c.bin >= p.table.rule.nbins) continue;
    BinChoice bc;
    bc.kind = c.kind;
    bc.group = c.group;
    bc.cap = c.cap;
    bc.variant = c.variant;
    apply_choice(p.table, phase, c.bin, bc, p.table.rule.bound[c.bin]);
  }
  clamp_to_device(p.table);
  return RVSP_SUCCESS;
}

// Decides whether the count phase's LDS-bitmap rows keep their sorted
// columns for the fill phase (one bitmap pass instead of two): offsets from
// the rows' column bounds min(Op_i, span_i), storage within
// min(16 GiB, free / 4); otherwise those rows are filled as usual.
rvsp_status_t plan_keep(rvsp_spgemm_descr* d, int M, hipStream_t s) {
  d->keep_ok = false;
  d->keep_bins = 0;
  unsigned mask = 0;
  for (int b = 2; b < d->count.table.rule.nbins; ++b)
    if (d->count.table.spec[b].kind == KIND_BITMAP && bin_rows(d->count, b) > 0) mask |= 1u << b;
  if (mask == 0) return RVSP_SUCCESS;
  RVSP_TRY(ensure(d->koff, ((size_t)M + 1) * 8));
  RVSP_TRY(ensure(d->keep_w, (size_t)M * 8));
  RVSP_HIP_TRY(launch_map_offsets(M, (const unsigned char*)d->bin_of.p, mask, row_stats(d),
                                  (long long*)d->keep_w.p, (long long*)d->koff.p,
                                  (long long*)d->scan_tiles.p, (long long*)d->total.p, s,
                                  true));
  long long total = 0;
  RVSP_HIP_TRY(hipMemcpyAsync(&total, d->total.p, 8, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  const size_t free_b = device_free_bytes();
  const size_t bytes = (size_t)total * 4;
  const size_t budget = std::min<size_t>((size_t)16 << 30, free_b / 4);
  if (bytes > budget || ensure(d->keep, bytes) != RVSP_SUCCESS) return RVSP_SUCCESS;
  d->keep_bins = mask;
  d->keep_ok = true;
  return RVSP_SUCCESS;
}

// The symbolic phase on the device: analysis, binning, exact count, scan,
// then the fill and numeric plans. On success *total holds nnz(C).
rvsp_status_t estimate(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                       long long* total, hipStream_t s) {
  const int M = A.rows;
  RVSP_TRY(ensure(d->ops, (size_t)M * 8));
  RVSP_TRY(ensure(d->lo, (size_t)M * 4));
  RVSP_TRY(ensure(d->hi, (size_t)M * 4));
  RVSP_TRY(ensure(d->maxb, (size_t)M * 4));
  RVSP_TRY(ensure(d->long_list, (size_t)M * 4));
  RVSP_TRY(ensure(d->n_long, 4));
  RVSP_TRY(ensure(d->cnt, (size_t)M * 4));
  RVSP_TRY(ensure(d->c_rp, ((size_t)M + 1) * 4));
  RVSP_TRY(ensure(d->scan_tiles, scan_tile_count(M) * 8));
  RVSP_TRY(ensure(d->total, 8));
  RVSP_TRY(ensure(d->bin_of, (size_t)M));
  d->tile_rows = 1024;
  while ((long long)d->tile_rows * 4096 < M) d->tile_rows *= 2;
  d->ntiles = (M + d->tile_rows - 1) / d->tile_rows;
  RVSP_TRY(ensure(d->tile_hist, (size_t)d->ntiles * kMaxBins * 4));
  RVSP_TRY(ensure(d->tile_base, (size_t)d->ntiles * kMaxBins * 4));
  RVSP_TRY(ensure(d->bin_offsets, 3 * (kMaxBins + 1) * 4));

  const RowStats st = row_stats(d);
  RVSP_HIP_TRY(launch_analyze(A, B, st, (int*)d->long_list.p, (int*)d->n_long.p, s));

  const Strategy strat = fixed_strategy();
  d->count.table = strat.count;
  d->fill.table = strat.fill;
  d->numeric.table = strat.numeric;
  if (d->algo == RVSP_SPGEMM_ALGO_HIP_TUNED) {
    // Tuned bin thresholds, keyed by B's mean row length.
    const int key = rule_len_class(B.rows, B.nnz);
    PhaseTable* tabs[] = {&d->count.table, &d->fill.table, &d->numeric.table};
    for (int ph = 0; ph < 3; ++ph) {
      const int param = tuned_rule(ph, key);
      if (param > 0) apply_rule(*tabs[ph], ph, param);
    }
  }

  // Count.
  RVSP_HIP_TRY(launch_classify_count(A, st, d->count.table.rule,
                                     (unsigned char*)d->bin_of.p, (int*)d->cnt.p, s));
  RVSP_TRY(partition_phase(d, d->count, M, 0, s));
  PhasePlan* cp[] = {&d->count};
  RVSP_TRY(fetch_offsets(d, cp, 1, 0, s));
  RVSP_TRY(select_phase(d, d->count, PHASE_COUNT, A, nullptr, s));
  RVSP_TRY(ensure_slots(d, d->count, B.cols, false, s));
  RVSP_TRY(plan_keep(d, M, s));
  for (int b = 2; b < d->count.table.rule.nbins; ++b) {
    const int n = bin_rows(d->count, b);
    if (n == 0) continue;
    const bool kb = d->keep_ok && ((d->keep_bins >> b) & 1u);
    RVSP_HIP_TRY(launch_count_bin(d->count.table.spec[b],
                                  (const int*)d->count.rows.p + d->count.offsets[b],
                                  n, A, B, st, (int*)d->cnt.p, d->gs_bits, s,
                                  kb ? (int*)d->keep.p : nullptr,
                                  kb ? (const long long*)d->koff.p : nullptr));
  }
  RVSP_HIP_TRY(launch_scan((const int*)d->cnt.p, M, (int*)d->c_rp.p,
                           (long long*)d->total.p, (long long*)d->scan_tiles.p, s));
  RVSP_HIP_TRY(hipMemcpyAsync(total, d->total.p, 8, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  // Reference guard (rvsp_symbolic.c:69-72): an nnz(C) beyond int32 cannot
  // be described by the CSR struct.
  if (*total > INT_MAX) return RVSP_ERROR_ALLOCATION_FAILED;

  // Fill and numeric plans from the exact per-row nnz.
  RVSP_HIP_TRY(launch_classify_nnz(A, st, (const int*)d->c_rp.p, d->fill.table.rule,
                                   (unsigned char*)d->bin_of.p, s,
                                   d->keep_ok ? (const long long*)d->koff.p : nullptr));
  RVSP_TRY(partition_phase(d, d->fill, M, 1, s));
  RVSP_HIP_TRY(launch_classify_nnz(A, st, (const int*)d->c_rp.p,
                                   d->numeric.table.rule,
                                   (unsigned char*)d->bin_of.p, s));
  RVSP_TRY(partition_phase(d, d->numeric, M, 2, s));
  PhasePlan* fp[] = {&d->fill, &d->numeric};
  RVSP_TRY(fetch_offsets(d, fp, 2, 1, s));
  RVSP_TRY(fetch_row_bases(d->numeric, s));
  RVSP_TRY(select_phase(d, d->fill, PHASE_FILL, A, (const int*)d->c_rp.p, s));
  if (d->shared_cu)
    for (int b = 2; b < d->fill.table.rule.nbins; ++b) {
      BinSpec& f = d->fill.table.spec[b];
      if (f.kind == KIND_BITMAP && f.lds_bytes > kWideBitmapBytes) {
        f.kind = KIND_GLOBAL;
        f.group = kBlock;
        f.cap = 0;
        f.lds_bytes = 0;
      }
    }
  RVSP_TRY(select_phase(d, d->numeric, PHASE_NUMERIC, A, (const int*)d->c_rp.p, s));
  RVSP_TRY(plan_map(d, s));
  RVSP_TRY(ensure_slots(d, d->fill, B.cols, false, s));
  RVSP_TRY(ensure_slots(d, d->numeric, B.cols, true, s));
  return RVSP_SUCCESS;
}

// Re-bins one phase of an estimated descriptor under table t (its rows'
// sizes are known): classification, partition, offsets, row bases and the
// global scratch its GLOBAL bins need; t becomes the phase's table.
rvsp_status_t repartition(rvsp_spgemm_descr* d, int phase, const DevCsr& A, const DevCsr& B,
                          const PhaseTable& t, hipStream_t s) {
  const int M = A.rows;
  const RowStats st = row_stats(d);
  PhasePlan* plans[] = {&d->count, &d->fill, &d->numeric};
  PhasePlan& p = *plans[phase];
  p.table = t;
  clamp_to_device(p.table);
  if (phase == PHASE_COUNT)
    RVSP_HIP_TRY(launch_classify_count(A, st, p.table.rule, (unsigned char*)d->bin_of.p,
                                       (int*)d->cnt.p, s));
  else
    RVSP_HIP_TRY(launch_classify_nnz(A, st, (const int*)d->c_rp.p, p.table.rule,
                                     (unsigned char*)d->bin_of.p, s,
                                     phase == PHASE_FILL && d->keep_ok
                                         ? (const long long*)d->koff.p : nullptr));
  RVSP_TRY(partition_phase(d, p, M, phase, s));
  PhasePlan* one[] = {&p};
  RVSP_TRY(fetch_offsets(d, one, 1, phase, s));
  if (phase == PHASE_NUMERIC) RVSP_TRY(fetch_row_bases(p, s));
  RVSP_TRY(ensure_slots(d, p, B.cols, phase == PHASE_NUMERIC, s));
  return RVSP_SUCCESS;
}

}  // namespace

namespace rvsp {

rvsp_status_t repartition_phase(rvsp_spgemm_descr* d, int phase, const DevCsr& A,
                                const DevCsr& B, const PhaseTable& t, hipStream_t s) {
  return repartition(d, phase, A, B, t, s);
}

int rule_len_class(int b_rows, int b_nnz) {
  const double mean = b_rows > 0 ? (double)b_nnz / b_rows : 0.0;
  int c = 0;
  while (c < 10 && mean >= (double)(2 << c)) ++c;
  return c;  // floor(log2(mean)), clamped to [0, 10]
}

rvsp_status_t run_fill(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                       int* c_ci, hipStream_t s) {
  const RowStats st = row_stats(d);
  if (d->keep_ok)
    for (int b = 2; b < d->count.table.rule.nbins; ++b) {
      if (!((d->keep_bins >> b) & 1u)) continue;
      RVSP_HIP_TRY(launch_copy_kept((const int*)d->count.rows.p + d->count.offsets[b],
                                    bin_rows(d->count, b), (const int*)d->keep.p,
                                    (const long long*)d->koff.p, (const int*)d->c_rp.p,
                                    c_ci, s));
    }
  for (int b = 1; b < d->fill.table.rule.nbins; ++b) {
    const int n = bin_rows(d->fill, b);
    if (n == 0) continue;
    RVSP_HIP_TRY(launch_fill_bin(d->fill.table.spec[b],
                                 (const int*)d->fill.rows.p + d->fill.offsets[b], n,
                                 A, B, st, (const int*)d->c_rp.p, c_ci, d->gs_bits, s));
  }
  return RVSP_SUCCESS;
}

rvsp_status_t ensure_split(rvsp_spgemm_descr* d, const DevCsr& B, hipStream_t s) {
  if (d->split_valid) return RVSP_SUCCESS;
  if (d->split_owner) {
    RVSP_TRY(ensure_split(d->split_owner, B, s));
    d->split = d->split_owner->split;
    d->split_valid = d->split_owner->split_valid;
    return RVSP_SUCCESS;
  }
  d->split = SplitTable();
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RVSP_HIP_TRY(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return RVSP_SUCCESS;  // searches run on B
  if (B.rows <= 0 || B.nnz < kSplitMin) {
    d->split_valid = true;  // no row is long enough
    return RVSP_SUCCESS;
  }
  RVSP_TRY(ensure(d->split_cnt, (size_t)B.rows * sizeof(int)));
  RVSP_TRY(ensure(d->split_off, ((size_t)B.rows + 1) * sizeof(int)));
  RVSP_TRY(ensure(d->scan_tiles, scan_tile_count(B.rows) * sizeof(long long)));
  RVSP_TRY(ensure(d->total, sizeof(long long)));
  RVSP_HIP_TRY(launch_split_counts(B, (int*)d->split_cnt.p, s));
  RVSP_HIP_TRY(launch_scan((const int*)d->split_cnt.p, B.rows, (int*)d->split_off.p,
                           (long long*)d->total.p, (long long*)d->scan_tiles.p, s));
  long long entries = 0;
  RVSP_HIP_TRY(hipMemcpyAsync(&entries, d->total.p, sizeof(entries), hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  // The table is an accelerator: without room for it the searches run on B.
  if (entries <= 0 || entries > INT32_MAX / 2 ||
      ensure(d->split_sp, (size_t)entries * sizeof(int)) != RVSP_SUCCESS) {
    d->split_valid = true;
    return RVSP_SUCCESS;
  }
  SplitTable t;
  t.off = (const int*)d->split_off.p;
  t.sp = (const int*)d->split_sp.p;
  RVSP_HIP_TRY(launch_split_build(B, t, (int*)d->split_sp.p, s));
  // Complete before any stream uses it (the groups of a batch may run on
  // other streams and share this table).
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  d->split = t;
  d->split_valid = true;
  return RVSP_SUCCESS;
}

rvsp_status_t run_numeric(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                          const int* c_ci, float* c_val, hipStream_t s) {
  const RowStats st = row_stats(d);
  for (int b = 1; b < d->numeric.table.rule.nbins; ++b)
    if (d->numeric.table.spec[b].kind == KIND_BLOCK && bin_rows(d->numeric, b) > 0 &&
        !((d->stream_bins >> b) & 1u)) {
      RVSP_TRY(ensure_split(d, B, s));
      break;
    }
  for (int b = 1; b < d->numeric.table.rule.nbins; ++b) {
    const int n = bin_rows(d->numeric, b);
    if (n == 0 || ((d->stream_bins >> b) & 1u)) continue;
    const BinSpec& spec = d->numeric.table.spec[b];
    const int* rows = (const int*)d->numeric.rows.p + d->numeric.offsets[b];
    if (d->defer_block && spec.kind == KIND_BLOCK && spec.variant == 0) {
      DeferredBin db;
      db.rows = rows;
      db.nrows = n;
      db.max_row = d->numeric.table.rule.bound[b];
      db.spec = spec;
      db.a_rp = A.rp;
      db.c_rp = (const int*)d->c_rp.p;
      db.c_ci = c_ci;
      db.c_val = c_val;
      d->deferred.push_back(db);
      continue;
    }
    if (spec.kind == KIND_MAP) {
      RVSP_TRY(ensure_map_built(d, A, B, c_ci, s));
      MapView mv = map_view(d);
      mv.row_base = numeric_row_base(d, rows);
      RVSP_HIP_TRY(launch_numeric_map_bin(spec, mv, rows, n, A, B, (const int*)d->c_rp.p,
                                          c_val, s));
    } else {
      RVSP_HIP_TRY(launch_numeric_bin(spec, rows, n, A, B, st, (const int*)d->c_rp.p, c_ci,
                                      c_val, d->gs_dense, d->numeric.table.rule.bound[b],
                                      d->split, s));
    }
  }
  return RVSP_SUCCESS;
}

MapView map_view(const rvsp_spgemm_descr* d) {
  MapView mv;
  mv.map = d->map.p;
  mv.poff = (const long long*)d->poff.p;
  mv.width = d->map_width;
  if (d->brow_ok) {
    mv.brow = (const int*)d->brow.p;
    mv.blen = (const unsigned short*)((const char*)d->brow.p + (size_t)d->a_nnz * 4);
  }
  return mv;
}

int numeric_row_base(const rvsp_spgemm_descr* d, const int* rows) {
  const PhasePlan& p = d->numeric;
  for (size_t b = 1; b < p.row_base.size() && b < p.offsets.size(); ++b)
    if ((const int*)p.rows.p + p.offsets[b] == rows && bin_rows(p, (int)b) > 0)
      return p.row_base[b];
  return -1;
}

rvsp_status_t plan_map(rvsp_spgemm_descr* d, hipStream_t s) {
  PhasePlan& p = d->numeric;
  unsigned mask = 0;
  int max_bound = 0;
  for (int b = 2; b < p.table.rule.nbins; ++b) {
    if (p.table.spec[b].kind != KIND_MAP || bin_rows(p, b) == 0) continue;
    mask |= 1u << b;
    max_bound = max_bound > p.table.rule.bound[b] ? max_bound : p.table.rule.bound[b];
  }
  d->map_valid = false;
  d->map_bins = 0;
  d->map_width = 0;
  d->brow_ok = false;
  if (mask == 0) return RVSP_SUCCESS;
  const int M = p.offsets.empty() ? 0 : p.offsets.back();  // rows partitioned
  const int width = max_bound <= 256 ? 1 : 2;  // positions are < the row's nnz
  RVSP_TRY(ensure(d->poff, ((size_t)M + 1) * 8));
  RVSP_TRY(ensure(d->map_weights, (size_t)M * 8));
  long long total = 0;
  RVSP_HIP_TRY(launch_map_offsets(M, (const unsigned char*)d->bin_of.p, mask, row_stats(d),
                                  (long long*)d->map_weights.p, (long long*)d->poff.p,
                                  (long long*)d->scan_tiles.p, (long long*)d->total.p, s));
  RVSP_HIP_TRY(hipMemcpyAsync(&total, d->total.p, 8, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  const size_t bytes = (size_t)total * width;
  const size_t free_b = device_free_bytes();
  const size_t budget = std::min<size_t>((size_t)8 << 30, free_b / 4);
  // 64 bytes of slack: the map kernels stage a segment's entries with
  // aligned 16-byte loads that may read up to 15 bytes past its end.
  if (bytes > budget || ensure(d->map, bytes + 64) != RVSP_SUCCESS) {
    // Over budget: these bins keep searching C's columns instead.
    for (int b = 2; b < p.table.rule.nbins; ++b)
      if ((mask >> b) & 1u) {
        BinChoice c;
        c.kind = KIND_SEARCH;
        apply_choice(p.table, PHASE_NUMERIC, b, c, p.table.rule.bound[b]);
      }
    return RVSP_SUCCESS;
  }
  d->map_bins = mask;
  d->map_width = width;
  // The B-row cache: a B row a map row reaches is no longer than the row's
  // product count (at most max_bound), so 16-bit lengths hold it.
  const size_t brow_bytes = (size_t)d->a_nnz * 6 + 16;
  if (max_bound <= 65535 && d->a_nnz > 0 && bytes + brow_bytes <= budget &&
      ensure(d->brow, brow_bytes) == RVSP_SUCCESS)
    d->brow_ok = true;  // filled with the map entries
  return RVSP_SUCCESS;
}

rvsp_status_t ensure_map_built(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                               const int* c_ci, hipStream_t s) {
  if (d->map_valid) return RVSP_SUCCESS;
  if (d->map_width == 0) return RVSP_ERROR_DEVICE;  // a MAP bin without a planned map
  PhasePlan& p = d->numeric;
  for (int b = 2; b < p.table.rule.nbins; ++b) {
    if (!((d->map_bins >> b) & 1u)) continue;
    RVSP_HIP_TRY(launch_build_map(d->map_width, (const int*)p.rows.p + p.offsets[b],
                                  bin_rows(p, b), A, B, (const int*)d->c_rp.p, c_ci,
                                  (const long long*)d->poff.p, d->map.p, s));
  }
  if (d->brow_ok) {
    const MapView mv = map_view(d);
    RVSP_HIP_TRY(launch_brow_cache(A, B, const_cast<int*>(mv.brow),
                                   const_cast<unsigned short*>(mv.blen), s));
  }
  d->map_valid = true;
  return RVSP_SUCCESS;
}

rvsp_status_t release_position_map(rvsp_spgemm_descr* d) {
  PhaseTable& t = d->numeric.table;
  for (int b = 2; b < t.rule.nbins; ++b)
    if (t.spec[b].kind == KIND_MAP) {
      BinChoice c;
      c.kind = KIND_SEARCH;
      apply_choice(t, PHASE_NUMERIC, b, c, t.rule.bound[b]);
    }
  release(d->map);
  release(d->poff);
  release(d->map_weights);
  release(d->brow);
  d->map_valid = false;
  d->map_bins = 0;
  d->map_width = 0;
  d->map_promoted = false;
  d->brow_ok = false;
  return RVSP_SUCCESS;
}

constexpr int kMapFromTiledMax = 2048;  // longest C row a tiled bin may move to the map with

rvsp_status_t promote_to_map(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                             const int* c_ci, hipStream_t s) {
  d->map_promoted = true;
  PhaseTable& t = d->numeric.table;
  bool any = false;
  for (int b = 2; b < t.rule.nbins; ++b) {
    if (bin_rows(d->numeric, b) == 0 || ((d->stream_bins >> b) & 1u)) continue;
    // Searching bins keep their shape; tiled bins of rows short enough for
    // a workgroup-sized accumulator take the widest map shape.
    const bool block = t.spec[b].kind == KIND_BLOCK && t.rule.bound[b] <= kMapFromTiledMax;
    if (t.spec[b].kind != KIND_SEARCH && !block) continue;
    BinChoice c;
    c.kind = KIND_MAP;
    c.group = block ? 64 : t.spec[b].group;
    c.cap = block ? t.rule.bound[b] : t.spec[b].cap;
    c.variant = 2;
    PhaseTable trial = t;  // a refused choice would turn the bin GLOBAL
    if (apply_choice(trial, PHASE_NUMERIC, b, c, t.rule.bound[b])) {
      t.spec[b] = trial.spec[b];
      any = true;
    }
  }
  if (!any) return RVSP_SUCCESS;
  if (d->algo == RVSP_SPGEMM_ALGO_HIP_LEARNED) {
    // The network picks each moved bin's map shape (lanes, variant) with its
    // reuse input set; apply_choice still vetoes what cannot hold the rows.
    std::vector<BinFeatures> f;
    RVSP_TRY(bin_features(d, d->numeric, PHASE_NUMERIC, A.rp, (const int*)d->c_rp.p, s, &f));
    RVSP_TRY(select_learned(d, t, PHASE_NUMERIC, f, s, true));
  }
  RVSP_TRY(plan_map(d, s));  // bins over the memory budget go back to SEARCH
  if (d->map_width) RVSP_TRY(ensure_map_built(d, A, B, c_ci, s));
  return RVSP_SUCCESS;
}

}  // namespace rvsp

extern "C" {

// Reference: src/core/spgemm_descr.c:66-89.
rvsp_status_t rvsp_spgemm_descr_create(rvsp_spgemm_descr_t* descr) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  *descr = NULL;
  rvsp_spgemm_descr* d = new (std::nothrow) rvsp_spgemm_descr();
  if (d == NULL) return RVSP_ERROR_ALLOCATION_FAILED;
  d->algo = RVSP_SPGEMM_ALGO_DEFAULT;
  *descr = d;
  return RVSP_SUCCESS;
}

// Reference: src/core/spgemm_descr.c:90-100 (NULL accepted).
void rvsp_spgemm_descr_destroy(rvsp_spgemm_descr_t descr) {
  if (descr == NULL) return;
  rvsp_spgemm_descr* d = descr;
  if (d->own_stream) (void)hipStreamSynchronize(d->own_stream);
  if (d->user_stream) (void)hipStreamSynchronize(d->user_stream);
  DevBuf* bufs[] = {&d->a_rp, &d->a_ci, &d->a_val, &d->b_rp, &d->b_ci, &d->b_val,
                    &d->ops, &d->lo, &d->hi, &d->maxb, &d->long_list, &d->n_long,
                    &d->cnt, &d->c_rp, &d->scan_tiles, &d->total, &d->bin_of,
                    &d->tile_hist, &d->tile_base, &d->bin_offsets, &d->count.rows,
                    &d->fill.rows, &d->numeric.rows, &d->c_ci_dev, &d->c_val_dev,
                    &d->gbits, &d->gdense, &d->learn_feat, &d->bin_stats,
                    &d->map, &d->poff, &d->map_weights, &d->brow, &d->keep, &d->koff,
                    &d->keep_w, &d->split_cnt, &d->split_off, &d->split_sp,
                    &d->check_key};
  // The streams are drained above: the buffers go back to the cache idle.
  for (DevBuf* b : bufs) release_idle(*b);
  destroy_graph(d);
  destroy_batch(d);
  if (d->own_stream) (void)hipStreamDestroy(d->own_stream);
  free(d->c_row_ptr);
  delete d;
}

// Reference: src/core/spgemm_descr.c:102-125.
rvsp_status_t rvsp_spgemm_set_algo(rvsp_spgemm_descr_t descr,
                                   rvsp_spgemm_algo_t algo) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  if (descr->estimated || descr->batch) return RVSP_ERROR_INVALID_ARGUMENT;
  if (!algo_available(algo)) return RVSP_ERROR_UNSUPPORTED_BACKEND;
  descr->algo = algo;
  return RVSP_SUCCESS;
}

// Reference: src/core/spgemm_descr.c:129-135.
void rvsp_spgemm_invalidate_structure(rvsp_spgemm_descr_t descr) {
  if (descr != NULL) {
    descr->filled_into = NULL;
    descr->dev_cols_valid = false;
    descr->map_valid = false;
    descr->split_valid = false;
    invalidate_batch(descr);
  }
}

// Reference: src/core/spgemm_descr.c:162-255 (check order, re-estimation
// discards the previous structure, int32 guard -> ALLOCATION_FAILED).
rvsp_status_t rvsp_spgemm_work_estimation(rvsp_spgemm_descr_t descr,
                                          const rvsp_csr_matrix_t* A,
                                          const rvsp_csr_matrix_t* B,
                                          size_t* workspace_bytes,
                                          int32_t* c_nnz_out) {
  if (descr == NULL || workspace_bytes == NULL || c_nnz_out == NULL)
    return RVSP_ERROR_NULL_POINTER;
  RVSP_TRY(check_operand(A));
  RVSP_TRY(check_operand(B));
  if (A->cols != B->rows) return RVSP_ERROR_INVALID_ARGUMENT;
  rvsp_spgemm_descr* d = descr;
  drop_structure(d);

  RVSP_TRY(device_ready());
  hipStream_t s;
  RVSP_TRY(descr_stream(d, &s));
  DevCsr dA, dB;
  RVSP_TRY(operand_views(d, A, B, true, false, &dA, &dB, s));
  if (d->check_csr) {
    rvsp_csr_status_t va = RVSP_CSR_OK, vb = RVSP_CSR_OK;
    int bad = -1;
    RVSP_TRY(csr_check_device(A->rows, A->cols, A->nnz, dA.rp, dA.ci, d->check_key, s, &va,
                              &bad));
    if (va == RVSP_CSR_OK && dB.rp != dA.rp)
      RVSP_TRY(csr_check_device(B->rows, B->cols, B->nnz, dB.rp, dB.ci, d->check_key, s, &vb,
                                &bad));
    if (va != RVSP_CSR_OK || vb != RVSP_CSR_OK) return RVSP_ERROR_INVALID_CSR;
  }
  long long total = 0;
  rvsp_status_t st = estimate(d, dA, dB, &total, s);
  if (st != RVSP_SUCCESS) {
    (void)hipStreamSynchronize(s);
    return st;
  }
  if (d->memory == RVSP_MEMORY_HOST) {
    d->c_row_ptr = (int32_t*)malloc(((size_t)A->rows + 1) * sizeof(int32_t));
    if (d->c_row_ptr == NULL) return RVSP_ERROR_ALLOCATION_FAILED;
    RVSP_HIP_TRY(hipMemcpyAsync(d->c_row_ptr, d->c_rp.p, ((size_t)A->rows + 1) * 4,
                                hipMemcpyDeviceToHost, s));
    RVSP_HIP_TRY(hipStreamSynchronize(s));
  }
  d->c_nnz = (int32_t)total;
  d->a_rows = A->rows; d->a_cols = A->cols; d->a_nnz = A->nnz;
  d->b_rows = B->rows; d->b_cols = B->cols; d->b_nnz = B->nnz;
  // compute() keeps its scratch on the device; the host workspace argument
  // is only checked for NULL, as in the reference.
  d->workspace_bytes = 64;
  d->estimated = 1;
  *workspace_bytes = d->workspace_bytes;
  *c_nnz_out = d->c_nnz;
  return RVSP_SUCCESS;
}

// Reference: src/core/spgemm_descr.c:257-355 (check order at :263-318, then
// row_ptr copy, fill unless this C.col_idx already holds the structure,
// numeric).
rvsp_status_t rvsp_spgemm_compute(rvsp_spgemm_descr_t descr,
                                  const rvsp_csr_matrix_t* A,
                                  const rvsp_csr_matrix_t* B,
                                  rvsp_csr_matrix_t* C, void* workspace) {
  if (descr == NULL || C == NULL || workspace == NULL) return RVSP_ERROR_NULL_POINTER;
  rvsp_spgemm_descr* d = descr;
  if (!d->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  RVSP_TRY(check_operand(A));
  RVSP_TRY(check_operand(B));
  if (A->rows != d->a_rows || A->cols != d->a_cols || A->nnz != d->a_nnz ||
      B->rows != d->b_rows || B->cols != d->b_cols || B->nnz != d->b_nnz)
    return RVSP_ERROR_INVALID_ARGUMENT;
  if (C->row_ptr == NULL || C->col_idx == NULL || C->values == NULL)
    return RVSP_ERROR_NULL_POINTER;
  if (C->dtype != RVSP_DTYPE_FP32) return RVSP_ERROR_UNSUPPORTED_DTYPE;
  if (C->rows != d->a_rows || C->cols != d->b_cols) return RVSP_ERROR_INVALID_ARGUMENT;
  if (C->nnz < d->c_nnz) return RVSP_ERROR_INVALID_ARGUMENT;
  if (!algo_available(d->algo)) return RVSP_ERROR_UNSUPPORTED_BACKEND;

  hipStream_t s;
  RVSP_TRY(descr_stream(d, &s));
  DevCsr dA, dB;
  RVSP_TRY(operand_views(d, A, B, false, true, &dA, &dB, s));
  const size_t nnzc = (size_t)(d->c_nnz > 0 ? d->c_nnz : 1);

  if (d->memory == RVSP_MEMORY_DEVICE) {
    if (d->rowptr_into != C->row_ptr) {
      RVSP_HIP_TRY(hipMemcpyAsync(C->row_ptr, d->c_rp.p, ((size_t)d->a_rows + 1) * 4,
                                  hipMemcpyDeviceToDevice, s));
      d->rowptr_into = C->row_ptr;
    }
    if (d->filled_into != C->col_idx) {
      RVSP_TRY(run_fill(d, dA, dB, C->col_idx, s));
      d->filled_into = C->col_idx;
    } else if (d->position_map && !d->map_promoted && d->algo != RVSP_SPGEMM_ALGO_HIP_TUNED) {
      // A numeric-only compute: the structure is being reused, so its
      // searching bins switch to the position map from here on (tuned plans
      // keep the shapes their measurements chose).
      RVSP_TRY(promote_to_map(d, dA, dB, C->col_idx, s));
    }
    RVSP_TRY(run_numeric(d, dA, dB, C->col_idx, (float*)C->values, s));
    if (!d->async) RVSP_HIP_TRY(hipStreamSynchronize(s));
  } else {
    RVSP_TRY(ensure(d->c_ci_dev, nnzc * 4));
    RVSP_TRY(ensure(d->c_val_dev, nnzc * 4));
    if (!d->dev_cols_valid) {
      RVSP_TRY(run_fill(d, dA, dB, (int*)d->c_ci_dev.p, s));
      d->dev_cols_valid = true;
      d->filled_into = NULL;
    } else if (d->position_map && !d->map_promoted && d->algo != RVSP_SPGEMM_ALGO_HIP_TUNED) {
      RVSP_TRY(promote_to_map(d, dA, dB, (const int*)d->c_ci_dev.p, s));
    }
    RVSP_TRY(run_numeric(d, dA, dB, (const int*)d->c_ci_dev.p, (float*)d->c_val_dev.p, s));
    if (d->filled_into != C->col_idx && d->c_nnz > 0)
      RVSP_HIP_TRY(hipMemcpyAsync(C->col_idx, d->c_ci_dev.p, (size_t)d->c_nnz * 4,
                                  hipMemcpyDeviceToHost, s));
    if (d->c_nnz > 0)
      RVSP_HIP_TRY(hipMemcpyAsync(C->values, d->c_val_dev.p, (size_t)d->c_nnz * 4,
                                  hipMemcpyDeviceToHost, s));
    RVSP_HIP_TRY(hipStreamSynchronize(s));
    memcpy(C->row_ptr, d->c_row_ptr, ((size_t)d->a_rows + 1) * sizeof(int32_t));
    d->filled_into = C->col_idx;
  }
  C->nnz = d->c_nnz;
  C->format = RVSP_FORMAT_CSR;
  return RVSP_SUCCESS;
}

// Reference: src/core/spgemm_descr.c:357-380.
rvsp_status_t rvsp_spgemm_get_op_counts(rvsp_spgemm_descr_t descr,
                                        int64_t* op_counts_out, int32_t n) {
  if (descr == NULL || op_counts_out == NULL) return RVSP_ERROR_NULL_POINTER;
  if (!descr->estimated || descr->ops.p == NULL) return RVSP_ERROR_INVALID_ARGUMENT;
  if (n < descr->a_rows) return RVSP_ERROR_INVALID_ARGUMENT;
  hipStream_t s;
  RVSP_TRY(descr_stream(descr, &s));
  RVSP_HIP_TRY(hipMemcpyAsync(op_counts_out, descr->ops.p, (size_t)descr->a_rows * 8,
                              hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  return RVSP_SUCCESS;
}

// --- one-shot ------------------------------------------------------------

// Reference: src/core/spgemm.c:26-133.
static rvsp_status_t oneshot(const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B,
                             rvsp_csr_matrix_t* C, rvsp_spgemm_algo_t algo) {
  if (A->dtype != RVSP_DTYPE_FP32 || B->dtype != RVSP_DTYPE_FP32)
    return RVSP_ERROR_UNSUPPORTED_DTYPE;
  if (A->cols != B->rows) return RVSP_ERROR_INVALID_ARGUMENT;
  // The canonical check runs on the device once the estimation has staged
  // the operands, before any kernel reads them. Operands the estimation
  // would refuse first, or a process without a usable device, are checked
  // here on the host, so the reference's order (INVALID_CSR before the
  // descriptor's own argument errors and before device errors) holds.
  const bool on_device = check_operand(A) == RVSP_SUCCESS &&
                         check_operand(B) == RVSP_SUCCESS && device_ready() == RVSP_SUCCESS;
  if (!on_device &&
      (rvsp_csr_check(A->rows, A->cols, A->row_ptr, A->col_idx, NULL) != RVSP_CSR_OK ||
       rvsp_csr_check(B->rows, B->cols, B->row_ptr, B->col_idx, NULL) != RVSP_CSR_OK))
    return RVSP_ERROR_INVALID_CSR;
  rvsp_spgemm_descr_t d = NULL;
  rvsp_status_t st = rvsp_spgemm_descr_create(&d);
  if (st != RVSP_SUCCESS) return st;
  d->check_csr = on_device;
  st = rvsp_spgemm_set_algo(d, algo);
  if (st != RVSP_SUCCESS) {
    rvsp_spgemm_descr_destroy(d);
    return st;
  }
  size_t ws_bytes = 0;
  int32_t c_nnz = 0;
  st = rvsp_spgemm_work_estimation(d, A, B, &ws_bytes, &c_nnz);
  if (st != RVSP_SUCCESS) {
    rvsp_spgemm_descr_destroy(d);
    return st;
  }
  const size_t alloc_nnz = c_nnz > 0 ? (size_t)c_nnz : 1;
  int32_t* rp = (int32_t*)malloc(((size_t)A->rows + 1) * sizeof(int32_t));
  int32_t* ci = (int32_t*)malloc(alloc_nnz * sizeof(int32_t));
  float* val = (float*)malloc(alloc_nnz * sizeof(float));
  void* ws = malloc(ws_bytes);
  if (!rp || !ci || !val || !ws) {
    free(rp); free(ci); free(val); free(ws);
    rvsp_spgemm_descr_destroy(d);
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
  C->rows = A->rows;
  C->cols = B->cols;
  C->nnz = c_nnz;
  C->row_ptr = rp;
  C->col_idx = ci;
  C->values = val;
  C->dtype = RVSP_DTYPE_FP32;
  C->format = RVSP_FORMAT_CSR;
  C->owns_data = 1;
  st = rvsp_spgemm_compute(d, A, B, C, ws);
  free(ws);
  rvsp_spgemm_descr_destroy(d);
  if (st != RVSP_SUCCESS) {
    free(rp); free(ci); free(val);
    C->row_ptr = NULL;
    C->col_idx = NULL;
    C->values = NULL;
    C->nnz = 0;
    C->owns_data = 0;
  }
  return st;
}

// Reference: src/core/spgemm.c:135-211. The FP32 scalar backends and the HIP
// backend run the HIP descriptor path; the RVV backend selects the RVV
// strategy, which this build reports as unsupported after the same checks
// the reference makes first. FP64 x FP64 -> FP64 and INT8 x INT8 -> INT32
// (the reference's legacy kernels) run on the GPU through legacy.cpp
// (oneshot_legacy), with each legacy wrapper's status codes.
rvsp_status_t rvsp_spgemm_csr(const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B,
                              rvsp_csr_matrix_t* C, const rvsp_spgemm_options_t* options) {
  if (!A || !B || !C) return RVSP_ERROR_NULL_POINTER;
  rvsp_backend_t backend = RVSP_BACKEND_SCALAR;
  rvsp_dtype_t in = RVSP_DTYPE_FP32, out = RVSP_DTYPE_FP32;
  if (options) {
    backend = options->backend;
    in = options->input_dtype;
    out = options->output_dtype;
  }
  if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_SCALAR_UNROLL4 ||
       backend == RVSP_BACKEND_HIP) &&
      in == RVSP_DTYPE_FP32 && out == RVSP_DTYPE_FP32)
    return oneshot(A, B, C, RVSP_SPGEMM_ALGO_DEFAULT);
  if (backend == RVSP_BACKEND_RVV_INTRINSICS && A->dtype == RVSP_DTYPE_FP32 &&
      B->dtype == RVSP_DTYPE_FP32)
    return oneshot(A, B, C, RVSP_SPGEMM_ALGO_RVV);
  // The legacy dtypes (spgemm.c:182-196), on the GPU (legacy.cpp). The
  // reference's vector variants of them need a RISC-V vector build.
  if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_HIP) &&
      in == RVSP_DTYPE_INT8 && out == RVSP_DTYPE_INT32)
    return oneshot_legacy(A, B, C, true);
  if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_HIP) &&
      in == RVSP_DTYPE_FP64 && out == RVSP_DTYPE_FP64)
    return oneshot_legacy(A, B, C, false);
  return RVSP_ERROR_UNSUPPORTED_BACKEND;
}

// Raw one-shot kernel with the reference's RVSP_SPGEMM_PARAMS
// (gustavson_core_f32.inc:66-194): argument checks, then symbolic count,
// C allocation, fill and numeric on the GPU. rvsp_spgemm_scalar_f32 (the
// reference's scalar instance, rvsp_spgemm.h:79) is served by it, as the
// scalar backend is.
rvsp_status_t rvsp_spgemm_hip_f32(int32_t a_rows, int32_t a_cols, int32_t b_cols,
                                  const int32_t* a_row_ptr, const int32_t* a_col_idx,
                                  const float* a_values, const int32_t* b_row_ptr,
                                  const int32_t* b_col_idx, const float* b_values,
                                  void* workspace, size_t workspace_bytes,
                                  int32_t** c_row_ptr_out, int32_t** c_col_idx_out,
                                  float** c_values_out, int32_t* c_nnz_out,
                                  int64_t* op_counts_out) {
  if (a_row_ptr == NULL || a_col_idx == NULL || a_values == NULL ||
      b_row_ptr == NULL || b_col_idx == NULL || b_values == NULL ||
      c_row_ptr_out == NULL || c_col_idx_out == NULL || c_values_out == NULL ||
      c_nnz_out == NULL || workspace == NULL)
    return RVSP_ERROR_NULL_POINTER;
  if (a_rows <= 0 || a_cols <= 0 || b_cols <= 0) return RVSP_ERROR_INVALID_ARGUMENT;
  size_t need = 0;
  rvsp_spgemm_buffer_size(b_cols, &need);
  if (workspace_bytes < need) return RVSP_ERROR_INVALID_ARGUMENT;
  *c_row_ptr_out = NULL;
  *c_col_idx_out = NULL;
  *c_values_out = NULL;
  *c_nnz_out = 0;

  rvsp_csr_matrix_t A = {a_rows, a_cols, a_row_ptr[a_rows], (int32_t*)a_row_ptr,
                         (int32_t*)a_col_idx, (void*)a_values, RVSP_DTYPE_FP32,
                         RVSP_FORMAT_CSR, 0};
  rvsp_csr_matrix_t B = {a_cols, b_cols, b_row_ptr[a_cols], (int32_t*)b_row_ptr,
                         (int32_t*)b_col_idx, (void*)b_values, RVSP_DTYPE_FP32,
                         RVSP_FORMAT_CSR, 0};
  rvsp_spgemm_descr_t d = NULL;
  rvsp_status_t st = rvsp_spgemm_descr_create(&d);
  if (st != RVSP_SUCCESS) return st;
  size_t ws = 0;
  int32_t c_nnz = 0;
  st = rvsp_spgemm_work_estimation(d, &A, &B, &ws, &c_nnz);
  if (st == RVSP_SUCCESS && op_counts_out != NULL)
    st = rvsp_spgemm_get_op_counts(d, op_counts_out, a_rows);
  if (st != RVSP_SUCCESS) {
    rvsp_spgemm_descr_destroy(d);
    return st;
  }
  const size_t alloc_nnz = c_nnz > 0 ? (size_t)c_nnz : 1;
  int32_t* rp = (int32_t*)malloc(((size_t)a_rows + 1) * sizeof(int32_t));
  int32_t* ci = (int32_t*)malloc(alloc_nnz * sizeof(int32_t));
  float* val = (float*)malloc(alloc_nnz * sizeof(float));
  if (!rp || !ci || !val) {
    free(rp); free(ci); free(val);
    rvsp_spgemm_descr_destroy(d);
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
  rvsp_csr_matrix_t C = {a_rows, b_cols, c_nnz, rp, ci, val, RVSP_DTYPE_FP32,
                         RVSP_FORMAT_CSR, 0};
  st = rvsp_spgemm_compute(d, &A, &B, &C, workspace);
  rvsp_spgemm_descr_destroy(d);
  if (st != RVSP_SUCCESS) {
    free(rp); free(ci); free(val);
    return st;
  }
  *c_row_ptr_out = rp;
  *c_col_idx_out = ci;
  *c_values_out = val;
  *c_nnz_out = c_nnz;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_scalar_f32(int32_t a_rows, int32_t a_cols, int32_t b_cols,
                                     const int32_t* a_row_ptr, const int32_t* a_col_idx,
                                     const float* a_values, const int32_t* b_row_ptr,
                                     const int32_t* b_col_idx, const float* b_values,
                                     void* workspace, size_t workspace_bytes,
                                     int32_t** c_row_ptr_out, int32_t** c_col_idx_out,
                                     float** c_values_out, int32_t* c_nnz_out,
                                     int64_t* op_counts_out) {
  return rvsp_spgemm_hip_f32(a_rows, a_cols, b_cols, a_row_ptr, a_col_idx, a_values, b_row_ptr,
                             b_col_idx, b_values, workspace, workspace_bytes, c_row_ptr_out,
                             c_col_idx_out, c_values_out, c_nnz_out, op_counts_out);
}

// --- HIP extensions --------------------------------------------------------

rvsp_status_t rvsp_spgemm_set_memory(rvsp_spgemm_descr_t descr, rvsp_memory_t memory) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  if (descr->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  if (memory != RVSP_MEMORY_HOST && memory != RVSP_MEMORY_DEVICE)
    return RVSP_ERROR_INVALID_ARGUMENT;
  descr->memory = memory;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_set_stream(rvsp_spgemm_descr_t descr, void* hip_stream) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  descr->user_stream = (hipStream_t)hip_stream;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_get_stream(rvsp_spgemm_descr_t descr, void** hip_stream_out) {
  if (descr == NULL || hip_stream_out == NULL) return RVSP_ERROR_NULL_POINTER;
  hipStream_t s;
  RVSP_TRY(descr_stream(descr, &s));
  *hip_stream_out = (void*)s;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_set_async(rvsp_spgemm_descr_t descr, int async) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  descr->async = async != 0;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_set_position_map(rvsp_spgemm_descr_t descr, int on) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  descr->position_map = on != 0;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_get_plan(rvsp_spgemm_descr_t descr, rvsp_bin_info_t* out,
                                   int32_t max, int32_t* n_out) {
  if (descr == NULL || n_out == NULL) return RVSP_ERROR_NULL_POINTER;
  if (!descr->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  const PhasePlan* plans[] = {&descr->count, &descr->fill, &descr->numeric};
  int32_t n = 0;
  for (int ph = 0; ph < 3; ++ph) {
    const PhasePlan& p = *plans[ph];
    for (int b = 0; b < p.table.rule.nbins; ++b, ++n) {
      if (out == NULL || n >= max) continue;
      rvsp_bin_info_t& o = out[n];
      o.phase = ph;
      o.bin = b;
      o.kind = p.table.spec[b].kind;
      o.group = p.table.spec[b].group;
      o.cap = p.table.spec[b].cap;
      o.rows = (int)p.offsets.size() > b + 1 ? p.offsets[b + 1] - p.offsets[b] : 0;
      o.lds_bytes = p.table.spec[b].lds_bytes;
      o.bound = p.table.rule.bound[b];
    }
  }
  *n_out = n;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_set_bin_choices(rvsp_spgemm_descr_t descr,
                                          const rvsp_bin_choice_t* choices,
                                          int32_t n) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  if (n < 0 || (n > 0 && choices == NULL)) return RVSP_ERROR_INVALID_ARGUMENT;
  if (descr->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  descr->choices.assign(choices, choices + n);
  return RVSP_SUCCESS;
}

// Autotuner: compute-equivalent call that additionally times every shape of
// every non-empty bin and keeps the fastest (see selector.cpp).
rvsp_status_t rvsp_spgemm_autotune(rvsp_spgemm_descr_t descr, const rvsp_csr_matrix_t* A,
                                   const rvsp_csr_matrix_t* B, rvsp_csr_matrix_t* C,
                                   void* workspace, int32_t reps) {
  RVSP_TRY(rvsp_spgemm_compute(descr, A, B, C, workspace));
  rvsp_spgemm_descr* d = descr;
  hipStream_t s;
  RVSP_TRY(descr_stream(d, &s));
  DevCsr dA, dB;
  RVSP_TRY(operand_views(d, A, B, false, false, &dA, &dB, s));
  int* c_ci = d->memory == RVSP_MEMORY_DEVICE ? C->col_idx : (int*)d->c_ci_dev.p;
  float* c_val = d->memory == RVSP_MEMORY_DEVICE ? (float*)C->values : (float*)d->c_val_dev.p;
  RVSP_TRY(autotune_plan(d, dA, dB, c_ci, c_val, reps, s));
  if (d->memory == RVSP_MEMORY_HOST && d->c_nnz > 0) {
    RVSP_HIP_TRY(hipMemcpyAsync(C->values, d->c_val_dev.p, (size_t)d->c_nnz * 4,
                                hipMemcpyDeviceToHost, s));
    RVSP_HIP_TRY(hipStreamSynchronize(s));
  }
  return RVSP_SUCCESS;
}

int32_t rvsp_tuning_table_save(const char* path) { return tuning_table_save(path); }
int32_t rvsp_tuning_table_load(const char* path) { return tuning_table_load(path); }
void rvsp_tuning_table_clear(void) { tuning_table_clear(); }
int32_t rvsp_tuning_table_size(void) { return tuning_table_size(); }

rvsp_status_t rvsp_hip_malloc(void** ptr, size_t bytes) {
  if (ptr == NULL) return RVSP_ERROR_NULL_POINTER;
  RVSP_TRY(device_ready());
  RVSP_HIP_TRY(hipMalloc(ptr, bytes ? bytes : 4));
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_free(void* ptr) {
  if (ptr == NULL) return RVSP_SUCCESS;
  RVSP_HIP_TRY(hipFree(ptr));
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  if (bytes == 0) return RVSP_SUCCESS;
  if (dst == NULL || src == NULL) return RVSP_ERROR_NULL_POINTER;
  hipMemcpyKind k;
  switch (kind) {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: return RVSP_ERROR_INVALID_ARGUMENT;
  }
  RVSP_TRY(device_ready());
  RVSP_HIP_TRY(hipMemcpy(dst, src, bytes, k));
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_synchronize(void) {
  RVSP_TRY(device_ready());
  RVSP_HIP_TRY(hipDeviceSynchronize());
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_csr_check_device(int32_t rows, int32_t cols, int32_t nnz,
                                    const int32_t* row_ptr, const int32_t* col_idx,
                                    rvsp_csr_status_t* status_out, int32_t* bad_row_out) {
  if (status_out == NULL) return RVSP_ERROR_NULL_POINTER;
  RVSP_TRY(device_ready());
  DevBuf key;
  int bad = -1;
  const rvsp_status_t st =
      csr_check_device(rows, cols, nnz, row_ptr, col_idx, key, nullptr, status_out, &bad);
  release(key);
  if (bad_row_out) *bad_row_out = bad;
  return st;
}

int32_t rvsp_hip_device_count(void) {
  return device_ready() == RVSP_SUCCESS ? 1 : 0;
}

}  // extern "C"
