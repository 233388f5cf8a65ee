// batch.cpp — batched SpGEMM: C[p] = A[p] * B[p] for many independent pairs
// in one set of launches per phase, repeated numeric-only computes replayed
// from a captured HIP graph. No counterpart in the reference; every pair's
// result is exactly what rvsp_spgemm_compute gives for that pair alone.
//
// The batch is packed into one block-diagonal product. The distinct B
// operands are stacked (rows concatenated, column indices kept local), each
// A is stacked with its column indices moved to the rows of its B in the
// stack, and the stacked A times the stacked B is the stacked C: a row of
// pair p reaches only rows of B[p], and its columns stay pair-local. Every
// kernel of the single-product path then serves all pairs at once, and row
// order, accumulation order and output bits are those of each pair alone.
//
// The output row pointers are int32, so the stack is cut into groups of
// consecutive pairs whose product count (an upper bound of their nnz) fits;
// each group is one inner descriptor over a row range of the stacked A.
// When the caller's C arrays of a group sit back to back in memory, the group
// writes them in place; otherwise it writes its own buffer and one segmented
// copy moves every pair's part out.
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <map>
#include <thread>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "descr_impl.hpp"

namespace rvsp {

struct BatchGroup {
  int p0 = 0, p1 = 0;         // pairs [p0, p1)
  long long r0 = 0, r1 = 0;   // rows of the stacked A
  rvsp_spgemm_descr* inner = nullptr;
  int32_t c_nnz = 0;
  std::vector<long long> off;  // nnz offset of pair p0 + i within the group
  DevBuf c_rp, c_ci, c_val;
  bool in_place = false;       // C arrays of the last compute were back to back
};

// A segment table uploaded to the device, with its host copy kept alive
// until the next rebuild (asynchronous uploads read it).
struct SegTable {
  std::vector<SegCopy> segs;
  std::vector<SegChunk> chunks;
  DevBuf d_segs, d_chunks;
  void clear() {
    segs.clear();
    chunks.clear();
  }
  void add(const int* src, int* dst, long long len, int add, int rebase) {
    if (len <= 0) return;
    SegCopy s;
    s.src = src;
    s.dst = dst;
    s.len = len;
    s.add = add;
    s.rebase = rebase;
    const int id = (int)segs.size();
    segs.push_back(s);
    for (long long b = 0; b < len; b += kSegChunk) {
      SegChunk c;
      c.seg = id;
      c.pad = 0;
      c.begin = b;
      chunks.push_back(c);
    }
  }
  rvsp_status_t upload(hipStream_t s) {
    if (segs.empty()) return RVSP_SUCCESS;
    RVSP_TRY(ensure(d_segs, segs.size() * sizeof(SegCopy)));
    RVSP_TRY(ensure(d_chunks, chunks.size() * sizeof(SegChunk)));
    RVSP_HIP_TRY(hipMemcpyAsync(d_segs.p, segs.data(), segs.size() * sizeof(SegCopy),
                                hipMemcpyHostToDevice, s));
    RVSP_HIP_TRY(hipMemcpyAsync(d_chunks.p, chunks.data(), chunks.size() * sizeof(SegChunk),
                                hipMemcpyHostToDevice, s));
    return RVSP_SUCCESS;
  }
  hipError_t launch(hipStream_t s) const {
    return launch_seg_copy((const SegCopy*)d_segs.p, (const SegChunk*)d_chunks.p,
                           (int)chunks.size(), s);
  }
  void release_all() {
    release(d_segs);
    release(d_chunks);
  }
};

constexpr int kMaxLanes = 2;

struct BatchState {
  bool estimated = false;
  int count = 0;
  std::vector<rvsp_csr_matrix_t> a, b;  // operands as estimated
  std::vector<int> b_of;                // stacked-B index of every pair
  std::vector<int> b_pair;              // first pair of every stacked B
  std::vector<long long> a_row_off, a_nnz_off, b_row_off, b_nnz_off;
  int b_cols = 1;
  DevBuf a_rp, a_ci, a_val, b_rp, b_ci, b_val;
  // Every pair multiplies a matrix by itself through the same value array
  // (C = A * A): the stacked B values are the stacked A values, and only
  // those are staged per compute.
  bool b_val_shared = false;
  const void* b_vals() const { return b_val_shared ? a_val.p : b_val.p; }
  DevBuf ops, gather_idx[kMaxLanes], gather_out[kMaxLanes];
  std::vector<BatchGroup> groups;
  std::vector<int32_t> c_nnz;
  SegTable t_struct, t_vals, t_out_full, t_out_vals;
  std::vector<const void*> struct_key, vals_key, out_key, last_key;
  bool out_filled = false;  // t_out_full already ran for out_key
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> graph_key;
  bool graph_failed = false;
  // Persistent row-stream numeric pass over every group's short-row numeric
  // bins (numeric-only computes): per segment its list-order product offsets,
  // queue-item starts and map; the segment table and the launch's queue.
  struct StreamPart {
    DevBuf rows, soff, items, map;
  };
  std::vector<StreamPart> parts;
  std::vector<StreamSeg> segs;
  std::vector<int> seg_group;  // group of every segment
  DevBuf d_segs, d_queue, s_weights, s_tiles, s_totals;
  int stream_items = 0;
  int build_capw = 0, map_capw = 0, search_capw = 0;  // per pass kind
  int map_grid = 0, search_grid = 0;
  bool stream_ready = false, stream_failed = false;  // segments prepared
  bool map_ready = false, map_failed = false;        // map built (or refused)
  bool segs_stale = true;  // output pointers changed since the last upload
  // The first numeric-only compute times the persistent pass against the
  // groups' own per-bin launches (each with its position map) and keeps the
  // faster; declined stays until the structure changes.
  bool path_chosen = false, stream_declined = false;
  float chosen_ms[2] = {0.0f, 0.0f};  // measured: persistent, per bin
  int graph_launches = 0;
  int repeats = 0;  // consecutive numeric-only computes with the same arrays
  // Computes run the groups on kLanes streams at once (fork and join
  // through events on the descriptor's stream, also inside a captured
  // graph), so one group's long-row tail overlaps the next group's work.
  // Two lanes: more streams than the process's hardware queues (four by
  // default, shared with the caller's streams) made some computes stall.
  static constexpr int kLanes = kMaxLanes;
  hipStream_t lane[kLanes - 1] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kLanes - 1] = {};

  bool lanes_ready() {
    if (fork_ev) return true;
    for (int k = 0; k < kLanes - 1; ++k) {
      if (hipStreamCreateWithFlags(&lane[k], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&join_ev[k], hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        drop_lanes();
        return false;
      }
    }
    if (hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      drop_lanes();
      return false;
    }
    return true;
  }
  void drop_lanes() {
    for (int k = 0; k < kLanes - 1; ++k) {
      if (lane[k]) (void)hipStreamSynchronize(lane[k]), (void)hipStreamDestroy(lane[k]);
      if (join_ev[k]) (void)hipEventDestroy(join_ev[k]);
      lane[k] = nullptr;
      join_ev[k] = nullptr;
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    fork_ev = nullptr;
  }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_key.clear();
  }
  void drop_groups() {
    for (BatchGroup& g : groups) {
      rvsp_spgemm_descr_destroy(g.inner);
      release(g.c_rp);
      release(g.c_ci);
      release(g.c_val);
    }
    groups.clear();
  }
  void drop_stream() {
    for (StreamPart& sp : parts) {
      release(sp.rows);
      release(sp.soff);
      release(sp.items);
      release(sp.map);
    }
    parts.clear();
    segs.clear();
    seg_group.clear();
    for (BatchGroup& g : groups)
      if (g.inner) g.inner->stream_bins = 0;
    stream_ready = false;
    stream_failed = false;
    map_ready = false;
    map_failed = false;
    segs_stale = true;
    path_chosen = false;
    stream_declined = false;
    stream_items = 0;
  }
  ~BatchState() {
    drop_graph();
    drop_stream();
    release(d_segs);
    release(d_queue);
    release(s_weights);
    release(s_tiles);
    release(s_totals);
    drop_groups();
    drop_lanes();
    DevBuf* bufs[] = {&a_rp, &a_ci, &a_val, &b_rp, &b_ci, &b_val, &ops};
    for (DevBuf* x : bufs) release(*x);
    for (int t = 0; t < kMaxLanes; ++t) {
      release(gather_idx[t]);
      release(gather_out[t]);
    }
    t_struct.release_all();
    t_vals.release_all();
    t_out_full.release_all();
    t_out_vals.release_all();
  }
};

void destroy_batch(rvsp_spgemm_descr* d) {
  delete d->batch;
  d->batch = nullptr;
}

void invalidate_batch(rvsp_spgemm_descr* d) {
  if (!d->batch) return;
  d->batch->drop_graph();
  d->batch->drop_stream();
  d->batch->out_filled = false;
  d->batch->last_key.clear();
  for (BatchGroup& g : d->batch->groups) rvsp_spgemm_invalidate_structure(g.inner);
}

namespace {

bool same_matrix(const rvsp_csr_matrix_t& x, const rvsp_csr_matrix_t& y) {
  return x.rows == y.rows && x.cols == y.cols && x.nnz == y.nnz && x.row_ptr == y.row_ptr &&
         x.col_idx == y.col_idx && x.values == y.values;
}

rvsp_csr_matrix_t view(int rows, int cols, long long nnz, const void* rp, const void* ci,
                       const void* val) {
  rvsp_csr_matrix_t m;
  memset(&m, 0, sizeof m);
  m.rows = rows;
  m.cols = cols;
  m.nnz = (int32_t)nnz;
  m.row_ptr = (int32_t*)rp;
  m.col_idx = (int32_t*)ci;
  m.values = (void*)val;
  m.dtype = RVSP_DTYPE_FP32;
  m.format = RVSP_FORMAT_CSR;
  m.owns_data = 0;
  return m;
}

// Stacked row pointers and (shifted) column indices of every A and B.
void build_struct_table(BatchState& bs, const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B) {
  SegTable& t = bs.t_struct;
  t.clear();
  const int n = bs.count;
  for (int p = 0; p < n; ++p) {
    const long long rows = A[p].rows + (p == n - 1 ? 1 : 0);
    t.add(A[p].row_ptr, (int*)bs.a_rp.p + bs.a_row_off[p], rows, (int)bs.a_nnz_off[p], 1);
    t.add(A[p].col_idx, (int*)bs.a_ci.p + bs.a_nnz_off[p], A[p].nnz,
          (int)bs.b_row_off[bs.b_of[p]], 0);
  }
  const int nb = (int)bs.b_pair.size();
  for (int j = 0; j < nb; ++j) {
    const rvsp_csr_matrix_t& Bj = B[bs.b_pair[j]];
    const long long rows = Bj.rows + (j == nb - 1 ? 1 : 0);
    t.add(Bj.row_ptr, (int*)bs.b_rp.p + bs.b_row_off[j], rows, (int)bs.b_nnz_off[j], 1);
    t.add(Bj.col_idx, (int*)bs.b_ci.p + bs.b_nnz_off[j], Bj.nnz, 0, 0);
  }
}

void build_value_table(BatchState& bs, const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B) {
  SegTable& t = bs.t_vals;
  t.clear();
  for (int p = 0; p < bs.count; ++p)
    t.add((const int*)A[p].values, (int*)bs.a_val.p + bs.a_nnz_off[p], A[p].nnz, 0, 0);
  // Stacked B j is pair j's B for every j, and each holds the values array
  // of its pair's A (then also the same entry count and stack offsets).
  bool shared = (int)bs.b_pair.size() == bs.count;
  for (int p = 0; p < bs.count && shared; ++p)
    shared = bs.b_pair[p] == p && bs.b_of[p] == p && B[p].values == A[p].values &&
             B[p].nnz == A[p].nnz && bs.b_nnz_off[p] == bs.a_nnz_off[p];
  bs.b_val_shared = shared;
  if (shared) return;
  for (int j = 0; j < (int)bs.b_pair.size(); ++j) {
    const rvsp_csr_matrix_t& Bj = B[bs.b_pair[j]];
    t.add((const int*)Bj.values, (int*)bs.b_val.p + bs.b_nnz_off[j], Bj.nnz, 0, 0);
  }
}

std::vector<const void*> struct_key(int n, const rvsp_csr_matrix_t* A,
                                    const rvsp_csr_matrix_t* B) {
  std::vector<const void*> k;
  k.reserve((size_t)n * 4);
  for (int p = 0; p < n; ++p) {
    k.push_back(A[p].row_ptr);
    k.push_back(A[p].col_idx);
    k.push_back(B[p].row_ptr);
    k.push_back(B[p].col_idx);
  }
  return k;
}

std::vector<const void*> value_key(int n, const rvsp_csr_matrix_t* A,
                                   const rvsp_csr_matrix_t* B) {
  std::vector<const void*> k;
  k.reserve((size_t)n * 2);
  for (int p = 0; p < n; ++p) {
    k.push_back(A[p].values);
    k.push_back(B[p].values);
  }
  return k;
}

std::vector<const void*> output_key(int n, const rvsp_csr_matrix_t* C) {
  std::vector<const void*> k;
  k.reserve((size_t)n * 3);
  for (int p = 0; p < n; ++p) {
    k.push_back(C[p].row_ptr);
    k.push_back(C[p].col_idx);
    k.push_back(C[p].values);
  }
  return k;
}

rvsp_status_t check_pairs(int count, const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B) {
  for (int p = 0; p < count; ++p) {
    const rvsp_csr_matrix_t* ops[] = {&A[p], &B[p]};
    for (const rvsp_csr_matrix_t* M : ops) {
      if (M->row_ptr == NULL || M->col_idx == NULL || M->values == NULL)
        return RVSP_ERROR_NULL_POINTER;
      if (M->dtype != RVSP_DTYPE_FP32) return RVSP_ERROR_UNSUPPORTED_DTYPE;
      if (M->rows <= 0 || M->cols <= 0 || M->nnz < 0) return RVSP_ERROR_INVALID_ARGUMENT;
    }
    if (A[p].cols != B[p].rows) return RVSP_ERROR_INVALID_ARGUMENT;
  }
  return RVSP_SUCCESS;
}

// Output table of the current C arrays: row pointers of every pair always;
// column indices and values of the groups that do not write in place.
void build_output_tables(BatchState& bs, const rvsp_csr_matrix_t* C) {
  bs.t_out_full.clear();
  bs.t_out_vals.clear();
  for (BatchGroup& g : bs.groups) {
    for (int p = g.p0; p < g.p1; ++p) {
      const long long lr = bs.a_row_off[p] - g.r0;
      bs.t_out_full.add((const int*)g.c_rp.p + lr, C[p].row_ptr, (long long)C[p].rows + 1, 0, 1);
      if (g.in_place) continue;
      const long long o = g.off[p - g.p0];
      const long long n = g.off[p - g.p0 + 1] - o;
      bs.t_out_full.add((const int*)g.c_ci.p + o, C[p].col_idx, n, 0, 0);
      bs.t_out_full.add((const int*)g.c_val.p + o, (int*)C[p].values, n, 0, 0);
      bs.t_out_vals.add((const int*)g.c_val.p + o, (int*)C[p].values, n, 0, 0);
    }
  }
}

DevCsr stacked_a(const BatchState& bs) {
  DevCsr A;
  A.rows = (int)bs.a_row_off.back();
  A.cols = (int)bs.b_row_off.back();
  A.nnz = (int)bs.a_nnz_off.back();
  A.ci = (const int*)bs.a_ci.p;
  A.val = (const float*)bs.a_val.p;
  return A;
}

DevCsr stacked_b(const BatchState& bs) {
  DevCsr B;
  B.rows = (int)bs.b_row_off.back();
  B.cols = bs.b_cols;
  B.nnz = (int)bs.b_nnz_off.back();
  B.rp = (const int*)bs.b_rp.p;
  B.ci = (const int*)bs.b_ci.p;
  B.val = (const float*)bs.b_vals();
  return B;
}

// Prepares the persistent row-stream numeric pass once the groups are
// estimated: per group up to two segments, the rows fed by one B row (bin 1,
// positions need no map) and every other row of a numeric bin whose rows are
// at most kStreamMaxRow long, in ascending row order (one pair's rows next to
// each other, so the waves walking them read the same B). Per segment the
// row list, the list-order product offsets and the queue-item starts.
// Synchronous. On success the groups skip those bins in every compute and the
// batch runs them in one launch (search pass until the map is built); on
// failure nothing changes and the per-bin launches stay.
rvsp_status_t prepare_stream(BatchState& bs, hipStream_t s) {
  bs.drop_stream();
  bs.stream_failed = true;  // until every step below has succeeded
  struct Pick {
    int group, nrows, row_base;
    const int* rows;
    bool ident;
    unsigned bins;
  };
  std::vector<Pick> picks;
  int max_rows = 0, max_group_rows = 0;
  // The passes address B's columns and values by 32-bit byte offsets.
  if (bs.b_nnz_off.back() > kStreamMaxB) return RVSP_SUCCESS;
  for (size_t gi = 0; gi < bs.groups.size() && (int)picks.size() + 2 <= kStreamMaxSegs; ++gi) {
    const rvsp_spgemm_descr* in = bs.groups[gi].inner;
    const PhasePlan& p = in->numeric;
    const int M = (int)(bs.groups[gi].r1 - bs.groups[gi].r0);
    if (p.table.rule.nbins > 1 && bin_rows(p, 1) > 0) {
      Pick pk;
      pk.group = (int)gi;
      pk.nrows = bin_rows(p, 1);
      pk.rows = (const int*)p.rows.p + p.offsets[1];
      pk.row_base = p.row_base.size() > 1 ? p.row_base[1] : -1;
      pk.ident = true;
      pk.bins = 1u << 1;
      picks.push_back(pk);
      max_rows = std::max(max_rows, pk.nrows);
    }
    Pick pk;
    pk.group = (int)gi;
    pk.nrows = 0;
    pk.rows = nullptr;
    pk.row_base = -1;
    pk.ident = false;
    pk.bins = 0;
    // Rows of bin b are at most bound[b] long, except in the last bin,
    // which takes every row no other bin admits.
    for (int b = 2; b < p.table.rule.nbins - 1; ++b)
      if (bin_rows(p, b) > 0 && p.table.rule.bound[b] <= kStreamMaxRow) {
        pk.bins |= 1u << b;
        pk.nrows += bin_rows(p, b);
      }
    if (pk.bins) {
      picks.push_back(pk);
      max_rows = std::max(max_rows, pk.nrows);
      max_group_rows = std::max(max_group_rows, M);
    }
  }
  if (picks.empty()) return RVSP_SUCCESS;
  const int ns = (int)picks.size();
  bs.parts.resize(ns);
  RVSP_TRY(ensure(bs.s_tiles, scan_tile_count(std::max(max_rows, max_group_rows)) * 8));
  RVSP_TRY(ensure(bs.s_totals, (size_t)ns * 8));
  // Row lists of the merged segments: rows of their bins, in row order.
  if (max_group_rows > 0) {
    DevBuf flags, pos;
    RVSP_TRY(ensure(flags, (size_t)max_group_rows * 4));
    RVSP_TRY(ensure(pos, ((size_t)max_group_rows + 1) * 4));
    for (int j = 0; j < ns; ++j) {
      Pick& pk = picks[j];
      if (pk.ident) continue;
      const PhasePlan& p = bs.groups[pk.group].inner->numeric;
      const int M = (int)(bs.groups[pk.group].r1 - bs.groups[pk.group].r0);
      RVSP_HIP_TRY(hipMemsetAsync(flags.p, 0, (size_t)M * 4, s));
      for (int b = 2; b < p.table.rule.nbins; ++b)
        if ((pk.bins >> b) & 1u)
          RVSP_HIP_TRY(launch_stream_mark((const int*)p.rows.p + p.offsets[b],
                                          (size_t)b < p.row_base.size() ? p.row_base[b] : -1,
                                          bin_rows(p, b), (int*)flags.p, s));
      RVSP_HIP_TRY(launch_scan((const int*)flags.p, M, (int*)pos.p,
                               (long long*)bs.s_totals.p + j, (long long*)bs.s_tiles.p, s));
      RVSP_TRY(ensure(bs.parts[j].rows, (size_t)pk.nrows * 4));
      RVSP_HIP_TRY(launch_stream_compact((const int*)flags.p, (const int*)pos.p, M,
                                         (int*)bs.parts[j].rows.p, s));
      pk.rows = (const int*)bs.parts[j].rows.p;
    }
    RVSP_HIP_TRY(hipStreamSynchronize(s));
    release_idle(flags);
    release_idle(pos);
  }
  // Products per list position, their scan (the stream offsets) and totals.
  RVSP_TRY(ensure(bs.s_weights, (size_t)max_rows * 8));
  for (int j = 0; j < ns; ++j) {
    const Pick& pk = picks[j];
    RVSP_TRY(ensure(bs.parts[j].soff, ((size_t)pk.nrows + 1) * 8));
    RVSP_HIP_TRY(launch_stream_weights(pk.row_base >= 0 ? nullptr : pk.rows, pk.row_base,
                                       pk.nrows, (const long long*)bs.groups[pk.group].inner->ops.p,
                                       (long long*)bs.s_weights.p, s));
    RVSP_HIP_TRY(launch_scan64((const long long*)bs.s_weights.p, pk.nrows,
                               (long long*)bs.parts[j].soff.p, (long long*)bs.s_tiles.p,
                               (long long*)bs.s_totals.p + j, s));
  }
  std::vector<long long> totals(ns, 0);
  RVSP_HIP_TRY(hipMemcpyAsync(totals.data(), bs.s_totals.p, (size_t)ns * 8,
                              hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  long long items = 0;
  for (int j = 0; j < ns; ++j) items += (totals[j] + kStreamItem - 1) / kStreamItem;
  if (items > INT_MAX / 2) return RVSP_SUCCESS;
  bs.segs.assign(ns, StreamSeg());
  bs.seg_group.assign(ns, 0);
  int item0 = 0;
  for (int j = 0; j < ns; ++j) {
    const Pick& pk = picks[j];
    const BatchGroup& g = bs.groups[pk.group];
    BatchState::StreamPart& part = bs.parts[j];
    const int nitems = (int)((totals[j] + kStreamItem - 1) / kStreamItem);
    RVSP_TRY(ensure(part.items, ((size_t)nitems + 1) * 4));
    RVSP_HIP_TRY(launch_stream_items((const long long*)part.soff.p, pk.nrows, nitems,
                                     (int*)part.items.p, s));
    StreamSeg& sg = bs.segs[j];
    sg.rows = pk.row_base >= 0 ? nullptr : pk.rows;
    sg.row_base = pk.row_base >= 0 ? pk.row_base : 0;
    sg.nrows = pk.nrows;
    sg.a_rp = (const int*)bs.a_rp.p + g.r0;
    sg.c_rp = (const int*)g.inner->c_rp.p;
    sg.soff = (const long long*)part.soff.p;
    sg.item_first = (const int*)part.items.p;
    sg.total = totals[j];
    sg.c_nnz = g.c_nnz;
    sg.ident = pk.ident ? 1 : 0;
    sg.item0 = item0;
    item0 += nitems;
    bs.seg_group[j] = pk.group;
  }
  RVSP_TRY(ensure(bs.d_segs, (size_t)ns * sizeof(StreamSeg)));
  if (!bs.d_queue.p) RVSP_TRY(ensure(bs.d_queue, sizeof(StreamQueue)));
  RVSP_HIP_TRY(hipMemsetAsync(bs.d_queue.p, 0, sizeof(StreamQueue), s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  bs.stream_items = item0;
  // Accumulators of the full sizes whatever the rows' bound, so a sub-batch
  // holds as many short rows as fit; the build holds every row whole.
  bs.build_capw = kStreamMaxRow;
  bs.map_capw = kStreamNumericRow;
  bs.search_capw = kStreamSearchRow;
  bs.map_grid = stream_grid(kStreamMapPass, bs.map_capw, item0);
  bs.search_grid = stream_grid(kStreamSearchPass, bs.search_capw, item0);
  for (const Pick& pk : picks) bs.groups[pk.group].inner->stream_bins |= pk.bins;
  bs.segs_stale = true;
  bs.stream_failed = false;
  bs.stream_ready = true;
  return RVSP_SUCCESS;
}

// Output arrays of the segments (they follow the caller's C arrays) into the
// device table, before a launch that reads them.
rvsp_status_t upload_segments(BatchState& bs, const rvsp_csr_matrix_t* C, hipStream_t s) {
  if (!bs.segs_stale) return RVSP_SUCCESS;
  for (size_t j = 0; j < bs.segs.size(); ++j) {
    const BatchGroup& g = bs.groups[bs.seg_group[j]];
    StreamSeg& sg = bs.segs[j];
    sg.c_ci = g.in_place ? (const int*)C[g.p0].col_idx : (const int*)g.c_ci.p;
    sg.c_val = g.in_place ? (float*)C[g.p0].values : (float*)g.c_val.p;
    sg.map = sg.ident || !bs.map_ready ? nullptr : (unsigned short*)bs.parts[j].map.p;
  }
  RVSP_HIP_TRY(hipMemcpyAsync(bs.d_segs.p, bs.segs.data(), bs.segs.size() * sizeof(StreamSeg),
                              hipMemcpyHostToDevice, s));
  bs.segs_stale = false;
  return RVSP_SUCCESS;
}

// Builds the position map of every non-identity segment (positions and row
// marks by the build pass, then their order bits) for the numeric-only
// computes. Synchronous. Over the memory budget, or if the build refused a
// write, the map stays off and those computes keep the search pass.
rvsp_status_t build_stream_map(BatchState& bs, const rvsp_csr_matrix_t* C, hipStream_t s) {
  bs.map_failed = true;
  size_t bytes = 0;
  for (const StreamSeg& sg : bs.segs)
    if (!sg.ident) bytes += (size_t)sg.total * 2;
  const size_t budget = std::min<size_t>((size_t)32 << 30, device_free_bytes() / 3);
  if (bytes > budget) return RVSP_SUCCESS;
  for (size_t j = 0; j < bs.segs.size(); ++j)
    if (!bs.segs[j].ident && bs.segs[j].total > 0)
      // 256 entries of slack: the map pass loads whole windows ahead.
      RVSP_TRY(ensure(bs.parts[j].map, (size_t)bs.segs[j].total * 2 + 512));
  bs.map_ready = true;  // the table now points at the maps
  bs.segs_stale = true;
  RVSP_TRY(upload_segments(bs, C, s));
  const int grid = stream_grid(kStreamBuildPass, bs.build_capw, bs.stream_items);
  RVSP_HIP_TRY(launch_stream(kStreamBuildPass, (const StreamSeg*)bs.d_segs.p, (int)bs.segs.size(),
                             bs.stream_items, stacked_a(bs), stacked_b(bs),
                             (StreamQueue*)bs.d_queue.p, bs.build_capw, grid, s));
  for (size_t j = 0; j < bs.segs.size(); ++j)
    if (!bs.segs[j].ident)
      RVSP_HIP_TRY(launch_stream_rounds(bs.segs[j].map, bs.segs[j].total, s));
  unsigned bad = 0;
  RVSP_HIP_TRY(hipMemcpyAsync(&bad, (char*)bs.d_queue.p + offsetof(StreamQueue, bad),
                              sizeof(bad), hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  if (bad) {  // a map that does not match the structure: keep searching
    bs.map_ready = false;
    bs.segs_stale = true;
    return RVSP_ERROR_DEVICE;
  }
  bs.map_failed = false;
  return RVSP_SUCCESS;
}

// RVSP_STREAM_VERIFY=1: batch computes synchronise and report writes the
// persistent pass refused as RVSP_ERROR_DEVICE.
bool stream_verify() {
  static const bool on = [] {
    const char* e = getenv("RVSP_STREAM_VERIFY");
    return e && e[0] == '1';
  }();
  return on;
}

// One persistent launch over every segment: the map pass once the map is
// built, the search pass before (first compute) or without it.
rvsp_status_t launch_stream_numeric(BatchState& bs, const rvsp_csr_matrix_t* C, hipStream_t s) {
  RVSP_TRY(upload_segments(bs, C, s));
  const bool use_map = bs.map_ready;
  RVSP_HIP_TRY(launch_stream(use_map ? kStreamMapPass : kStreamSearchPass,
                             (const StreamSeg*)bs.d_segs.p, (int)bs.segs.size(), bs.stream_items,
                             stacked_a(bs), stacked_b(bs), (StreamQueue*)bs.d_queue.p,
                             use_map ? bs.map_capw : bs.search_capw,
                             use_map ? bs.map_grid : bs.search_grid, s));
  return RVSP_SUCCESS;
}

rvsp_status_t batch_body(BatchState& bs, const rvsp_csr_matrix_t* C, bool full_out,
                         hipStream_t s, bool concurrent);

// The first numeric-only compute with the persistent pass ready: one compute
// through it, then the groups' per-bin launches (one compute that builds
// their position maps, one timed), and the faster path stays; the other
// path's maps are released. Every compute writes the same values.
rvsp_status_t choose_numeric_path(BatchState& bs, const rvsp_csr_matrix_t* C, hipStream_t s) {
  hipEvent_t ev[4] = {};
  for (hipEvent_t& e : ev) RVSP_HIP_TRY(hipEventCreate(&e));
  auto destroy = [&] {
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  std::vector<unsigned> masks;
  rvsp_status_t st = RVSP_SUCCESS;
  hipError_t he = hipEventRecord(ev[0], s);
  if (he == hipSuccess) st = batch_body(bs, C, false, s, true);
  if (he == hipSuccess && st == RVSP_SUCCESS) he = hipEventRecord(ev[1], s);
  if (he == hipSuccess && st == RVSP_SUCCESS) {
    for (BatchGroup& g : bs.groups) {
      masks.push_back(g.inner->stream_bins);
      g.inner->stream_bins = 0;
      g.inner->map_promoted = false;  // the next numeric-only compute promotes
    }
    bs.stream_ready = false;
    st = batch_body(bs, C, false, s, true);
    if (st == RVSP_SUCCESS) he = hipEventRecord(ev[2], s);
    if (st == RVSP_SUCCESS && he == hipSuccess) st = batch_body(bs, C, false, s, true);
    if (st == RVSP_SUCCESS && he == hipSuccess) he = hipEventRecord(ev[3], s);
    if (he == hipSuccess) he = hipEventSynchronize(ev[3]);
  }
  float t_stream = 0.0f, t_bin = 0.0f;
  if (he == hipSuccess && st == RVSP_SUCCESS) he = hipEventElapsedTime(&t_stream, ev[0], ev[1]);
  if (he == hipSuccess && st == RVSP_SUCCESS) he = hipEventElapsedTime(&t_bin, ev[2], ev[3]);
  destroy();
  if (masks.size() == bs.groups.size()) {
    // Back to the persistent pass unless the per-bin path was measured faster.
    const bool keep = he != hipSuccess || st != RVSP_SUCCESS || t_stream <= t_bin;
    if (keep) {
      for (size_t i = 0; i < bs.groups.size(); ++i) {
        bs.groups[i].inner->stream_bins = masks[i];
        RVSP_TRY(release_position_map(bs.groups[i].inner));
      }
      bs.stream_ready = true;
    } else {
      bs.drop_stream();  // releases the persistent pass's maps
      bs.stream_declined = true;
      // The groups' promoted bins stay on their maps.
      for (BatchGroup& g : bs.groups) g.inner->map_promoted = true;
    }
  }
  bs.path_chosen = true;
  bs.chosen_ms[0] = t_stream;
  bs.chosen_ms[1] = t_bin;
  RVSP_HIP_TRY(he);
  return st;
}

// The groups' long-row numeric bins (workgroup tiles), deferred by their
// computes: one fused launch per tile size and up to kMaxFusedSegs bins, on
// the descriptor's stream after every group's fill and other bins.
rvsp_status_t launch_deferred(BatchState& bs, hipStream_t s) {
  std::vector<DeferredBin> all;
  for (BatchGroup& g : bs.groups) {
    all.insert(all.end(), g.inner->deferred.begin(), g.inner->deferred.end());
    g.inner->deferred.clear();
  }
  if (all.empty()) return RVSP_SUCCESS;
  const rvsp_spgemm_descr* owner = bs.groups.front().inner;
  DevCsr A;
  A.rows = (int)bs.a_row_off.back();
  A.cols = (int)bs.b_row_off.back();
  A.nnz = (int)bs.a_nnz_off.back();
  A.ci = (const int*)bs.a_ci.p;
  A.val = (const float*)bs.a_val.p;
  DevCsr B;
  B.rows = (int)bs.b_row_off.back();
  B.cols = bs.b_cols;
  B.nnz = (int)bs.b_nnz_off.back();
  B.rp = (const int*)bs.b_rp.p;
  B.ci = (const int*)bs.b_ci.p;
  B.val = (const float*)bs.b_vals();
  std::vector<bool> done(all.size(), false);
  for (size_t i = 0; i < all.size(); ++i) {
    if (done[i]) continue;
    const BinSpec& sp = all[i].spec;
    TiledSegs segs;
    int max_row = 0;
    for (size_t j = i; j < all.size() && segs.n < kMaxFusedSegs; ++j) {
      if (done[j] || all[j].spec.cap != sp.cap) continue;
      done[j] = true;
      TiledSeg& t = segs.seg[segs.n++];
      t.rows = all[j].rows;
      t.a_rp = all[j].a_rp;
      t.c_rp = all[j].c_rp;
      t.c_ci = all[j].c_ci;
      t.c_val = all[j].c_val;
      t.nrows = all[j].nrows;
      t.start = segs.total;
      segs.total += all[j].nrows;
      max_row = std::max(max_row, all[j].max_row);
    }
    RVSP_HIP_TRY(launch_numeric_tiled_fused(segs, sp.cap, max_row, A, B, owner->split,
                                            sp.lds_bytes, s));
  }
  return RVSP_SUCCESS;
}

// The compute body: values into the stack, every group's compute, results
// out. With `concurrent` the groups alternate between the descriptor's
// stream and the batch's extra lanes (forked from and joined back into the
// descriptor's stream); otherwise, as under capture, all run on it.
rvsp_status_t batch_body(BatchState& bs, const rvsp_csr_matrix_t* C, bool full_out,
                         hipStream_t s, bool concurrent) {
  for (BatchGroup& g : bs.groups) g.inner->deferred.clear();  // nothing left from a failed call
  RVSP_HIP_TRY(bs.t_vals.launch(s));
  const rvsp_csr_matrix_t Bst = view((int)bs.b_row_off.back(), bs.b_cols, bs.b_nnz_off.back(),
                                     bs.b_rp.p, bs.b_ci.p, bs.b_vals());
  const bool conc = concurrent && bs.groups.size() > 1 && bs.lanes_ready();
  if (conc) {
    RVSP_HIP_TRY(hipEventRecord(bs.fork_ev, s));
    for (int k = 0; k < BatchState::kLanes - 1; ++k)
      RVSP_HIP_TRY(hipStreamWaitEvent(bs.lane[k], bs.fork_ev, 0));
  }
  int gi = 0;
  for (BatchGroup& g : bs.groups) {
    const int lane = conc ? gi % BatchState::kLanes : 0;
    ++gi;
    RVSP_TRY(rvsp_spgemm_set_stream(g.inner, (void*)(lane ? bs.lane[lane - 1] : s)));
    const rvsp_csr_matrix_t Ast =
        view((int)(g.r1 - g.r0), (int)bs.b_row_off.back(),
             bs.a_nnz_off[g.p1] - bs.a_nnz_off[g.p0], (const int*)bs.a_rp.p + g.r0, bs.a_ci.p,
             bs.a_val.p);
    const void* ci = g.in_place ? (const void*)C[g.p0].col_idx : g.c_ci.p;
    const void* val = g.in_place ? (const void*)C[g.p0].values : g.c_val.p;
    rvsp_csr_matrix_t Cg = view(Ast.rows, bs.b_cols, g.c_nnz, g.c_rp.p, ci, val);
    char token = 0;
    RVSP_TRY(rvsp_spgemm_compute(g.inner, &Ast, &Bst, &Cg, &token));
  }
  // Numeric-only computes: every group's streamed bins in one launch (the
  // groups' own computes skip them), queued behind the groups' other
  // (long-row) launches so those get their CUs before the persistent waves
  // take the rest.
  if (bs.stream_ready && !full_out) RVSP_TRY(launch_stream_numeric(bs, C, s));
  if (conc) {
    for (int k = 0; k < BatchState::kLanes - 1; ++k) {
      RVSP_HIP_TRY(hipEventRecord(bs.join_ev[k], bs.lane[k]));
      RVSP_HIP_TRY(hipStreamWaitEvent(s, bs.join_ev[k], 0));
    }
  }
  // A full compute: the streamed bins once every group's columns are filled.
  if (bs.stream_ready && full_out) RVSP_TRY(launch_stream_numeric(bs, C, s));
  RVSP_TRY(launch_deferred(bs, s));
  const SegTable& out = full_out ? bs.t_out_full : bs.t_out_vals;
  if (!out.chunks.empty()) RVSP_HIP_TRY(out.launch(s));
  return RVSP_SUCCESS;
}

}  // namespace

}  // namespace rvsp

using namespace rvsp;

extern "C" {

rvsp_status_t rvsp_spgemm_batch_work_estimation(rvsp_spgemm_descr_t descr, int32_t count,
                                                const rvsp_csr_matrix_t* A,
                                                const rvsp_csr_matrix_t* B, int32_t* c_nnz) {
  if (descr == NULL || A == NULL || B == NULL || c_nnz == NULL) return RVSP_ERROR_NULL_POINTER;
  if (count < 1) return RVSP_ERROR_INVALID_ARGUMENT;
  if (descr->memory != RVSP_MEMORY_DEVICE) return RVSP_ERROR_INVALID_ARGUMENT;
  RVSP_TRY(check_pairs(count, A, B));
  rvsp_spgemm_descr* d = descr;
  destroy_batch(d);
  d->batch = new (std::nothrow) BatchState();
  if (!d->batch) return RVSP_ERROR_ALLOCATION_FAILED;
  BatchState& bs = *d->batch;
  RVSP_TRY(device_ready());
  hipStream_t s;
  RVSP_TRY(descr_stream(d, &s));

  bs.count = count;
  bs.a.assign(A, A + count);
  bs.b.assign(B, B + count);
  // Stack offsets; B operands that are the same arrays are stacked once.
  bs.b_of.resize(count);
  {
    using BKey = std::tuple<const void*, const void*, const void*, int, int, int>;
    std::map<BKey, int> seen;
    for (int p = 0; p < count; ++p) {
      const BKey k(B[p].row_ptr, B[p].col_idx, B[p].values, B[p].rows, B[p].cols, B[p].nnz);
      auto it = seen.find(k);
      if (it == seen.end()) {
        it = seen.emplace(k, (int)bs.b_pair.size()).first;
        bs.b_pair.push_back(p);
      }
      bs.b_of[p] = it->second;
    }
  }
  const int nb = (int)bs.b_pair.size();
  bs.a_row_off.assign(count + 1, 0);
  bs.a_nnz_off.assign(count + 1, 0);
  bs.b_row_off.assign(nb + 1, 0);
  bs.b_nnz_off.assign(nb + 1, 0);
  for (int p = 0; p < count; ++p) {
    bs.a_row_off[p + 1] = bs.a_row_off[p] + A[p].rows;
    bs.a_nnz_off[p + 1] = bs.a_nnz_off[p] + A[p].nnz;
  }
  bs.b_cols = 1;
  for (int j = 0; j < nb; ++j) {
    const rvsp_csr_matrix_t& Bj = B[bs.b_pair[j]];
    bs.b_row_off[j + 1] = bs.b_row_off[j] + Bj.rows;
    bs.b_nnz_off[j + 1] = bs.b_nnz_off[j] + Bj.nnz;
    bs.b_cols = std::max(bs.b_cols, Bj.cols);
  }
  // The stacks are int32-indexed like every operand.
  if (bs.a_row_off[count] >= INT_MAX || bs.a_nnz_off[count] > INT_MAX ||
      bs.b_row_off[nb] >= INT_MAX || bs.b_nnz_off[nb] > INT_MAX)
    return RVSP_ERROR_INVALID_ARGUMENT;
  RVSP_TRY(ensure(bs.a_rp, ((size_t)bs.a_row_off[count] + 1) * 4));
  RVSP_TRY(ensure(bs.a_ci, (size_t)bs.a_nnz_off[count] * 4));
  RVSP_TRY(ensure(bs.a_val, (size_t)bs.a_nnz_off[count] * 4));
  RVSP_TRY(ensure(bs.b_rp, ((size_t)bs.b_row_off[nb] + 1) * 4));
  RVSP_TRY(ensure(bs.b_ci, (size_t)bs.b_nnz_off[nb] * 4));
  RVSP_TRY(ensure(bs.b_val, (size_t)bs.b_nnz_off[nb] * 4));
  build_struct_table(bs, A, B);
  RVSP_TRY(bs.t_struct.upload(s));
  RVSP_HIP_TRY(bs.t_struct.launch(s));
  bs.struct_key = struct_key(count, A, B);

  // Products per pair, then groups of consecutive pairs under the limit.
  std::vector<OpsChunk> oc;
  for (int p = 0; p < count; ++p)
    for (long long b = bs.a_nnz_off[p]; b < bs.a_nnz_off[p + 1]; b += 16 * kSegChunk) {
      OpsChunk c;
      c.pair = p;
      c.pad = 0;
      c.begin = b;
      c.end = std::min(bs.a_nnz_off[p + 1], b + 16 * kSegChunk);
      oc.push_back(c);
    }
  std::vector<unsigned long long> ops(count, 0);
  RVSP_TRY(ensure(bs.ops, (size_t)count * 8));
  RVSP_HIP_TRY(hipMemsetAsync(bs.ops.p, 0, (size_t)count * 8, s));
  DevBuf d_oc;
  if (!oc.empty()) {
    RVSP_TRY(ensure(d_oc, oc.size() * sizeof(OpsChunk)));
    hipError_t e = hipMemcpyAsync(d_oc.p, oc.data(), oc.size() * sizeof(OpsChunk),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = launch_pair_ops((const OpsChunk*)d_oc.p, (int)oc.size(), (const int*)bs.a_ci.p,
                          (const int*)bs.b_rp.p, (unsigned long long*)bs.ops.p, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ops.data(), bs.ops.p, (size_t)count * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    release(d_oc);
    RVSP_HIP_TRY(e);
  }
  const long long limit = d->batch_limit > 0 ? d->batch_limit : INT_MAX;
  {
    BatchGroup g;
    long long acc = 0;
    for (int p = 0; p < count; ++p) {
      if (p > g.p0 && acc + (long long)ops[p] > limit) {
        g.p1 = p;
        bs.groups.push_back(g);
        g = BatchGroup();
        g.p0 = p;
        acc = 0;
      }
      acc += (long long)ops[p];
    }
    g.p1 = count;
    bs.groups.push_back(g);
  }

  // One inner descriptor per group over its rows of the stacked A.
  bs.c_nnz.assign(count, 0);
  const rvsp_csr_matrix_t Bst = view((int)bs.b_row_off[nb], bs.b_cols, bs.b_nnz_off[nb],
                                     bs.b_rp.p, bs.b_ci.p, bs.b_vals());
  for (BatchGroup& g : bs.groups) {
    g.r0 = bs.a_row_off[g.p0];
    g.r1 = bs.a_row_off[g.p1];
    RVSP_TRY(rvsp_spgemm_descr_create(&g.inner));
    if (&g != &bs.groups.front()) g.inner->split_owner = bs.groups.front().inner;
    RVSP_TRY(rvsp_spgemm_set_algo(g.inner, d->algo));
    RVSP_TRY(rvsp_spgemm_set_memory(g.inner, RVSP_MEMORY_DEVICE));
    RVSP_TRY(rvsp_spgemm_set_stream(g.inner, (void*)s));
    RVSP_TRY(rvsp_spgemm_set_async(g.inner, 1));
    g.inner->defer_block = d->batch_fuse;
    if (!d->choices.empty())
      RVSP_TRY(rvsp_spgemm_set_bin_choices(g.inner, d->choices.data(),
                                           (int32_t)d->choices.size()));
  }
  // The groups are estimated kLanes at a time: host thread t estimates
  // groups t, t + kLanes, ... on lane t (each estimation waits for its own
  // counts, so the lanes' host round trips and kernels overlap). Different
  // descriptors are independent; the library's shared tables are locked.
  const int lanes = bs.groups.size() > 1 && bs.lanes_ready() ? BatchState::kLanes : 1;
  if (lanes > 1)
    for (BatchGroup& g : bs.groups) g.inner->shared_cu = true;
  auto estimate_lane = [&](int t) -> rvsp_status_t {
    hipStream_t ls = t ? bs.lane[t - 1] : s;
    DevBuf& gidx = bs.gather_idx[t];
    DevBuf& gout = bs.gather_out[t];
    for (size_t gi = t; gi < bs.groups.size(); gi += lanes) {
      BatchGroup& g = bs.groups[gi];
      RVSP_TRY(rvsp_spgemm_set_stream(g.inner, (void*)ls));
      const rvsp_csr_matrix_t Ast =
          view((int)(g.r1 - g.r0), (int)bs.b_row_off[nb], bs.a_nnz_off[g.p1] - bs.a_nnz_off[g.p0],
               (const int*)bs.a_rp.p + g.r0, bs.a_ci.p, bs.a_val.p);
      size_t ws = 0;
      RVSP_TRY(rvsp_spgemm_work_estimation(g.inner, &Ast, &Bst, &ws, &g.c_nnz));
      // Pair boundaries inside the group's row pointers.
      const int np = g.p1 - g.p0;
      std::vector<long long> idx(np + 1);
      for (int i = 0; i <= np; ++i) idx[i] = bs.a_row_off[g.p0 + i] - g.r0;
      std::vector<int> at(np + 1);
      RVSP_TRY(ensure(gidx, (size_t)(np + 1) * 8));
      RVSP_TRY(ensure(gout, (size_t)(np + 1) * 4));
      RVSP_HIP_TRY(hipMemcpyAsync(gidx.p, idx.data(), (size_t)(np + 1) * 8,
                                  hipMemcpyHostToDevice, ls));
      RVSP_HIP_TRY(launch_gather_i32((const int*)g.inner->c_rp.p, (const long long*)gidx.p,
                                     np + 1, (int*)gout.p, ls));
      RVSP_HIP_TRY(hipMemcpyAsync(at.data(), gout.p, (size_t)(np + 1) * 4,
                                  hipMemcpyDeviceToHost, ls));
      RVSP_HIP_TRY(hipStreamSynchronize(ls));
      g.off.assign(np + 1, 0);
      for (int i = 0; i <= np; ++i) g.off[i] = at[i];
      for (int i = 0; i < np; ++i) bs.c_nnz[g.p0 + i] = at[i + 1] - at[i];
      RVSP_TRY(ensure(g.c_rp, (size_t)(g.r1 - g.r0 + 1) * 4));
    }
    return RVSP_SUCCESS;
  };
  std::vector<rvsp_status_t> lane_status(lanes, RVSP_SUCCESS);
  {
    std::vector<std::thread> helpers;
    for (int t = 1; t < lanes; ++t)
      helpers.emplace_back([&, t] { lane_status[t] = estimate_lane(t); });
    lane_status[0] = estimate_lane(0);
    for (std::thread& h : helpers) h.join();
  }
  for (rvsp_status_t st : lane_status) RVSP_TRY(st);
  for (int p = 0; p < count; ++p) c_nnz[p] = bs.c_nnz[p];
  bs.estimated = true;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_batch_compute(rvsp_spgemm_descr_t descr, int32_t count,
                                        const rvsp_csr_matrix_t* A, const rvsp_csr_matrix_t* B,
                                        rvsp_csr_matrix_t* C) {
  if (descr == NULL || A == NULL || B == NULL || C == NULL) return RVSP_ERROR_NULL_POINTER;
  rvsp_spgemm_descr* d = descr;
  if (!d->batch || !d->batch->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  BatchState& bs = *d->batch;
  if (count != bs.count) return RVSP_ERROR_INVALID_ARGUMENT;
  RVSP_TRY(check_pairs(count, A, B));
  for (int p = 0; p < count; ++p) {
    const rvsp_csr_matrix_t &a = bs.a[p], &b = bs.b[p];
    if (A[p].rows != a.rows || A[p].cols != a.cols || A[p].nnz != a.nnz ||
        B[p].rows != b.rows || B[p].cols != b.cols || B[p].nnz != b.nnz)
      return RVSP_ERROR_INVALID_ARGUMENT;
  }
  for (int p = 0; p < count; ++p) {
    if (C[p].row_ptr == NULL || C[p].col_idx == NULL || C[p].values == NULL)
      return RVSP_ERROR_NULL_POINTER;
    if (C[p].dtype != RVSP_DTYPE_FP32) return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (C[p].rows != A[p].rows || C[p].cols != B[p].cols) return RVSP_ERROR_INVALID_ARGUMENT;
    if (C[p].nnz < bs.c_nnz[p]) return RVSP_ERROR_INVALID_ARGUMENT;
  }
  // Stacked B rows are shared by the pairs that gave the same B arrays at
  // estimation; a compute must keep that sharing.
  for (int p = 0; p < count; ++p)
    if (!same_matrix(B[p], B[bs.b_pair[bs.b_of[p]]])) return RVSP_ERROR_INVALID_ARGUMENT;

  hipStream_t s;
  RVSP_TRY(descr_stream(d, &s));
  // Structure arrays moved: stack them again (same shapes, same offsets).
  std::vector<const void*> sk = struct_key(count, A, B);
  if (sk != bs.struct_key) {
    bs.drop_graph();
    bs.drop_stream();
    build_struct_table(bs, A, B);
    RVSP_TRY(bs.t_struct.upload(s));
    RVSP_HIP_TRY(bs.t_struct.launch(s));
    bs.struct_key = sk;
    for (BatchGroup& g : bs.groups) rvsp_spgemm_invalidate_structure(g.inner);
    bs.out_filled = false;
  }
  std::vector<const void*> vk = value_key(count, A, B);
  if (vk != bs.vals_key) {
    bs.drop_graph();
    build_value_table(bs, A, B);
    RVSP_TRY(bs.t_vals.upload(s));
    bs.vals_key = vk;
  }
  std::vector<const void*> ok = output_key(count, C);
  if (ok != bs.out_key) {
    bs.drop_graph();
    bs.segs_stale = true;
    for (BatchGroup& g : bs.groups) {
      bool in_place = true;
      for (int p = g.p0 + 1; p < g.p1 && in_place; ++p) {
        const long long o = g.off[p - g.p0];
        in_place = C[p].col_idx == C[g.p0].col_idx + o &&
                   (const float*)C[p].values == (const float*)C[g.p0].values + o;
      }
      g.in_place = in_place;
      if (!in_place) {
        RVSP_TRY(ensure(g.c_ci, (size_t)std::max(1, g.c_nnz) * 4));
        RVSP_TRY(ensure(g.c_val, (size_t)std::max(1, g.c_nnz) * 4));
      }
    }
    build_output_tables(bs, C);
    RVSP_TRY(bs.t_out_full.upload(s));
    RVSP_TRY(bs.t_out_vals.upload(s));
    bs.out_key = ok;
    bs.out_filled = false;
  }

  std::vector<const void*> key = sk;
  key.insert(key.end(), vk.begin(), vk.end());
  key.insert(key.end(), ok.begin(), ok.end());
  const bool repeat = bs.out_filled && key == bs.last_key;
  bs.last_key = key;
  bs.repeats = repeat ? bs.repeats + 1 : 0;
  // The persistent row-stream pass: segments once per structure (before the
  // first compute), the map at the first numeric-only compute.
  if (d->batch_persistent && !bs.stream_ready && !bs.stream_failed && !bs.stream_declined) {
    if (prepare_stream(bs, s) != RVSP_SUCCESS) {
      bs.drop_stream();
      bs.stream_failed = true;
      (void)hipGetLastError();
    }
  }
  if (repeat && bs.stream_ready && !bs.map_ready && !bs.map_failed) {
    if (build_stream_map(bs, C, s) != RVSP_SUCCESS) {
      bs.map_ready = false;
      bs.map_failed = true;
      bs.segs_stale = true;
      (void)hipGetLastError();
    }
  }
  // The first numeric-only compute runs eagerly (the groups may switch to
  // their position maps, which synchronises); the second is captured.
  if (repeat && bs.repeats >= 2 && d->batch_graph && !bs.graph_failed) {
    if (!bs.exec) {
      // The segment table goes to the device before the capture (a host
      // copy inside it would be replayed from pageable memory each time).
      if (bs.stream_ready && bs.segs_stale) {
        RVSP_TRY(upload_segments(bs, C, s));
        RVSP_HIP_TRY(hipStreamSynchronize(s));
      }
      // Capture one numeric-only compute; later ones replay it.
      RVSP_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      // The lanes are forked from and joined into the captured stream by
      // events, so the graph keeps their overlap.
      const rvsp_status_t st = batch_body(bs, C, false, s, true);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(s, &g);
      hipGraphExec_t ex = nullptr;
      hipError_t ei = hipErrorUnknown;
      if (st == RVSP_SUCCESS && ec == hipSuccess && g)
        ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      if (ei == hipSuccess) {
        bs.graph = g;
        bs.exec = ex;
      } else {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        bs.graph_failed = true;
      }
    }
    if (bs.exec) {
      RVSP_HIP_TRY(hipGraphLaunch(bs.exec, s));
      ++bs.graph_launches;
    } else {
      RVSP_TRY(batch_body(bs, C, false, s, true));
    }
  } else if (repeat && bs.stream_ready && bs.map_ready && !bs.path_chosen &&
             d->batch_persistent == 1) {
    RVSP_TRY(choose_numeric_path(bs, C, s));
  } else {
    RVSP_TRY(batch_body(bs, C, !bs.out_filled, s, true));
    bs.out_filled = true;
  }
  if (!d->async) RVSP_HIP_TRY(hipStreamSynchronize(s));
  if (bs.stream_ready && stream_verify()) {
    // Diagnostic mode: every compute checks that the persistent pass
    // refused no write.
    unsigned bad = 0;
    RVSP_HIP_TRY(hipMemcpyAsync(&bad, (char*)bs.d_queue.p + offsetof(StreamQueue, bad),
                                sizeof(bad), hipMemcpyDeviceToHost, s));
    RVSP_HIP_TRY(hipStreamSynchronize(s));
    if (bad) return RVSP_ERROR_DEVICE;
  }
  for (int p = 0; p < count; ++p) {
    C[p].nnz = bs.c_nnz[p];
    C[p].format = RVSP_FORMAT_CSR;
  }
  return RVSP_SUCCESS;
}

int32_t rvsp_spgemm_batch_groups(rvsp_spgemm_descr_t descr) {
  if (descr == NULL || !descr->batch || !descr->batch->estimated) return -1;
  return (int32_t)descr->batch->groups.size();
}

int32_t rvsp_spgemm_batch_graph_launches(rvsp_spgemm_descr_t descr) {
  if (descr == NULL || !descr->batch) return 0;
  return descr->batch->graph_launches;
}

rvsp_status_t rvsp_spgemm_batch_set_graph(rvsp_spgemm_descr_t descr, int on) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  descr->batch_graph = on != 0;
  if (descr->batch) descr->batch->drop_graph();
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_batch_set_persistent(rvsp_spgemm_descr_t descr, int mode) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  if (mode < 0 || mode > 2) return RVSP_ERROR_INVALID_ARGUMENT;
  descr->batch_persistent = mode;
  if (descr->batch) {
    descr->batch->drop_graph();
    descr->batch->drop_stream();
    for (BatchGroup& g : descr->batch->groups)
      if (g.inner) g.inner->map_promoted = false;  // per-bin maps are promoted anew
  }
  return RVSP_SUCCESS;
}

int32_t rvsp_spgemm_batch_stream_segments(rvsp_spgemm_descr_t descr) {
  if (descr == NULL || !descr->batch || !descr->batch->stream_ready) return 0;
  return (int32_t)descr->batch->segs.size();
}

int32_t rvsp_spgemm_batch_numeric_path(rvsp_spgemm_descr_t descr, float* persistent_ms,
                                       float* per_bin_ms) {
  if (descr == NULL || !descr->batch) return 0;
  const BatchState& bs = *descr->batch;
  if (persistent_ms) *persistent_ms = bs.chosen_ms[0];
  if (per_bin_ms) *per_bin_ms = bs.chosen_ms[1];
  if (bs.stream_declined) return 2;
  return bs.stream_ready ? 1 : 0;
}

rvsp_status_t rvsp_spgemm_batch_set_fuse(rvsp_spgemm_descr_t descr, int on) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  if (descr->batch && descr->batch->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  descr->batch_fuse = on != 0;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_batch_set_group_limit(rvsp_spgemm_descr_t descr,
                                                int64_t max_products) {
  if (descr == NULL) return RVSP_ERROR_NULL_POINTER;
  if (max_products < 0 || max_products > INT_MAX) return RVSP_ERROR_INVALID_ARGUMENT;
  if (descr->batch && descr->batch->estimated) return RVSP_ERROR_INVALID_ARGUMENT;
  descr->batch_limit = max_products;
  return RVSP_SUCCESS;
}

}  // extern "C"
