// kernels.hpp — host-side launch interface of the gfx950 SpGEMM kernels.
//
// Phases (reference counterparts in parentheses):
//   analyze    per-row product count Op_i and B-column span
//              (rvsp_symbolic.c:27-73 op counting; bench.c:215-240)
//   classify   row -> bin for one phase, from a BinRule table
//   partition  stable counting sort of rows by bin (ascending row order
//              inside every bin)
//   count      exact distinct-column count per row (rvsp_symbolic.c:13-88)
//   scan       exclusive scan into C.row_ptr with an int64 total
//              (rvsp_symbolic.c:75-85, overflow guard :69-72)
//   fill       sorted column indices per row (rvsp_symbolic.c:91-138)
//   numeric    values in reference accumulation order
//              (gustavson_core_f32.inc:18-64, accum_scalar_f32.c:19-71)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rvsp {

// Device view of a CSR operand. Pointers are device addresses.
struct DevCsr {
  int rows = 0, cols = 0, nnz = 0;
  const int* rp = nullptr;
  const int* ci = nullptr;
  const float* val = nullptr;
};

// Per-row analysis results (device arrays of length rows(A)).
struct RowStats {
  long long* ops = nullptr;  // Op_i = sum of |B_k| over A's row i
  int* lo = nullptr;         // smallest B column reached by row i (INT_MAX if none)
  int* hi = nullptr;         // largest B column reached by row i (-1 if none)
  int* maxb = nullptr;       // longest B row reached by row i
};

// Accumulator kinds. The same enum names the kernel family used for a bin.
enum Kind : int {
  KIND_NONE = 0,     // rows with nothing to do (empty output rows)
  KIND_DIRECT = 1,   // one A entry: the C row is a scaled copy of one B row
  KIND_HASH = 2,     // LDS hash table per row, G lanes per row (count, fill)
  KIND_SEARCH = 3,   // LDS sorted columns + binary search, G lanes per row (numeric)
  KIND_BITMAP = 4,   // one workgroup per row, LDS bitmap over the row's span (count, fill)
  KIND_BLOCK = 5,    // one workgroup per tile of a long row, LDS columns (numeric)
  KIND_GLOBAL = 6,   // one workgroup per row, dense scratch in global memory
  KIND_MAP = 7       // position map + LDS accumulators, G lanes per row (numeric)
};

enum Phase : int { PHASE_COUNT = 0, PHASE_FILL = 1, PHASE_NUMERIC = 2 };

constexpr int kMaxBins = 24;

// Row -> bin rule. Bin 0 is reserved for rows that need no table (empty or
// single-A-entry rows). Other rows go to the first bin b >= 1 whose
// `bound[b]` >= the row's size metric and whose `span_max[b]` >= the row's
// column span; rows matching none go to the last bin, whose kind must take
// rows of any length (GLOBAL, or tiled BLOCK for numeric). The size metric is the distinct-column upper bound
// min(Op_i, span_i) for the count phase and the exact nnz of the C row for
// fill and numeric.
struct BinRule {
  int nbins = 0;
  int bound[kMaxBins];
  int span_max[kMaxBins];
};

// Launch parameters of one bin.
struct BinSpec {
  int kind = KIND_NONE;
  int group = 1;     // lanes per row (power of two <= 64, or 256 for block kinds)
  int cap = 0;       // per-row LDS capacity: table slots (HASH), columns
                     // (SEARCH, MAP), positions per tile (BLOCK), bitmap bits (BITMAP)
  int lds_bytes = 0; // dynamic LDS per workgroup
  int variant = 0;   // kernel variant of the kind (MAP: 1 = A metadata staged in LDS)
};

// Scratch for the partition step.
struct PartitionScratch {
  unsigned char* bin_of = nullptr;  // rows
  int* tile_hist = nullptr;         // ntiles * kMaxBins
  int* tile_base = nullptr;         // ntiles * kMaxBins
  int* bin_offsets = nullptr;       // kMaxBins + 1
  int tile_rows = 0;
  int ntiles = 0;
};

// Global-memory scratch for GLOBAL bins (one slot per resident workgroup).
struct GlobalSlots {
  void* base = nullptr;
  size_t slot_bytes = 0;
  int nslots = 0;
};

// --- analysis ---------------------------------------------------------------
hipError_t launch_analyze(const DevCsr& A, const DevCsr& B, RowStats st,
                          int* long_list, int* n_long, hipStream_t s);

// --- binning ----------------------------------------------------------------
// Count phase: classifies by min(Op_i, span_i); writes cnt[i] directly for
// rows whose count needs no table (Op_i == 0, one A entry, Op_i == 1).
hipError_t launch_classify_count(const DevCsr& A, RowStats st, const BinRule& r,
                                 unsigned char* bin_of, int* cnt, hipStream_t s);
// Fill/numeric phases: classifies by the exact C-row nnz.
// kept (may be null): rows with kept[i + 1] > kept[i] had their columns kept
// by the count phase and go to bin 0.
hipError_t launch_classify_nnz(const DevCsr& A, RowStats st, const int* c_rp,
                               const BinRule& r, unsigned char* bin_of,
                               hipStream_t s, const long long* kept = nullptr);
hipError_t launch_partition(int rows, int nbins, PartitionScratch& ps,
                            int* rows_out, hipStream_t s);

// Per-bin statistics of a partitioned phase: out[b * kBinStats + s] with
// s = rows, sum nnz(A row), sum Op_i, sum size metric, max size metric,
// sum span, max longest B row, max Op_i. `offsets` is the device copy of the
// phase's bin offsets; c_rp is null for the count phase.
constexpr int kBinStats = 8;
hipError_t launch_bin_stats(const int* rows, const int* offsets, int nbins, int total_rows,
                            const int* a_rp, RowStats st, const int* c_rp,
                            unsigned long long* out, hipStream_t s);

// --- symbolic -----------------------------------------------------------
// keep (may be null): LDS bitmap rows also write their sorted columns to
// keep + koff[row] (koff: exclusive scan of min(Op_i, span_i) over them).
hipError_t launch_count_bin(const BinSpec& b, const int* rows, int nrows,
                            const DevCsr& A, const DevCsr& B, RowStats st,
                            int* cnt, const GlobalSlots& gs, hipStream_t s,
                            int* keep = nullptr, const long long* koff = nullptr);
// Copies the kept columns of the listed rows to C.
hipError_t launch_copy_kept(const int* rows, int nrows, const int* keep,
                            const long long* koff, const int* c_rp, int* c_ci,
                            hipStream_t s);
hipError_t launch_scan(const int* cnt, int rows, int* c_rp, long long* total,
                       long long* tile_sums, hipStream_t s);
// LDS bitmaps above this size fit once per CU; their workgroups get 1024
// threads.
constexpr int kWideBitmapBytes = 80 * 1024;

// Canonical-CSR check of device arrays (reference rvsp_support.c:36-126).
// Failures are keyed (row << 32) | code and *key keeps the smallest (the
// caller sets it to ~0 first): code 0 a row pointer that decreases or (with
// column indices) passes nnz, 1 a non-empty row without column indices,
// 2 + j the row's entry j
// (out of range, or not above entry j - 1). The row pass covers all rows;
// the column pass the entries of rows [0, limit), which it reads from
// row_ptr[limit] (<= nnz when the row pass found nothing before limit).
hipError_t launch_csr_check_rows(int rows, int nnz, const int* row_ptr, bool has_cols,
                                 unsigned long long* key, hipStream_t s);
// FP64 (i8 false: double values, double sums) and INT8 -> INT32 (i8 true)
// numeric phase over C's structure (c_rp, sorted c_ci): sums per position in
// the reference's order into c_acc, non-zero sums per row into nonzero; then
// the rows' non-zero entries to out_rp / out_ci / out_val (legacy.hip).
hipError_t launch_numeric_legacy(bool i8, int rows, const int* a_rp, const int* a_ci,
                                 const void* a_val, const int* b_rp, const int* b_ci,
                                 const void* b_val, const int* c_rp, const int* c_ci,
                                 void* c_acc, int* nonzero, hipStream_t s);
hipError_t launch_compact_legacy(bool i8, int rows, const int* c_rp, const int* c_ci,
                                 const void* c_acc, const int* out_rp, int* out_ci,
                                 void* out_val, hipStream_t s);
hipError_t launch_csr_check_cols(int limit, int cols, int nnz, const int* row_ptr,
                                 const int* col_idx, unsigned long long* key, hipStream_t s);
size_t scan_tile_count(int rows);
hipError_t launch_fill_bin(const BinSpec& b, const int* rows, int nrows,
                           const DevCsr& A, const DevCsr& B, RowStats st,
                           const int* c_rp, int* c_ci, const GlobalSlots& gs,
                           hipStream_t s);

// --- numeric ------------------------------------------------------------
// max_row: the largest C-row length the bin admits (its bound); KIND_BLOCK
// sizes its tile grid from it (tiles of b.cap positions).
// Split table of B's long rows (launch_split_*). A row k with at least
// kSplitMin entries and columns lo..hi gets nb = clamp(len / kSplitPer, 8,
// 256) granules of 2^g columns (the smallest g with (hi - lo) >> g < nb).
// Its entries at sp[off[k]]: lo, g | nb << 5, then for j = 1..nb the first
// position of B_k whose column is >= lo + (j << g) (padded to an even
// count). Rows without a table have off[k] == off[k + 1]. KIND_BLOCK narrows
// long B rows through it: a search then covers one granule.
struct SplitTable {
  const int* off = nullptr;  // B.rows + 1, or null: no table
  const int* sp = nullptr;
};
constexpr int kSplitMin = 64;
constexpr int kSplitPer = 8;
__host__ __device__ inline int split_granules(int len) {
  const int nb = len / kSplitPer;
  return nb < 8 ? 8 : (nb > 256 ? 256 : nb);
}
// cnt (B.rows ints): table entries per row; scan it with launch_scan into
// off, then launch_split_build fills sp.
hipError_t launch_split_counts(const DevCsr& B, int* cnt, hipStream_t s);
hipError_t launch_split_build(const DevCsr& B, const SplitTable& t, int* sp, hipStream_t s);

hipError_t launch_numeric_bin(const BinSpec& b, const int* rows, int nrows,
                              const DevCsr& A, const DevCsr& B, RowStats st,
                              const int* c_rp, const int* c_ci, float* c_val,
                              const GlobalSlots& gs, int max_row, const SplitTable& split,
                              hipStream_t s);

// Dynamic LDS of the tiled long-row numeric kernel for tiles of T positions.
int tiled_lds_bytes(int T);

// Long-row bins of several descriptors in one tiled launch (a batch's
// groups): per segment its row list and count, the first fused workgroup
// row it owns, its A row pointers and its C arrays. Passed by value.
struct TiledSeg {
  const int* rows;
  const int* a_rp;
  const int* c_rp;
  const int* c_ci;
  float* c_val;
  int nrows;
  int start;
};
constexpr int kMaxFusedSegs = 32;
struct TiledSegs {
  int n = 0;
  int total = 0;
  TiledSeg seg[kMaxFusedSegs];
};
// One launch over every segment (tiles of T positions, max_row the longest
// row any segment's bin admits); A's columns and values and B are shared.
hipError_t launch_numeric_tiled_fused(const TiledSegs& segs, int T, int max_row, const DevCsr& A,
                                      const DevCsr& B, const SplitTable& split, int lds_bytes,
                                      hipStream_t s);
// Default tile of the long-row numeric kernel: three workgroups fit a CU.
constexpr int kTiledPositions = 6144;
// Runs (workgroups) a long row is cut into at most (more runs measured
// slower on R-MAT rows: every run walks the row's A entries once per tile).
constexpr int kTiledRuns = 16;
// Dynamic LDS of the wave-tile long-row numeric kernel (KIND_BLOCK variant
// 1) for tiles of Tw positions (a power of two, 256..2048); a workgroup is
// one wave, and a row gets up to kWaveRuns runs of tiles.
int wave_tile_lds_bytes(int Tw);
constexpr int kWaveRuns = 64;
constexpr int kWavePositions = 1024;

// --- position map (numeric structure cache) -------------------------------
struct MapView {
  const void* map = nullptr;        // uint8 or uint16 entries
  const long long* poff = nullptr;  // rows + 1 entries
  int width = 0;                    // bytes per entry (1 or 2); 0 = no map
  int row_base = -1;  // >= 0: the listed rows are row_base, row_base + 1, ...
  // Optional B-row cache (null: none): for A's entry e, slot e - A.rp[0]
  // holds the first position (brow) and length (blen) of the B row that
  // A's column at e names, so the map kernels read no column of A and no
  // row pointer of B.
  const int* brow = nullptr;
  const unsigned short* blen = nullptr;
};
// Fills the B-row cache of A's nnz entries from A.rp[0] on (lengths capped
// at 65535; only rows whose product count fits 16 bits may use it).
hipError_t launch_brow_cache(const DevCsr& A, const DevCsr& B, int* brow, unsigned short* blen,
                             hipStream_t s);
// poff = exclusive scan over rows of (bin_of[i] in bin_mask ? Op_i : 0);
// *total = poff[rows]. `weights` holds rows int64 values of scratch.
// span_bound: weights min(Op_i, span_i) instead of Op_i.
hipError_t launch_map_offsets(int rows, const unsigned char* bin_of, unsigned bin_mask,
                              RowStats st, long long* weights, long long* poff,
                              long long* tile_sums, long long* total, hipStream_t s,
                              bool span_bound = false);
hipError_t launch_build_map(int width, const int* rows, int nrows, const DevCsr& A,
                            const DevCsr& B, const int* c_rp, const int* c_ci,
                            const long long* poff, void* map, hipStream_t s);
hipError_t launch_numeric_map_bin(const BinSpec& b, const MapView& mv, const int* rows,
                                  int nrows, const DevCsr& A, const DevCsr& B,
                                  const int* c_rp, float* c_val, hipStream_t s);

// --- persistent row-stream numeric pass (stream.hip) ------------------------
// Numeric bins of one or several descriptors (a batch's groups) served by one
// persistent launch. A segment is one bin: its rows in list order, whose
// products form one stream (soff[i] = products of list positions before i).
// The stream is cut into queue items of kStreamItem products; an item owns
// the rows whose first product lies in it (item_first[k] = first such list
// position, item_first[items] = nrows). Each wave takes items from a queue
// and walks its rows' products 64 at a time across A entries. A map segment
// reads every product's output position from `map` (one 16-bit entry per
// product of the stream; bits 0-10 the position in the C row, bit 11 set on
// the row's first product, bits 12-15 the product's order among earlier
// products of the same C entry in the same 64-product window, 15 = apply in
// lane order); an identity segment (rows fed by one B row, map null) needs
// no map. Rows of map segments hold at most kStreamMaxRow entries.
struct StreamSeg {
  const int* rows = nullptr;  // list position -> row, or null: row_base + i
  const int* a_rp = nullptr;
  const int* c_rp = nullptr;
  const int* c_ci = nullptr;  // read when building the map
  float* c_val = nullptr;
  const long long* soff = nullptr;  // nrows + 1
  const int* item_first = nullptr;  // items + 1
  unsigned short* map = nullptr;    // null: identity segment
  long long total = 0;              // products of the stream (map entries)
  int item0 = 0;                    // first queue item of the segment
  int nrows = 0;
  int row_base = 0;
  int c_nnz = 0;                    // entries of the C arrays the rows index
  int ident = 0;                    // rows fed by one B row: positions need no map
};
constexpr int kStreamMapPass = 0;
constexpr int kStreamBuildPass = 1;
constexpr int kStreamSearchPass = 2;
// Work queue of one launch; next and done are zero before the first launch,
// and every launch leaves them zero again (its last wave resets them).
// `bad` collects kStreamBad* bits of writes the kernels refused because
// their index fell outside the arrays (never set when segments and map
// match the descriptors' structure).
// The items are split into kStreamQueues contiguous ranges, one per XCD
// (workgroup w runs on XCD w mod 8): a wave takes items of its own XCD's
// range first, so neighbouring rows (one pair's, reading one B) share that
// XCD's L2, and then helps the other ranges.
constexpr int kStreamQueues = 8;
struct StreamQueue {
  unsigned next[kStreamQueues];
  unsigned done;
  unsigned bad;
};
constexpr unsigned kStreamBadValue = 1u;   // C value outside the segment's C arrays
constexpr unsigned kStreamBadMap = 2u;     // map entry outside the stream
constexpr unsigned kStreamBadLds = 4u;     // accumulator outside the sub-batch
constexpr unsigned kStreamBadColumn = 8u;  // product column not in its C row
constexpr int kStreamItem = 8192;
// Stacked-B entries the passes serve: they address B's columns and values
// by 32-bit byte offsets.
constexpr long long kStreamMaxB = 1LL << 30;
// Longest rows (products) the passes serve; longer ones stay with the
// long-row kernels.
constexpr int kStreamMaxRow = 4096;
// Accumulators per wave of the numeric pass; longer rows are walked once
// per range of this many positions.
constexpr int kStreamNumericRow = 1024;
// The same for the search pass (accumulators, columns and order tags).
constexpr int kStreamSearchRow = 512;
constexpr int kStreamMaxSegs = 256;
// Map entry (16 bits): position in the C row (bits 0-11, < kStreamMaxRow),
// row mark (bit 12: the row's first product), order among the window's
// earlier products of the same C entry (bits 13-15).
constexpr int kStreamPosBits = 12;
static_assert((1 << kStreamPosBits) >= kStreamMaxRow, "map positions must cover the rows");
// w[i] = ops[row of list position i].
hipError_t launch_stream_weights(const int* rows, int row_base, int nrows, const long long* ops,
                                 long long* w, hipStream_t s);
// Exclusive scan of n int64 values into out[0..n] (out[n] = *total);
// tile_sums holds scan_tile_count(n) values.
hipError_t launch_scan64(const long long* in, int n, long long* out, long long* tile_sums,
                         long long* total, hipStream_t s);
hipError_t launch_stream_items(const long long* soff, int nrows, int items, int* item_first,
                               hipStream_t s);
// flags[row] = 1 for the listed rows (rows null: row_base + i).
hipError_t launch_stream_mark(const int* rows, int row_base, int nrows, int* flags, hipStream_t s);
// list[pos[i]] = i for rows i < M with flags[i] (pos: exclusive scan of flags).
hipError_t launch_stream_compact(const int* flags, const int* pos, int M, int* list,
                                 hipStream_t s);
// Waves of a persistent launch of the given kStream*Pass: as many as the
// device holds at once (one wave per workgroup), at most `items`.
int stream_grid(int mode, int capw, int items);
// One persistent launch of a kStream*Pass over the segments. capw: the
// longest row a sub-batch's accumulators (or columns) hold; the build needs
// room for every row of its segments; grid from stream_grid.
hipError_t launch_stream(int mode, const StreamSeg* segs, int nseg, int items, const DevCsr& A,
                         const DevCsr& B, StreamQueue* q, int capw, int grid, hipStream_t s);
// Order bits of a map segment's entries (after the build).
hipError_t launch_stream_rounds(unsigned short* map, long long total, hipStream_t s);

// One-time kernel attribute setup: allows `lds_bytes` of dynamic LDS.
hipError_t init_kernel_attributes(int lds_bytes);

// Largest dynamic LDS a workgroup may request on gfx950.
constexpr int kMaxLdsBytes = 160 * 1024;
// LDS kept back from plans for the kernels' static arrays (the largest is the
// 1024-thread bitmap kernel's, about 12.3 KB).
constexpr int kStaticLdsReserve = 16 * 1024;

// A entries per row the SEARCH kernel's segment mode stages in LDS (a
// workgroup whose rows hold more A entries reads them from global memory).
constexpr int kSegAPerRow = 6;

// Bytes of map entries the MAP kernel's segment mode stages in LDS: per row
// the bin's accumulator capacity, within 64..256 bytes, and at most 8 KB per
// workgroup (a segment whose entries need more reads them from global
// memory). 1-byte maps hold twice the entries of 2-byte ones.
__host__ __device__ constexpr int map_stage_cap_bytes(int groups, int cap) {
  const int per_row = cap < 64 ? 64 : (cap > 256 ? 256 : cap);
  return groups * per_row < 8192 ? groups * per_row : 8192;
}
// LDS bytes of a segment's staged map entries, with room for the 16-byte
// alignment of the staging loads.
__host__ __device__ constexpr int map_stage_bytes(int groups, int cap) {
  return (map_stage_cap_bytes(groups, cap) + 32 + 15) / 16 * 16;
}
// A entries per row whose scale and B-row bounds the MAP kernel's variant 1
// stages in LDS (12 bytes each).
__host__ __device__ constexpr int map_stage_a_per_row(int cap) {
  return cap / 16 < kSegAPerRow ? kSegAPerRow : (cap / 16 > 16 ? 16 : cap / 16);
}

}  // namespace rvsp
