// kernels.hip — gfx950 kernels of the CSR SpGEMM C = A * B (FP32).
//
// Row-wise Gustavson, as in the reference (src/kernels/spgemm/
// gustavson_core_f32.inc:18-64): output row i is the union of the B rows
// selected by A's row i. The GPU version splits the work into phases
// (analysis, binning, count, scan, fill, numeric) and assigns every output
// row to one accumulator kernel by its size.
//
// Bit-exactness. The reference accumulates C(i,j) as acc = acc + a*b with a
// separately rounded product and sum, starting from +0.0f, taking the A
// entries of row i in ascending position (accum_scalar_f32.c:67-70 under
// -std=c11). Inside one B row every column appears once (canonical CSR), so
// the products of one A entry never collide. Every numeric kernel here walks
// A's row in order and lets lanes split only the B row of the current A
// entry, or splits the output columns between waves; the additions into any
// one C(i,j) therefore happen in the reference's order and the result is
// bit-identical, whatever kernel a row is assigned to.
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace rvsp {

// ---------------------------------------------------------------------------
// Small helpers.

// Lane `l` of a G-lane group visits, in lockstep with the rest of the group,
// every (A entry, B entry) product of A's row segment [a_s, a_e) and calls
// f(col) for the B column of each product it owns. A entries are taken G at
// a time; their B-row lengths are prefix-summed across the group and the
// flattened products are dealt out round-robin (load-balanced search), so a
// group stays busy whatever the mix of long and short B rows. Order of the
// calls is unspecified: only for order-free work (count, fill).
template <int G, class F>
__device__ __forceinline__ void group_for_each_col(int a_s, int a_e,
                                                   const int* __restrict__ a_ci,
                                                   const int* __restrict__ b_rp,
                                                   const int* __restrict__ b_ci,
                                                   F&& f) {
  const int l = threadIdx.x & (G - 1);
  for (int base = a_s; base < a_e; base += G) {
    const int ap = base + l;
    int bs = 0, len = 0;
    if (ap < a_e) {
      const int k = a_ci[ap];
      bs = b_rp[k];
      len = b_rp[k + 1] - bs;
    }
    const int incl = group_incl_scan<G>(len);
    const int total = __shfl(incl, G - 1, G);
    const int excl = incl - len;
    for (int p0 = 0; p0 < total; p0 += G) {
      const int p = p0 + l;
      // Owner: the last lane whose exclusive offset is <= p.
      int j = 0;
#pragma unroll
      for (int step = G >> 1; step > 0; step >>= 1) {
        const int e = __shfl(excl, j + step, G);
        if (e <= p) j += step;
      }
      const int oe = __shfl(excl, j, G);
      const int ob = __shfl(bs, j, G);
      if (p < total) f(b_ci[ob + (p - oe)]);
    }
  }
}

// The last entry e >= j of an N-long nondecreasing prefix array with
// pre[e] <= p, given pre[j] <= p: gallop forward, then halve.
template <int N = kBlock>
__device__ __forceinline__ int advance_entry(const int* pre, int j, int p) {
  int step = 1;
  while (j + step < N && pre[j + step] <= p) {
    j += step;
    step <<= 1;
  }
  for (step >>= 1; step > 0; step >>= 1)
    if (j + step < N && pre[j + step] <= p) j += step;
  return j;
}

// Whole-workgroup version of group_for_each_col (blockDim.x == NT).
// `sh_excl` holds NT + kBlockPad ints, `sh_bs` and `etab` NT ints each
// (etab: the entry of every (1 << shift)-th product) and `tmp` NT / kWave
// ints. Must be reached by every thread of the block.
constexpr int kBlockPad = 8;      // prefixes probed past an entry-table hit
constexpr int kBlockStreams = 4;  // products in flight per lane
template <int NT, class F>
__device__ __forceinline__ void block_for_each_col(int a_s, int a_e,
                                                   const int* __restrict__ a_ci,
                                                   const int* __restrict__ b_rp,
                                                   const int* __restrict__ b_ci,
                                                   int* sh_excl, int* sh_bs, int* etab,
                                                   int* tmp, F&& f, int len_cap = INT_MAX) {
  const int t = threadIdx.x;
  if (t < kBlockPad) sh_excl[NT + t] = INT_MAX;  // probes past the chunk end
  for (int base = a_s; base < a_e; base += NT) {
    const int ap = base + t;
    int bs = 0, len = 0;
    if (ap < a_e) {
      const int k = a_ci[ap];
      bs = b_rp[k];
      len = b_rp[k + 1] - bs;
      if (len >= len_cap) len = 0;  // served elsewhere
    }
    int total;
    const int excl = block_excl_scan<int, NT>(len, tmp, &total);
    sh_excl[t] = excl;
    sh_bs[t] = bs;
    // Entry table: slot m holds the entry of product m << shift.
    int shift = 5;
    while ((total >> shift) >= NT) ++shift;
    for (int m = (excl + (1 << shift) - 1) >> shift; (m << shift) < excl + len; ++m) etab[m] = t;
    __syncthreads();
    // kBlockStreams interleaved streams per lane (products t + (S k + s) NT),
    // each product's entry found independently (table slot, then the next
    // kBlockPad prefixes at once), so all their gathers are in flight
    // together. The loop is uniform over the block (f sees every lane, with
    // a flag for lanes past the end), so f may use cross-lane operations.
    for (int p0 = 0; p0 < total; p0 += kBlockStreams * NT) {
      int c[kBlockStreams];
#pragma unroll
      for (int st = 0; st < kBlockStreams; ++st) {
        const int p = p0 + st * NT + t;
        c[st] = -1;
        if (p < total) {
          int e = etab[p >> shift];
          int n = 0;
#pragma unroll
          for (int j = 1; j <= kBlockPad; ++j) n += sh_excl[e + j] <= p ? 1 : 0;
          e += n;
          if (n == kBlockPad) e = advance_entry<NT>(sh_excl, e, p);
          c[st] = b_ci[sh_bs[e] + (p - sh_excl[e])];
        }
      }
#pragma unroll
      for (int st = 0; st < kBlockStreams; ++st) f(c[st], p0 + st * NT + t < total);
    }
    __syncthreads();
  }
}

// Order-free whole-workgroup product stream (the bitmap kernels: a set
// union needs no order). Per chunk of NT A entries, entries whose B row has
// at least kItemMin entries are cut into items of kItem consecutive B
// positions that the waves take in turn: a wave finds its item's entry by
// a binary search of the item offsets (one LDS address per step, shared by
// all lanes) and gathers kItem / 64 columns per lane with no per-product
// entry lookup. The short entries' products then go through
// block_for_each_col's flattened stream (the long ones count as empty
// there). f(col, valid) is called uniformly per wave. Must be reached by
// every thread of the block; ends with a barrier.
constexpr int kItemMin = 64;
constexpr int kItem = 16 * kWave;
template <int NT, class F>
__device__ __forceinline__ void block_for_each_col_items(int a_s, int a_e,
                                                         const int* __restrict__ a_ci,
                                                         const int* __restrict__ b_rp,
                                                         const int* __restrict__ b_ci,
                                                         int* sh_excl, int* sh_bs, int* etab,
                                                         int* tmp, F&& f) {
  const int t = threadIdx.x;
  const int lane = t & (kWave - 1);
  const int wid = t >> 6;
  constexpr int kWaves = NT / kWave;
  bool any_short = false;
  for (int base = a_s; base < a_e; base += NT) {
    const int ap = base + t;
    int bs = 0, len = 0;
    if (ap < a_e) {
      const int k = a_ci[ap];
      bs = b_rp[k];
      len = b_rp[k + 1] - bs;
    }
    const bool lng = len >= kItemMin;
    any_short |= len > 0 && !lng;
    const int items = lng ? (len + kItem - 1) / kItem : 0;
    int total;
    const int off = block_excl_scan<int, NT>(items, tmp, &total);
    sh_excl[t] = off;
    sh_bs[t] = bs;
    etab[t] = len;
    __syncthreads();
    // This wave's items wid, wid + kWaves, ...: each one's gathers are in
    // flight while the previous one is applied. The gathers are
    // unconditional (inside the entry's B row; a fetch past the wave's last
    // item reads that item again and applies nothing), and every gathered
    // column is used on every path, so the compiler waits only for the item
    // being applied.
    const int nmine = total > wid ? (total - wid + kWaves - 1) / kWaves : 0;  // wave-uniform
    if (nmine > 0) {
      struct Item {
        int c[kItem / kWave];
        int lim;  // valid positions of the item
      };
      auto fetch = [&](int j) {
        Item x;
        const int it = wid + min(j, nmine - 1) * kWaves;
        // The entry of item `it`: the last e with sh_excl[e] <= it.
        int lo = 0, n = NT;
        while (n > 1) {
          const int half = n >> 1;
          lo = sh_excl[lo + half] <= it ? lo + half : lo;
          n -= half;
        }
        const int eb = sh_bs[lo], el = etab[lo];  // el >= kItemMin
        const int o0 = (it - sh_excl[lo]) * kItem;
        x.lim = j < nmine ? el - o0 : 0;
#pragma unroll
        for (int u = 0; u < kItem / kWave; ++u) x.c[u] = b_ci[eb + min(o0 + u * kWave + lane, el - 1)];
        return x;
      };
      auto apply = [&](const Item& x) {
#pragma unroll
        for (int u = 0; u < kItem / kWave; ++u) {
          const bool v = u * kWave + lane < x.lim;
          f(v ? x.c[u] : -1, v);
        }
      };
      Item xa = fetch(0);
      Item xb;
      int j = 0;
      for (; j + 2 <= nmine; j += 2) {
        xb = fetch(j + 1);
        apply(xa);
        xa = fetch(j + 2);
        apply(xb);
      }
      if (j < nmine) apply(xa);
    }
    __syncthreads();
  }
  // Short entries, if the row has any: the flattened stream.
  if (__syncthreads_or(any_short)) {
    block_for_each_col<NT>(a_s, a_e, a_ci, b_rp, b_ci, sh_excl, sh_bs, etab, tmp,
                           std::forward<F>(f), kItemMin);
  }
}

__device__ __forceinline__ int pick_bin(const BinRule& r, int metric, int span) {
  for (int b = 2; b < r.nbins; ++b)
    if (metric <= r.bound[b] && span <= r.span_max[b]) return b;
  return r.nbins - 1;
}

// ---------------------------------------------------------------------------
// Analysis: Op_i, span and longest B row per output row.

constexpr int kLongA = 64;  // A rows longer than this are analysed by a wave

__global__ __launch_bounds__(kBlock) void k_analyze(
    int M, const int* __restrict__ a_rp, const int* __restrict__ a_ci,
    const int* __restrict__ b_rp, const int* __restrict__ b_ci, RowStats st,
    int* __restrict__ long_list, int* __restrict__ n_long) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int s = a_rp[i], e = a_rp[i + 1];
  if (e - s > kLongA) {
    long_list[atomicAdd(n_long, 1)] = i;
    return;
  }
  long long o = 0;
  int lo = INT_MAX, hi = -1, mb = 0;
  for (int ap = s; ap < e; ++ap) {
    const int k = a_ci[ap];
    const int bs = b_rp[k], be = b_rp[k + 1];
    const int len = be - bs;
    if (len > 0) {
      o += len;
      lo = min(lo, b_ci[bs]);
      hi = max(hi, b_ci[be - 1]);
      mb = max(mb, len);
    }
  }
  st.ops[i] = o;
  st.lo[i] = lo;
  st.hi[i] = hi;
  st.maxb[i] = mb;
}

__global__ __launch_bounds__(kBlock) void k_analyze_long(
    const int* __restrict__ a_rp, const int* __restrict__ a_ci,
    const int* __restrict__ b_rp, const int* __restrict__ b_ci, RowStats st,
    const int* __restrict__ long_list, const int* __restrict__ n_long) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nwaves = (gridDim.x * blockDim.x) / kWave;
  const int n = *n_long;
  for (int t = wave; t < n; t += nwaves) {
    const int i = long_list[t];
    const int s = a_rp[i], e = a_rp[i + 1];
    long long o = 0;
    int lo = INT_MAX, hi = -1, mb = 0;
    for (int ap = s + lane; ap < e; ap += kWave) {
      const int k = a_ci[ap];
      const int bs = b_rp[k], be = b_rp[k + 1];
      const int len = be - bs;
      if (len > 0) {
        o += len;
        lo = min(lo, b_ci[bs]);
        hi = max(hi, b_ci[be - 1]);
        mb = max(mb, len);
      }
    }
    o = group_sum<kWave>(o);
    lo = group_min<kWave>(lo);
    hi = group_max<kWave>(hi);
    mb = group_max<kWave>(mb);
    if (lane == 0) {
      st.ops[i] = o;
      st.lo[i] = lo;
      st.hi[i] = hi;
      st.maxb[i] = mb;
    }
  }
}

hipError_t launch_analyze(const DevCsr& A, const DevCsr& B, RowStats st,
                          int* long_list, int* n_long, hipStream_t s) {
  if (A.rows <= 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(n_long, 0, sizeof(int), s);
  if (e != hipSuccess) return e;
  const int grid = (A.rows + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_analyze, dim3(grid), dim3(kBlock), 0, s, A.rows, A.rp,
                     A.ci, B.rp, B.ci, st, long_list, n_long);
  hipLaunchKernelGGL(k_analyze_long, dim3(1024), dim3(kBlock), 0, s, A.rp,
                     A.ci, B.rp, B.ci, st, long_list, n_long);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Classification.

__global__ __launch_bounds__(kBlock) void k_classify_count(
    int M, const int* __restrict__ a_rp, RowStats st, BinRule r,
    unsigned char* __restrict__ bin_of, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const long long o = st.ops[i];
  int b = 0;
  if (o == 0) {
    cnt[i] = 0;
  } else if ((long long)st.maxb[i] == o) {
    // Every product comes from one B row, whose columns are distinct.
    cnt[i] = (int)o;
  } else {
    const int span = st.hi[i] - st.lo[i] + 1;
    const int ub = (int)min(o, (long long)span);
    b = pick_bin(r, ub, span);
  }
  bin_of[i] = (unsigned char)b;
}

__global__ __launch_bounds__(kBlock) void k_classify_nnz(
    int M, RowStats st, const int* __restrict__ c_rp, BinRule r,
    unsigned char* __restrict__ bin_of, const long long* __restrict__ kept) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int n = c_rp[i + 1] - c_rp[i];
  int b;
  if (n == 0 || (kept && kept[i + 1] > kept[i]))
    b = 0;  // nothing to do, or columns kept by the count phase
  else if ((long long)st.maxb[i] == st.ops[i])
    b = 1;
  else
    b = pick_bin(r, n, st.hi[i] - st.lo[i] + 1);
  bin_of[i] = (unsigned char)b;
}

hipError_t launch_classify_count(const DevCsr& A, RowStats st, const BinRule& r,
                                 unsigned char* bin_of, int* cnt, hipStream_t s) {
  const int grid = (A.rows + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_classify_count, dim3(grid), dim3(kBlock), 0, s, A.rows,
                     A.rp, st, r, bin_of, cnt);
  return hipGetLastError();
}

hipError_t launch_classify_nnz(const DevCsr& A, RowStats st, const int* c_rp,
                               const BinRule& r, unsigned char* bin_of,
                               hipStream_t s, const long long* kept) {
  const int grid = (A.rows + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_classify_nnz, dim3(grid), dim3(kBlock), 0, s, A.rows, st,
                     c_rp, r, bin_of, kept);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable partition of rows by bin (three passes: per-tile histogram, scan of
// the histograms, ordered scatter). Rows keep ascending order inside a bin,
// so neighbouring groups work on neighbouring rows.

__global__ __launch_bounds__(kBlock) void k_part_hist(
    int M, int nb, int tile_rows, const unsigned char* __restrict__ bin_of,
    int* __restrict__ tile_hist) {
  __shared__ int h[kMaxBins];
  const int lane = threadIdx.x & (kWave - 1);
  if (threadIdx.x < kMaxBins) h[threadIdx.x] = 0;
  __syncthreads();
  const int t0 = blockIdx.x * tile_rows;
  const int t1 = min(M, t0 + tile_rows);
  for (int c = t0; c < t1; c += kBlock) {
    const int i = c + threadIdx.x;
    const int b = (i < t1) ? (int)bin_of[i] : -1;
    for (int bb = 0; bb < nb; ++bb) {
      const unsigned long long m = __ballot(b == bb);
      if (lane == 0 && m) atomicAdd(&h[bb], __popcll(m));
    }
  }
  __syncthreads();
  if (threadIdx.x < kMaxBins)
    tile_hist[blockIdx.x * kMaxBins + threadIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_part_scan(
    int ntiles, int nb, const int* __restrict__ tile_hist,
    int* __restrict__ tile_base, int* __restrict__ bin_offsets) {
  __shared__ int tmp[kBlock / kWave];
  __shared__ int tot[kMaxBins];
  for (int b = 0; b < nb; ++b) {
    int running = 0;
    for (int t0 = 0; t0 < ntiles; t0 += kBlock) {
      const int t = t0 + threadIdx.x;
      const int v = (t < ntiles) ? tile_hist[t * kMaxBins + b] : 0;
      int total;
      const int ex = block_excl_scan(v, tmp, &total);
      if (t < ntiles) tile_base[t * kMaxBins + b] = running + ex;
      running += total;
    }
    if (threadIdx.x == 0) tot[b] = running;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int b = 0; b < nb; ++b) {
      bin_offsets[b] = acc;
      acc += tot[b];
    }
    bin_offsets[nb] = acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_part_scatter(
    int M, int nb, int tile_rows, const unsigned char* __restrict__ bin_of,
    const int* __restrict__ tile_base, const int* __restrict__ bin_offsets,
    int* __restrict__ rows_out) {
  __shared__ int run[kMaxBins];
  __shared__ int wcnt[kBlock / kWave][kMaxBins];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  if (threadIdx.x < nb)
    run[threadIdx.x] =
        bin_offsets[threadIdx.x] + tile_base[blockIdx.x * kMaxBins + threadIdx.x];
  __syncthreads();
  const int t0 = blockIdx.x * tile_rows;
  const int t1 = min(M, t0 + tile_rows);
  const unsigned long long below = __lanemask_lt();
  for (int c = t0; c < t1; c += kBlock) {
    const int i = c + threadIdx.x;
    const int b = (i < t1) ? (int)bin_of[i] : -1;
    int rank = 0;
    for (int bb = 0; bb < nb; ++bb) {
      const unsigned long long m = __ballot(b == bb);
      if (b == bb) rank = __popcll(m & below);
      if (lane == 0) wcnt[wid][bb] = __popcll(m);
    }
    __syncthreads();
    if (b >= 0) {
      int pos = run[b] + rank;
      for (int w = 0; w < wid; ++w) pos += wcnt[w][b];
      rows_out[pos] = i;
    }
    __syncthreads();
    if (threadIdx.x < nb) {
      int add = 0;
      for (int w = 0; w < kBlock / kWave; ++w) add += wcnt[w][threadIdx.x];
      run[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

hipError_t launch_partition(int rows, int nbins, PartitionScratch& ps,
                            int* rows_out, hipStream_t s) {
  if (rows <= 0) return hipMemsetAsync(ps.bin_offsets, 0,
                                       (kMaxBins + 1) * sizeof(int), s);
  hipLaunchKernelGGL(k_part_hist, dim3(ps.ntiles), dim3(kBlock), 0, s, rows,
                     nbins, ps.tile_rows, ps.bin_of, ps.tile_hist);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(kBlock), 0, s, ps.ntiles,
                     nbins, ps.tile_hist, ps.tile_base, ps.bin_offsets);
  hipLaunchKernelGGL(k_part_scatter, dim3(ps.ntiles), dim3(kBlock), 0, s,
                     rows, nbins, ps.tile_rows, ps.bin_of, ps.tile_base,
                     ps.bin_offsets, rows_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-bin statistics of a partitioned phase (features of the tuned and learned
// selectors). out[b * kBinStats + s], s = rows, sum nnz(A row), sum Op_i,
// sum size metric, max size metric, sum span, max longest-B-row, max Op_i.
// The size metric is min(Op_i, span) when c_rp is null (count phase), else
// the C row's nnz.

__global__ __launch_bounds__(kBlock) void k_bin_stats(
    const int* __restrict__ rows, const int* __restrict__ offsets, int nbins,
    const int* __restrict__ a_rp, RowStats st, const int* __restrict__ c_rp,
    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long acc[kMaxBins * kBinStats];
  __shared__ int off[kMaxBins + 1];
  for (int i = threadIdx.x; i < kMaxBins * kBinStats; i += kBlock) acc[i] = 0;
  if (threadIdx.x <= nbins) off[threadIdx.x] = offsets[threadIdx.x];
  __syncthreads();
  const int total = off[nbins];
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < total; p += gridDim.x * kBlock) {
    int b = 0;
    while (b + 1 < nbins && off[b + 1] <= p) ++b;
    const int row = rows[p];
    const long long ops = st.ops[row];
    const int span = ops > 0 ? st.hi[row] - st.lo[row] + 1 : 0;
    const long long metric = c_rp ? (long long)(c_rp[row + 1] - c_rp[row])
                                  : min(ops, (long long)span);
    unsigned long long* a = acc + b * kBinStats;
    atomicAdd(&a[0], 1ull);
    atomicAdd(&a[1], (unsigned long long)(a_rp[row + 1] - a_rp[row]));
    atomicAdd(&a[2], (unsigned long long)ops);
    atomicAdd(&a[3], (unsigned long long)metric);
    atomicMax(&a[4], (unsigned long long)metric);
    atomicAdd(&a[5], (unsigned long long)span);
    atomicMax(&a[6], (unsigned long long)st.maxb[row]);
    atomicMax(&a[7], (unsigned long long)ops);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins * kBinStats; i += kBlock) {
    if (acc[i] == 0) continue;
    const int s = i % kBinStats;
    if (s == 4 || s == 6 || s == 7)
      atomicMax(&out[i], acc[i]);
    else
      atomicAdd(&out[i], acc[i]);
  }
}

hipError_t launch_bin_stats(const int* rows, const int* offsets, int nbins, int total_rows,
                            const int* a_rp, RowStats st, const int* c_rp,
                            unsigned long long* out, hipStream_t s) {
  hipError_t e = hipMemsetAsync(out, 0, sizeof(unsigned long long) * kMaxBins * kBinStats, s);
  if (e != hipSuccess || total_rows <= 0) return e;
  int grid = (total_rows + kBlock * 8 - 1) / (kBlock * 8);
  grid = grid < 1024 ? grid : 1024;
  hipLaunchKernelGGL(k_bin_stats, dim3(grid), dim3(kBlock), 0, s, rows, offsets, nbins,
                     a_rp, st, c_rp, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Symbolic count (reference: rvsp_symbolic.c:13-88, distinct columns of the
// union of B rows).

// LDS open-addressing hash of 2^log2T keys per row, G lanes per row.
template <int G>
__global__ __launch_bounds__(kBlock) void k_count_hash(
    const int* __restrict__ rows, int nrows, int log2T, DevCsr A, DevCsr B,
    int* __restrict__ cnt) {
  extern __shared__ int lds[];
  constexpr int kGroups = kBlock / G;
  const int T = 1 << log2T;
  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  const int r = blockIdx.x * kGroups + g;
  if (r >= nrows) return;  // uniform per group; no block barriers below
  int* keys = lds + g * T;
  const int row = rows[r];
  for (int j = l; j < T; j += G) keys[j] = kEmptyKey;
  int mine = 0;
  group_for_each_col<G>(A.rp[row], A.rp[row + 1], A.ci, B.rp, B.ci,
                        [&](int col) {
                          int slot = hash_slot(col, log2T);
                          while (true) {
                            const int old = atomicCAS(&keys[slot], kEmptyKey, col);
                            if (old == kEmptyKey) { ++mine; break; }
                            if (old == col) break;
                            slot = (slot + 1) & (T - 1);
                          }
                        });
  mine = group_sum<G>(mine);
  if (l == 0) cnt[row] = mine;
}

// k_bitmap_row modes: count only; write the sorted columns to C; count and
// also keep the sorted columns at koff[row] of a side buffer, so the fill
// phase copies them instead of building the bitmap again.
constexpr int kBitmapCount = 0, kBitmapFill = 1, kBitmapCountKeep = 2;
constexpr int kWideBlock = 1024;              // threads of a workgroup holding a large bitmap

// One workgroup per row. The row's columns lie in [lo, hi]; a bitmap over that
// span lives in LDS (GLOBAL_BM == false, span <= 32 * words) or in this
// workgroup's slot of global scratch (GLOBAL_BM == true, absolute column
// bits, words cleared again after use). FILL == false counts set bits;
// FILL == true writes the set columns in ascending order to C.
template <bool GLOBAL_BM, int MODE, int NT = kBlock>
__global__ __launch_bounds__(NT) void k_bitmap_row(
    const int* __restrict__ rows, int nrows, DevCsr A, DevCsr B, RowStats st,
    int* __restrict__ cnt, const int* __restrict__ c_rp, int* __restrict__ c_ci,
    unsigned* __restrict__ gbits, size_t slot_words, int* __restrict__ keep,
    const long long* __restrict__ koff) {
  constexpr bool FILL = MODE == kBitmapFill;
  constexpr int kWaves = NT / kWave;
  extern __shared__ unsigned lbits[];
  __shared__ int sh_excl[NT + kBlockPad], sh_bs[NT], sh_etab[NT], tmp[kWaves];
  unsigned* bits = GLOBAL_BM ? gbits + (size_t)blockIdx.x * slot_words : lbits;
  for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
    const int row = rows[r];
    const int lo = st.lo[row], hi = st.hi[row];
    // LDS bitmaps are offset by lo; global ones are absolute.
    const int w0 = GLOBAL_BM ? (lo >> 5) : 0;
    const int w1 = GLOBAL_BM ? (hi >> 5) + 1 : ((hi - lo) >> 5) + 1;
    const int base_col = GLOBAL_BM ? 0 : lo;
    if (!GLOBAL_BM) {
      for (int w = threadIdx.x; w < w1; w += NT) bits[w] = 0u;
      __syncthreads();
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x >> 6;
    block_for_each_col_items<NT>(A.rp[row], A.rp[row + 1], A.ci, B.rp, B.ci, sh_excl,
                       sh_bs, sh_etab, tmp, [&](int col, bool valid) {
                         if (valid) {
                           // Bits are only ever set: a bit seen set needs no
                           // atomic, which spares the serialised atomics
                           // of the many products that land on columns
                           // already seen (hub columns in every B row).
                           const int d = col - base_col;
                           unsigned* w = &bits[d >> 5];
                           const unsigned m = 1u << (d & 31);
                           if (!(__hip_atomic_load(w, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP) & m))
                             __hip_atomic_fetch_or(w, m, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                         }
                       });
    // block_for_each_col_items ends with a barrier, which also orders the global
    // atomics above before the reads below (workgroup scope).
    // The span's words are split into one part per wave; a wave reads its
    // part 64 consecutive words at a time (no bank conflicts).
    const int nw = w1 - w0;
    const int q0 = w0 + (int)((long long)nw * wid / kWaves);
    const int q1 = w0 + (int)((long long)nw * (wid + 1) / kWaves);
    int c = 0;
    for (int w = q0 + lane; w < q1; w += kWave) {
      c += __popc(bits[w]);
      if (GLOBAL_BM && !FILL) bits[w] = 0u;
    }
    c = group_sum<kWave>(c);
    if (lane == 0) tmp[wid] = c;
    __syncthreads();
    int total = 0, base = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      total += tmp[v];
      base += v < wid ? tmp[v] : 0;
    }
    if (MODE != kBitmapFill && threadIdx.x == 0) cnt[row] = total;
    if (MODE != kBitmapCount) {
      // Ascending output: a wave's 64 words at a time, each lane's columns
      // after those of the lanes below it (a wave prefix sum of popcounts).
      int* out = FILL ? c_ci + c_rp[row] : keep + koff[row];
      for (int wb = q0; wb < q1; wb += kWave) {
        const int w = wb + lane;
        unsigned m = w < q1 ? bits[w] : 0u;
        if (GLOBAL_BM && w < q1) bits[w] = 0u;
        const int pc = __popc(m);
        const int incl = group_incl_scan<kWave>(pc);
        int off = base + incl - pc;
        while (m) {
          const int b = __ffs(m) - 1;
          out[off++] = base_col + w * 32 + b;
          m &= m - 1;
        }
        base += __shfl(incl, kWave - 1, kWave);
      }
    }
    __syncthreads();
  }
}

template <int G>
static hipError_t launch_count_hash_g(const BinSpec& b, const int* rows,
                                      int nrows, const DevCsr& A,
                                      const DevCsr& B, int* cnt, hipStream_t s) {
  const int groups = kBlock / G;
  const int grid = (nrows + groups - 1) / groups;
  int log2T = 0;
  while ((1 << log2T) < b.cap) ++log2T;
  hipLaunchKernelGGL(k_count_hash<G>, dim3(grid), dim3(kBlock), b.lds_bytes, s,
                     rows, nrows, log2T, A, B, cnt);
  return hipGetLastError();
}

static int row_grid(int nrows) { return nrows < 65535 ? nrows : 65535; }

hipError_t launch_count_bin(const BinSpec& b, const int* rows, int nrows,
                            const DevCsr& A, const DevCsr& B, RowStats st,
                            int* cnt, const GlobalSlots& gs, hipStream_t s, int* keep,
                            const long long* koff) {
  if (nrows <= 0) return hipSuccess;
  switch (b.kind) {
    case KIND_HASH:
      switch (b.group) {
        case 1: return launch_count_hash_g<1>(b, rows, nrows, A, B, cnt, s);
        case 2: return launch_count_hash_g<2>(b, rows, nrows, A, B, cnt, s);
        case 4: return launch_count_hash_g<4>(b, rows, nrows, A, B, cnt, s);
        case 8: return launch_count_hash_g<8>(b, rows, nrows, A, B, cnt, s);
        case 16: return launch_count_hash_g<16>(b, rows, nrows, A, B, cnt, s);
        case 32: return launch_count_hash_g<32>(b, rows, nrows, A, B, cnt, s);
        case 64: return launch_count_hash_g<64>(b, rows, nrows, A, B, cnt, s);
        default: return hipErrorInvalidValue;
      }
    case KIND_BITMAP:
      // A bitmap that leaves room for one workgroup per CU gets 16 waves.
      if (b.lds_bytes > kWideBitmapBytes) {
        if (keep)
          hipLaunchKernelGGL((k_bitmap_row<false, kBitmapCountKeep, kWideBlock>),
                             dim3(row_grid(nrows)), dim3(kWideBlock), b.lds_bytes, s, rows,
                             nrows, A, B, st, cnt, nullptr, nullptr, nullptr, (size_t)0, keep,
                             koff);
        else
          hipLaunchKernelGGL((k_bitmap_row<false, kBitmapCount, kWideBlock>),
                             dim3(row_grid(nrows)), dim3(kWideBlock), b.lds_bytes, s, rows,
                             nrows, A, B, st, cnt, nullptr, nullptr, nullptr, (size_t)0,
                             nullptr, nullptr);
      } else if (keep) {
        hipLaunchKernelGGL((k_bitmap_row<false, kBitmapCountKeep>), dim3(row_grid(nrows)),
                           dim3(kBlock), b.lds_bytes, s, rows, nrows, A, B, st, cnt,
                           nullptr, nullptr, nullptr, (size_t)0, keep, koff);
      } else {
        hipLaunchKernelGGL((k_bitmap_row<false, kBitmapCount>), dim3(row_grid(nrows)),
                           dim3(kBlock), b.lds_bytes, s, rows, nrows, A, B, st, cnt,
                           nullptr, nullptr, nullptr, (size_t)0, nullptr, nullptr);
      }
      return hipGetLastError();
    case KIND_GLOBAL:
      if (!gs.base || gs.nslots <= 0) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_bitmap_row<true, kBitmapCount>),
                         dim3(min(nrows, gs.nslots)), dim3(kBlock), 0, s, rows,
                         nrows, A, B, st, cnt, nullptr, nullptr,
                         (unsigned*)gs.base, gs.slot_bytes / sizeof(unsigned), nullptr,
                         nullptr);
      return hipGetLastError();
    default:
      return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Canonical-CSR check on the device (reference: rvsp_support.c:36-126). Each
// failing row pointer or entry proposes (row << 32) | code; the smallest
// proposal is the failure a scan of rows in order, entries in order, would
// stop at. An entry is judged against its predecessor in the row only: at
// the first failure every earlier entry passed, which is all the scan needs.

__global__ __launch_bounds__(kBlock) void k_csr_check_rows(
    int rows, int nnz, const int* __restrict__ rp, int has_cols,
    unsigned long long* __restrict__ key) {
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < rows; r += gridDim.x * kBlock) {
    const int s = rp[r], e = rp[r + 1];
    int code = -1;
    if (e < s)
      code = 0;
    else if (e > s && !has_cols)
      code = 1;
    else if (has_cols && e > nnz)
      code = 0;  // past the column array (without one nothing is read)
    if (code >= 0) atomicMin(key, ((unsigned long long)r << 32) | (unsigned)code);
  }
}

constexpr int kCheckTile = 4 * kBlock;  // entries per workgroup step

// Last row r in [lo, hi] with rp[r] <= p (rp[lo] <= p holds).
__device__ __forceinline__ int row_of_entry(const int* __restrict__ rp, int lo, int hi, int p) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo + 1) >> 1);
    if (rp[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void k_csr_check_cols(
    int limit, int cols, const int* __restrict__ rp, const int* __restrict__ ci,
    unsigned long long* __restrict__ key) {
  __shared__ int range[2];
  const int P = rp[limit];  // entries of rows [0, limit)
  for (long long b = (long long)blockIdx.x * kCheckTile; b < P;
       b += (long long)gridDim.x * kCheckTile) {
    const int b0 = (int)b;
    const int b1 = (int)min((long long)P, b + kCheckTile) - 1;
    // Rows of the step's first and last entries bound every search below.
    if (threadIdx.x < 2) range[threadIdx.x] = row_of_entry(rp, 0, limit - 1, threadIdx.x ? b1 : b0);
    __syncthreads();
    const int rlo = range[0], rhi = range[1];
    for (int p = b0 + threadIdx.x; p <= b1; p += kBlock) {
      const int r = row_of_entry(rp, rlo, rhi, p);
      const int s = rp[r];
      const int c = ci[p];
      bool bad = c < 0 || c >= cols;
      if (!bad && p > s) bad = c <= ci[p - 1];
      if (bad) atomicMin(key, ((unsigned long long)r << 32) | (2u + (unsigned)(p - s)));
    }
    __syncthreads();  // range is rewritten by the next step
  }
}

hipError_t launch_csr_check_rows(int rows, int nnz, const int* row_ptr, bool has_cols,
                                 unsigned long long* key, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  const int grid = std::min((rows + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_csr_check_rows, dim3(grid), dim3(kBlock), 0, s, rows, nnz, row_ptr,
                     has_cols ? 1 : 0, key);
  return hipGetLastError();
}

hipError_t launch_csr_check_cols(int limit, int cols, int nnz, const int* row_ptr,
                                 const int* col_idx, unsigned long long* key, hipStream_t s) {
  if (limit <= 0 || nnz <= 0) return hipSuccess;
  const int grid = std::min((nnz + kCheckTile - 1) / kCheckTile, 8192);
  hipLaunchKernelGGL(k_csr_check_cols, dim3(grid), dim3(kBlock), 0, s, limit, cols, row_ptr,
                     col_idx, key);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Exclusive scan of the per-row counts into C.row_ptr (reference:
// rvsp_symbolic.c:75-85). Sums are kept in 64 bits so the caller can apply
// the reference's INT32_MAX guard (rvsp_symbolic.c:69-72) to the total.

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

size_t scan_tile_count(int rows) {
  return (size_t)((rows + kScanTile - 1) / kScanTile) + 1;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(
    const int* __restrict__ cnt, int n, long long* __restrict__ tile_sums) {
  __shared__ long long tmp[kBlock / kWave];
  const int base = blockIdx.x * kScanTile;
  long long s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int i = base + j * kBlock + threadIdx.x;
    if (i < n) s += cnt[i];
  }
  s = block_sum(s, tmp);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(
    long long* __restrict__ tile_sums, int ntiles, long long* __restrict__ total) {
  __shared__ long long tmp[kBlock / kWave];
  long long running = 0;
  for (int t0 = 0; t0 < ntiles; t0 += kBlock) {
    const int t = t0 + threadIdx.x;
    const long long v = (t < ntiles) ? tile_sums[t] : 0;
    long long tot;
    const long long ex = block_excl_scan(v, tmp, &tot);
    if (t < ntiles) tile_sums[t] = running + ex;
    running += tot;
  }
  if (threadIdx.x == 0) *total = running;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(
    const int* __restrict__ cnt, int n, const long long* __restrict__ tile_off,
    int* __restrict__ c_rp) {
  __shared__ long long tmp[kBlock / kWave];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  long long s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? cnt[base + j] : 0;
    s += v[j];
  }
  long long tot;
  long long run = block_excl_scan(s, tmp, &tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) {
      c_rp[base + j] = (int)run;
      run += v[j];
      if (base + j == n - 1) c_rp[n] = (int)run;
    }
  }
}

hipError_t launch_scan(const int* cnt, int rows, int* c_rp, long long* total,
                       long long* tile_sums, hipStream_t s) {
  const int ntiles = (rows + kScanTile - 1) / kScanTile;
  if (rows <= 0) return hipMemsetAsync(total, 0, sizeof(long long), s);
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kBlock), 0, s, cnt, rows,
                     tile_sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, s, tile_sums, ntiles,
                     total);
  hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(kBlock), 0, s, cnt, rows,
                     tile_sums, c_rp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fill: sorted column indices of every C row (reference:
// rvsp_symbolic.c:91-138 with rvsp_sort.h; only the ascending result is
// observable, not the sort algorithm).

// One nonempty B row: the C row's columns are that row's columns.
template <int G>
__global__ __launch_bounds__(kBlock) void k_fill_direct(
    const int* __restrict__ rows, int nrows, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, int* __restrict__ c_ci) {
  constexpr int kGroups = kBlock / G;
  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  const int r = blockIdx.x * kGroups + g;
  if (r >= nrows) return;
  const int row = rows[r];
  int* out = c_ci + c_rp[row];
  for (int ap = A.rp[row]; ap < A.rp[row + 1]; ++ap) {
    const int k = A.ci[ap];
    const int bs = B.rp[k], be = B.rp[k + 1];
    for (int p = bs + l; p < be; p += G) out[p - bs] = B.ci[p];
  }
}

// LDS hash of 2^log2T slots plus a compact list of the distinct columns;
// each column's output position is its rank among the row's columns.
template <int G>
__global__ __launch_bounds__(kBlock) void k_fill_hash(
    const int* __restrict__ rows, int nrows, int log2T, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, int* __restrict__ c_ci) {
  extern __shared__ int lds[];
  constexpr int kGroups = kBlock / G;
  const int T = 1 << log2T;
  const int stride = T + T / 2 + 4;
  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  const int r = blockIdx.x * kGroups + g;
  if (r >= nrows) return;
  int* keys = lds + g * stride;
  int* list = keys + T;
  int* counter = list + T / 2;
  const int row = rows[r];
  for (int j = l; j < T; j += G) keys[j] = kEmptyKey;
  if (l == 0) *counter = 0;
  group_for_each_col<G>(A.rp[row], A.rp[row + 1], A.ci, B.rp, B.ci,
                        [&](int col) {
                          int slot = hash_slot(col, log2T);
                          while (true) {
                            const int old = atomicCAS(&keys[slot], kEmptyKey, col);
                            if (old == kEmptyKey) {
                              list[atomicAdd(counter, 1)] = col;
                              break;
                            }
                            if (old == col) break;
                            slot = (slot + 1) & (T - 1);
                          }
                        });
  const int c_s = c_rp[row];
  const int n = c_rp[row + 1] - c_s;
  // The group's n distinct columns, in hash order, sorted ascending by a
  // bitonic network over the list padded to a power of two (at most T / 2,
  // the list's size). A group lies within one wave, whose LDS operations run
  // in program order, so its compare-exchange steps need no barrier.
  int P = 1;
  while (P < n) P <<= 1;
  for (int j = n + l; j < P; j += G) list[j] = INT_MAX;
  wave_lds_order();
  for (int k = 2; k <= P; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = l; i < (P >> 1); i += G) {
        const int lo = ((i & ~(jj - 1)) << 1) | (i & (jj - 1));
        const int hi = lo + jj;
        const int a = list[lo], b = list[hi];
        if ((a > b) == ((lo & k) == 0)) {
          list[lo] = b;
          list[hi] = a;
        }
      }
      wave_lds_order();
    }
  }
  for (int j = l; j < n; j += G) c_ci[c_s + j] = list[j];
}

template <int G>
static hipError_t launch_fill_g(const BinSpec& b, const int* rows, int nrows,
                                const DevCsr& A, const DevCsr& B,
                                const int* c_rp, int* c_ci, hipStream_t s) {
  const int groups = kBlock / G;
  const int grid = (nrows + groups - 1) / groups;
  if (b.kind == KIND_DIRECT) {
    hipLaunchKernelGGL(k_fill_direct<G>, dim3(grid), dim3(kBlock), 0, s, rows,
                       nrows, A, B, c_rp, c_ci);
  } else {
    int log2T = 0;
    while ((1 << log2T) < b.cap) ++log2T;
    hipLaunchKernelGGL(k_fill_hash<G>, dim3(grid), dim3(kBlock), b.lds_bytes, s,
                       rows, nrows, log2T, A, B, c_rp, c_ci);
  }
  return hipGetLastError();
}

hipError_t launch_fill_bin(const BinSpec& b, const int* rows, int nrows,
                           const DevCsr& A, const DevCsr& B, RowStats st,
                           const int* c_rp, int* c_ci, const GlobalSlots& gs,
                           hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  switch (b.kind) {
    case KIND_DIRECT:
    case KIND_HASH:
      switch (b.group) {
        case 1: return launch_fill_g<1>(b, rows, nrows, A, B, c_rp, c_ci, s);
        case 2: return launch_fill_g<2>(b, rows, nrows, A, B, c_rp, c_ci, s);
        case 4: return launch_fill_g<4>(b, rows, nrows, A, B, c_rp, c_ci, s);
        case 8: return launch_fill_g<8>(b, rows, nrows, A, B, c_rp, c_ci, s);
        case 16: return launch_fill_g<16>(b, rows, nrows, A, B, c_rp, c_ci, s);
        case 32: return launch_fill_g<32>(b, rows, nrows, A, B, c_rp, c_ci, s);
        case 64: return launch_fill_g<64>(b, rows, nrows, A, B, c_rp, c_ci, s);
        default: return hipErrorInvalidValue;
      }
    case KIND_BITMAP:
      if (b.lds_bytes > kWideBitmapBytes)
        hipLaunchKernelGGL((k_bitmap_row<false, kBitmapFill, kWideBlock>), dim3(row_grid(nrows)),
                           dim3(kWideBlock), b.lds_bytes, s, rows, nrows, A, B, st,
                           nullptr, c_rp, c_ci, nullptr, (size_t)0, nullptr, nullptr);
      else
        hipLaunchKernelGGL((k_bitmap_row<false, kBitmapFill>), dim3(row_grid(nrows)),
                           dim3(kBlock), b.lds_bytes, s, rows, nrows, A, B, st,
                           nullptr, c_rp, c_ci, nullptr, (size_t)0, nullptr, nullptr);
      return hipGetLastError();
    case KIND_GLOBAL:
      if (!gs.base || gs.nslots <= 0) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_bitmap_row<true, kBitmapFill>), dim3(min(nrows, gs.nslots)),
                         dim3(kBlock), 0, s, rows, nrows, A, B, st, nullptr,
                         c_rp, c_ci, (unsigned*)gs.base,
                         gs.slot_bytes / sizeof(unsigned), nullptr, nullptr);
      return hipGetLastError();
    default:
      return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Numeric (reference: gustavson_core_f32.inc:18-64, accum_scalar_f32.c:19-71).

// One nonempty B row: C(i, :) = 0 + a * B(k, :).
// C's values leave the numeric kernels with non-temporal stores, so they do
// not push B's values (gathered by later workgroups) out of the L2: on
// 16-entry rows of C = A * A for uniform random A (4096 x 4096, 64
// block-diagonal copies) that cut the position-map kernel from 165 to 130 us.
// Shapes of at most two lanes per row keep ordinary stores, measured 1.7 %
// faster on the 1024 x 1024 Laplacian, whose B rows are reread from the L2
// anyway.
template <int G>
__device__ __forceinline__ void c_store(float v, float* dst) {
  if constexpr (G > 2) __builtin_nontemporal_store(v, dst);
  else *dst = v;
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_numeric_direct(
    const int* __restrict__ rows, int nrows, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, float* __restrict__ c_val) {
  constexpr int kGroups = kBlock / G;
  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  const int r = blockIdx.x * kGroups + g;
  if (r >= nrows) return;
  const int row = rows[r];
  float* out = c_val + c_rp[row];
  for (int ap = A.rp[row]; ap < A.rp[row + 1]; ++ap) {
    const int k = A.ci[ap];
    const int bs = B.rp[k], be = B.rp[k + 1];
    const float a = A.val[ap];
    for (int p = bs + l; p < be; p += G) c_store<G>(acc_step(0.0f, a, B.val[p]), out + (p - bs));
  }
}

// Block b is dispatched to XCD b % 8, and each XCD has its own L2. Renumber
// blocks so every XCD works through one contiguous range of logical blocks:
// neighbouring rows, which read overlapping B rows, then share an L2.
// A bijection on [0, n) for any n.
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// G lanes per row. The row's sorted C columns sit in LDS next to an FP32
// accumulator; each product finds its column by binary search. A entries are
// visited in order and the lanes of the group split the current B row only.
//
// Segment mode. When the workgroup's rows are consecutive (always for a bin
// holding a banded or uniform matrix, whose rows stay in ascending order),
// its C rows and its A rows are each one contiguous range. The workgroup
// then stages, with coalesced loads and without the G-fold redundancy of
// per-group loads: the C and A row offsets, the C columns (and zeroed
// accumulators), and for every A entry its scale and B-row bounds. The
// per-entry loop is left with only the gathers of B's columns and values,
// and the values are stored back coalesced. Otherwise every group stages
// its own row in a fixed slot of `cap` entries and reads A directly.
//
// Loads run ahead of the arithmetic: entry ap+1's products (kU per lane,
// in registers) are in flight while entry ap is applied. Only loads move;
// the additions into any C(i,j) stay in A order.
template <int G>
__global__ __launch_bounds__(kBlock) void k_numeric_search(
    const int* __restrict__ rows, int nrows, int cap, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, const int* __restrict__ c_ci,
    float* __restrict__ c_val) {
  extern __shared__ int lds[];
  constexpr int kGroups = kBlock / G;
  constexpr int kACap = kGroups * kSegAPerRow;
  constexpr int kU = 2;
  int* cols = lds;
  float* vals = reinterpret_cast<float*>(lds + kGroups * cap);
  int* seg_crp = lds + 2 * kGroups * cap;  // kGroups + 1 entries
  int* seg_arp = seg_crp + kGroups + 1;     // kGroups + 1 entries
  float* m_a = reinterpret_cast<float*>(seg_arp + kGroups + 1);  // kACap each
  int* m_bs = reinterpret_cast<int*>(m_a + kACap);
  int* m_be = m_bs + kACap;

  const int r0 = xcd_block(blockIdx.x, gridDim.x) * kGroups;
  const int nr = min(kGroups, nrows - r0);
  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  const bool active = g < nr;
  const int first = rows[r0];
  const int last = rows[r0 + nr - 1];
  const bool seg = (last - first == nr - 1);  // uniform over the block

  int base, n, e_s, e_e;  // C slot base and length; A entries [e_s, e_e)
  bool staged = false;    // A metadata in LDS, indexed from 0
  int s0 = 0, s1 = 0;
  if (seg) {
    for (int i = threadIdx.x; i <= nr; i += kBlock) {
      seg_crp[i] = c_rp[first + i];
      seg_arp[i] = A.rp[first + i];
    }
    __syncthreads();
    s0 = seg_crp[0];
    s1 = seg_crp[nr];
    const int as0 = seg_arp[0];
    const int na = seg_arp[nr] - as0;
    staged = na <= kACap;
    for (int i = threadIdx.x; i < s1 - s0; i += kBlock) {
      cols[i] = c_ci[s0 + i];
      vals[i] = 0.0f;
    }
    if (staged) {
      for (int i = threadIdx.x; i < na; i += kBlock) {
        const int k = A.ci[as0 + i];
        m_a[i] = A.val[as0 + i];
        m_bs[i] = B.rp[k];
        m_be[i] = B.rp[k + 1];
      }
    }
    __syncthreads();
    base = active ? seg_crp[g] - s0 : 0;
    n = active ? seg_crp[g + 1] - seg_crp[g] : 0;
    e_s = active ? seg_arp[g] - (staged ? as0 : 0) : 0;
    e_e = active ? seg_arp[g + 1] - (staged ? as0 : 0) : 0;
  } else {
    const int row = active ? rows[r0 + g] : first;
    const int c_s = c_rp[row];
    base = g * cap;
    n = active ? c_rp[row + 1] - c_s : 0;
    e_s = active ? A.rp[row] : 0;
    e_e = active ? A.rp[row + 1] : 0;
    for (int j = l; j < n; j += G) {
      cols[base + j] = c_ci[c_s + j];
      vals[base + j] = 0.0f;
    }
  }

  // Scale and B-row bounds of A entry e.
  auto meta = [&](int e, float* a, int* bs, int* be) {
    if (staged) {
      *a = m_a[e];
      *bs = m_bs[e];
      *be = m_be[e];
    } else {
      const int k = A.ci[e];
      *a = A.val[e];
      *bs = B.rp[k];
      *be = B.rp[k + 1];
    }
  };

  const int* rc = cols + base;
  float* rv = vals + base;
  float a0 = 0.0f, a1 = 0.0f;
  int bs0 = 0, be0 = 0, bs1 = 0, be1 = 0;
  if (e_s < e_e) meta(e_s, &a0, &bs0, &be0);
  if (e_s + 1 < e_e) meta(e_s + 1, &a1, &bs1, &be1);
  int c0[kU], c1[kU];
  float v0[kU], v1[kU];
#pragma unroll
  for (int j = 0; j < kU; ++j) {
    const int p = bs0 + l + j * G;
    const bool h = p < be0;
    c0[j] = h ? B.ci[p] : 0;
    v0[j] = h ? B.val[p] : 0.0f;
  }
  for (int e = e_s; e < e_e; ++e) {
#pragma unroll
    for (int j = 0; j < kU; ++j) {
      const int p = bs1 + l + j * G;
      const bool h = p < be1;
      c1[j] = h ? B.ci[p] : 0;
      v1[j] = h ? B.val[p] : 0.0f;
    }
    float a2 = 0.0f;
    int bs2 = 0, be2 = 0;
    if (e + 2 < e_e) meta(e + 2, &a2, &bs2, &be2);
    // Entry e: the products held in registers, then any remainder of a B row
    // longer than kU * G. All belong to one A entry, so their columns
    // differ and their relative order is free.
#pragma unroll
    for (int j = 0; j < kU; ++j) {
      if (bs0 + l + j * G < be0) {
        const int pos = lds_find(rc, n, c0[j]);
        rv[pos] = acc_step(rv[pos], a0, v0[j]);
      }
    }
    for (int p = bs0 + kU * G + l; p < be0; p += G) {
      const int pos = lds_find(rc, n, B.ci[p]);
      rv[pos] = acc_step(rv[pos], a0, B.val[p]);
    }
#pragma unroll
    for (int j = 0; j < kU; ++j) {
      c0[j] = c1[j];
      v0[j] = v1[j];
    }
    a0 = a1;
    bs0 = bs1;
    be0 = be1;
    a1 = a2;
    bs1 = bs2;
    be1 = be2;
  }

  if (seg) {
    __syncthreads();
    for (int i = threadIdx.x; i < s1 - s0; i += kBlock) c_store<G>(vals[i], c_val + s0 + i);
  } else if (active) {
    const int c_s = c_rp[rows[r0 + g]];
    for (int j = l; j < n; j += G) c_store<G>(rv[j], c_val + c_s + j);
  }
}

// ---------------------------------------------------------------------------
// Position map: a structure cache for repeated numeric-only multiplies.
//
// Once C's sorted columns exist, the position inside C's row that each
// product (row i, A entry ap, B entry bp) lands on is fixed for as long as
// the sparsity pattern is. map[poff[i] + t] holds it for the t-th product
// of row i in (A entry, B entry) order, as uint8 when every mapped row has
// at most 256 columns, else uint16. A numeric pass with the map reads one
// byte or two per product instead of B's column and C's columns, and does
// no search. poff is an exclusive scan of Op_i over the mapped rows.

// w[i] = Op_i for rows whose numeric bin is in `bin_mask`, else 0.
__global__ __launch_bounds__(kBlock) void k_map_weights(int M, const unsigned char* __restrict__ bin_of,
                                                        unsigned bin_mask, RowStats st,
                                                        long long* __restrict__ w,
                                                        bool span_bound) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  long long v = 0;
  if ((bin_mask >> bin_of[i]) & 1u) {
    v = st.ops[i];
    if (span_bound && v > 0) v = min(v, (long long)(st.hi[i] - st.lo[i] + 1));
  }
  w[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_scan64_reduce(const long long* __restrict__ in, int n,
                                                          long long* __restrict__ tile_sums) {
  __shared__ long long tmp[kBlock / kWave];
  const int base = blockIdx.x * kScanTile;
  long long s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int i = base + j * kBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  s = block_sum(s, tmp);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_scan64_apply(const long long* __restrict__ in, int n,
                                                         const long long* __restrict__ tile_off,
                                                         long long* __restrict__ out) {
  __shared__ long long tmp[kBlock / kWave];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  long long v[kScanItems];
  long long s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0;
    s += v[j];
  }
  long long tot;
  long long run = block_excl_scan(s, tmp, &tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) {
      out[base + j] = run;
      run += v[j];
      if (base + j == n - 1) out[n] = run;
    }
  }
}

hipError_t launch_scan64(const long long* in, int n, long long* out, long long* tile_sums,
                         long long* total, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan64_reduce, dim3(ntiles), dim3(kBlock), 0, s, in, n, tile_sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, s, tile_sums, ntiles, total);
  hipLaunchKernelGGL(k_scan64_apply, dim3(ntiles), dim3(kBlock), 0, s, in, n, tile_sums, out);
  return hipGetLastError();
}

hipError_t launch_map_offsets(int M, const unsigned char* bin_of, unsigned bin_mask, RowStats st,
                              long long* weights, long long* poff, long long* tile_sums,
                              long long* total, hipStream_t s, bool span_bound) {
  if (M <= 0) return hipSuccess;
  const int ntiles = (M + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_map_weights, dim3((M + kBlock - 1) / kBlock), dim3(kBlock), 0, s, M,
                     bin_of, bin_mask, st, weights, span_bound);
  hipLaunchKernelGGL(k_scan64_reduce, dim3(ntiles), dim3(kBlock), 0, s, weights, M, tile_sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, s, tile_sums, ntiles, total);
  hipLaunchKernelGGL(k_scan64_apply, dim3(ntiles), dim3(kBlock), 0, s, weights, M, tile_sums, poff);
  return hipGetLastError();
}

// Builds the map entries of the listed rows: 16 lanes per row, positions by
// binary search in C's sorted row (global memory; runs once per structure).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_map(const int* __restrict__ rows, int nrows,
                                                      DevCsr A, DevCsr B,
                                                      const int* __restrict__ c_rp,
                                                      const int* __restrict__ c_ci,
                                                      const long long* __restrict__ poff,
                                                      T* __restrict__ map) {
  constexpr int G = 16;
  const int r = (blockIdx.x * kBlock + threadIdx.x) / G;
  const int l = threadIdx.x & (G - 1);
  if (r >= nrows) return;
  const int row = rows[r];
  const int c_s = c_rp[row];
  const int n = c_rp[row + 1] - c_s;
  const int* cc = c_ci + c_s;
  long long mo = poff[row];
  for (int ap = A.rp[row]; ap < A.rp[row + 1]; ++ap) {
    const int k = A.ci[ap];
    const int bs = B.rp[k], be = B.rp[k + 1];
    for (int p = bs + l; p < be; p += G) map[mo + (p - bs)] = (T)lds_find(cc, n, B.ci[p]);
    mo += be - bs;
  }
}

// Entry A.rp[0] + i of A (i < nnz) at slot i: A's entries may start past 0
// (a row range of a larger matrix).
__global__ __launch_bounds__(kBlock) void k_brow_cache(const int* __restrict__ a_rp,
                                                       const int* __restrict__ a_ci, int nnz,
                                                       const int* __restrict__ b_rp,
                                                       int* __restrict__ brow,
                                                       unsigned short* __restrict__ blen) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nnz) return;
  const int k = a_ci[a_rp[0] + i];
  const int bs = b_rp[k];
  brow[i] = bs;
  blen[i] = (unsigned short)min(b_rp[k + 1] - bs, 65535);
}

hipError_t launch_brow_cache(const DevCsr& A, const DevCsr& B, int* brow, unsigned short* blen,
                             hipStream_t s) {
  if (A.nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_brow_cache, dim3((A.nnz + kBlock - 1) / kBlock), dim3(kBlock), 0, s, A.rp,
                     A.ci, A.nnz, B.rp, brow, blen);
  return hipGetLastError();
}

hipError_t launch_build_map(int width, const int* rows, int nrows, const DevCsr& A,
                            const DevCsr& B, const int* c_rp, const int* c_ci,
                            const long long* poff, void* map, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  const int grid = (int)(((long long)nrows * 16 + kBlock - 1) / kBlock);
  if (width == 1)
    hipLaunchKernelGGL(k_build_map<unsigned char>, dim3(grid), dim3(kBlock), 0, s, rows, nrows,
                       A, B, c_rp, c_ci, poff, (unsigned char*)map);
  else
    hipLaunchKernelGGL(k_build_map<unsigned short>, dim3(grid), dim3(kBlock), 0, s, rows, nrows,
                       A, B, c_rp, c_ci, poff, (unsigned short*)map);
  return hipGetLastError();
}

// Numeric pass driven by the position map. Same row/group structure and
// segment mode as k_numeric_search, but LDS holds only the accumulators
// (and, in segment mode, the segment's slice of the map, up to
// map_stage_cap_bytes), and each product is one map entry, one B value and
// one LDS add. VAR 0 and VAR 1 walk a row's A entries in chunks of kC, the
// first kU B values of every entry of a chunk in flight together; VAR 1 also
// stages every A entry's scale and B-row bounds of the segment in LDS, which
// removes the per-group (G-fold redundant) scattered loads of A and B's row
// pointers at the cost of LDS and one more load round before the barrier.
// VAR 2 fetches the A entries of a chunk with one load per lane and shares
// them by lane shuffles (see below).
template <int G, typename T, int VAR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(G <= 2 ? 8 : 1))) void k_numeric_map(
    const int* __restrict__ rows, int nrows, int row_base, int cap, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, const long long* __restrict__ poff,
    const T* __restrict__ map, float* __restrict__ c_val, const int* __restrict__ brow,
    const unsigned short* __restrict__ blen) {
  // With the B-row cache, an entry's B row comes from brow/blen at its own
  // position relative to A's first (no load of A's column, then of B's row
  // pointers).
  const bool cached = brow != nullptr;  // uniform
  if (cached) {
    const int a_base = A.rp[0];
    brow -= a_base;
    blen -= a_base;
  }
  extern __shared__ int lds[];
  constexpr int kGroups = kBlock / G;
  const int kMapCap = map_stage_cap_bytes(kGroups, cap) / (int)sizeof(T);  // map entries staged
  const int kACap = kGroups * map_stage_a_per_row(cap);  // A entries staged per segment
  constexpr int kU = VAR == 2 ? 3 : 2;
  // Accumulators, then one 16-byte slot that the branch-free adds of lanes
  // without a product go to (VAR 2), the staged map slice and A metadata.
  float* vals = reinterpret_cast<float*>(lds);
  const int spare = kGroups * cap;  // that slot's index in vals
  T* mstage = reinterpret_cast<T*>(lds + kGroups * cap + 4);
  float* m_a = reinterpret_cast<float*>(lds + kGroups * cap + 4 + map_stage_bytes(kGroups, cap) / 4);
  int* m_bs = reinterpret_cast<int*>(m_a + kACap);
  int* m_be = m_bs + kACap;
  __shared__ int s_crp[kGroups + 1], s_arp[kGroups + 1];
  __shared__ long long s_poff[kGroups + 1];

  const int r0 = xcd_block(blockIdx.x, gridDim.x) * kGroups;
  const int nr = min(kGroups, nrows - r0);
  const int g = threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  const bool active = g < nr;
  // A bin of consecutive rows needs no look into its row list.
  const int first = row_base >= 0 ? row_base + r0 : rows[r0];
  const int last = row_base >= 0 ? first + nr - 1 : rows[r0 + nr - 1];
  const bool seg = (last - first == nr - 1);

  int c_s = 0, n = 0, a_s = 0, a_e = 0;
  long long mo = 0;
  int base;
  int s0 = 0, s1 = 0;
  int a_off = 0;          // subtracted from A positions to index staged metadata
  bool a_staged = false;  // uniform
  const T* rm;            // this row's map entries
  bool rm_lds = false;    // uniform: rm points into the staged copy in LDS
  int rm_shift = 0;       // bytes from mstage to the staged entry m0
  long long rm_idx = 0;   // this row's first entry relative to m0
  if (seg) {
    // Rows first..last: their row pointers of C and A and their map offsets
    // are staged with coalesced loads, then the C segment's accumulators are
    // cleared and the segment's map entries staged.
    for (int i = threadIdx.x; i <= nr; i += kBlock) {
      s_crp[i] = c_rp[first + i];
      s_arp[i] = A.rp[first + i];
      s_poff[i] = poff[first + i];
    }
    s0 = c_rp[first];
    s1 = c_rp[last + 1];
    const long long m0 = poff[first];
    const long long m1 = poff[last + 1];
    const bool staged = m1 - m0 <= kMapCap;  // uniform
    if (VAR == 1) {
      a_off = A.rp[first];
      a_staged = A.rp[last + 1] - a_off <= kACap;
    }
    for (int i = threadIdx.x; i < s1 - s0; i += kBlock) vals[i] = 0.0f;
    int mshift = 0;  // bytes from the staged copy's start to entry m0
    if (staged) {
      // Aligned 16-byte loads over the segment's bytes (the map buffer has
      // slack past its end for the last one).
      const long long b0 = m0 * (long long)sizeof(T);
      const long long ab0 = b0 & ~15LL;
      const int n16 = (int)((m1 * (long long)sizeof(T) - ab0 + 15) >> 4);
      const uint4* src = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(map) + ab0);
      unsigned* dst = reinterpret_cast<unsigned*>(mstage);  // 4-byte aligned in LDS
      for (int i = threadIdx.x; i < n16; i += kBlock) {
        const uint4 v = src[i];
        dst[4 * i] = v.x;
        dst[4 * i + 1] = v.y;
        dst[4 * i + 2] = v.z;
        dst[4 * i + 3] = v.w;
      }
      mshift = (int)(b0 - ab0);
    }
    if (a_staged) {
      const int na = A.rp[last + 1] - a_off;
      for (int i = threadIdx.x; i < na; i += kBlock) {
        m_a[i] = A.val[a_off + i];
        if (cached) {
          m_bs[i] = brow[a_off + i];
          m_be[i] = m_bs[i] + blen[a_off + i];
        } else {
          const int k = A.ci[a_off + i];
          m_bs[i] = B.rp[k];
          m_be[i] = B.rp[k + 1];
        }
      }
    }
    __syncthreads();
    c_s = s0;
    mo = m0;
    if (active) {
      c_s = s_crp[g];
      n = s_crp[g + 1] - c_s;
      mo = s_poff[g];
      a_s = s_arp[g];
      a_e = s_arp[g + 1];
    }
    base = c_s - s0;
    rm = staged ? reinterpret_cast<const T*>(reinterpret_cast<const char*>(mstage) + mshift) +
                      (mo - m0)
                : map + mo;
    rm_lds = staged;
    rm_shift = mshift;
    rm_idx = mo - m0;
  } else {
    const int row = active ? rows[r0 + g] : first;
    c_s = c_rp[row];
    n = active ? c_rp[row + 1] - c_s : 0;
    mo = poff[row];
    a_s = active ? A.rp[row] : 0;
    a_e = active ? A.rp[row + 1] : 0;
    rm = map + mo;
    base = g * cap;
    for (int j = l; j < n; j += G) vals[base + j] = 0.0f;
  }

  // Scale and B-row bounds of A entry e.
  auto meta = [&](int e, float* a, int* bs, int* be) {
    if (a_staged) {
      *a = m_a[e - a_off];
      *bs = m_bs[e - a_off];
      *be = m_be[e - a_off];
    } else if (cached) {
      *a = A.val[e];
      *bs = brow[e];
      *be = *bs + blen[e];
    } else {
      const int k = A.ci[e];
      *a = A.val[e];
      *bs = B.rp[k];
      *be = B.rp[k + 1];
    }
  };

  if constexpr (VAR == 2) {
    // Chunked loads: for up to kChunk A entries at once, their columns and
    // scales, then their B-row bounds, then the first kU products of each
    // are in flight together (three dependent rounds per chunk instead of
    // one or more per entry); the entries are then applied in order. A-entry
    // loads past the row's last entry read a valid address (the last entry)
    // and are masked, so every round issues at once. One lane's products of one
    // entry have distinct positions, so their accumulators are read, added
    // and written back together; entry u + 1's reads follow entry u's writes
    // in program order.
    constexpr int kChunk = 5;
    if (active && a_s < a_e && B.nnz > 0) {
      float* rv = vals + base;
      const int rv_spare = spare - base;
      const int bmax = B.nnz - 1;
      int t0 = 0;
      auto apply_chunk = [&](const auto* mp) {
        for (int c0 = a_s; c0 < a_e; c0 += kChunk) {
          int bb[kChunk], ee[kChunk];
          float aa[kChunk], vv[kChunk][kU];
          constexpr int kPer = (kChunk + G - 1) / G;
          int kl[kPer], bl[kPer], el[kPer];
          float al[kPer];
          if (cached) {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
              const int e = min(c0 + l + u * G, a_e - 1);
              al[u] = A.val[e];
              bl[u] = brow[e];
              el[u] = bl[u] + blen[e];
            }
          } else {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
              const int e = min(c0 + l + u * G, a_e - 1);
              kl[u] = A.ci[e];
              al[u] = A.val[e];
            }
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
              bl[u] = B.rp[kl[u]];
              el[u] = B.rp[kl[u] + 1];
            }
          }
#pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            const bool h = c0 + u < a_e;
            aa[u] = __shfl(al[u / G], u % G, G);
            bb[u] = __shfl(bl[u / G], u % G, G);
            ee[u] = h ? __shfl(el[u / G], u % G, G) : bb[u];
          }
          // The first kU products of each entry. With G <= 2 branch-free
          // (the kernel then fits 64 registers): every lane loads a B value
          // (past its entry's end, a valid entry of B that is dropped) and a
          // map entry (past its entry's end, the entry's last or, for an
          // empty entry, the one after the row's products, within the map's
          // slack), and adds into its position or, with no product, into the
          // spare slot. Wider groups keep the per-product branches, which
          // hold fewer registers.
#pragma unroll
          for (int u = 0; u < kChunk; ++u)
#pragma unroll
            for (int j = 0; j < kU; ++j) {
              const int p = bb[u] + l + j * G;
              if constexpr (G <= 2)
                vv[u][j] = B.val[min(p, bmax)];
              else
                vv[u][j] = p < ee[u] ? B.val[p] : 0.0f;
            }
#pragma unroll
          for (int u = 0; u < kChunk; ++u) {
            const int len = ee[u] - bb[u];  // 0 past the row's last entry
            int pos[kU];
            float cur[kU];
            if constexpr (G <= 2) {
#pragma unroll
              for (int j = 0; j < kU; ++j) {
                const int t = l + j * G;
                const int m = (int)mp[t0 + min(t, max(len - 1, 0))];
                pos[j] = t < len ? m : rv_spare;
              }
#pragma unroll
              for (int j = 0; j < kU; ++j) cur[j] = rv[pos[j]];
#pragma unroll
              for (int j = 0; j < kU; ++j) rv[pos[j]] = acc_step(cur[j], aa[u], vv[u][j]);
            } else {
#pragma unroll
              for (int j = 0; j < kU; ++j) {
                const int t = l + j * G;
                pos[j] = t < len ? (int)mp[t0 + t] : -1;
              }
#pragma unroll
              for (int j = 0; j < kU; ++j) cur[j] = pos[j] >= 0 ? rv[pos[j]] : 0.0f;
#pragma unroll
              for (int j = 0; j < kU; ++j)
                if (pos[j] >= 0) rv[pos[j]] = acc_step(cur[j], aa[u], vv[u][j]);
            }
            for (int t = kU * G + l; t < len; t += G) {
              const int q = mp[t0 + t];
              rv[q] = acc_step(rv[q], aa[u], B.val[bb[u] + t]);
            }
            t0 += len;
          }
        }
      };
      // Separate instantiations, so the staged map is read with LDS
      // instructions and the unstaged one with global loads.
      if (rm_lds)
        apply_chunk(reinterpret_cast<const T*>(reinterpret_cast<const char*>(mstage) + rm_shift) +
                    rm_idx);
      else
        apply_chunk(map + mo);
    }
  } else {
  // Chunks of kC entries: their scales and B-row bounds (from LDS when
  // staged), then the first kU B values of every entry of the chunk in
  // flight together, then the entries applied in order. Entries past the
  // row's last read the last one and are empty. (No load destination is
  // copied into another register while in flight: such a copy waits for it.)
  constexpr int kC = G <= 2 ? 4 : 8;
  if (active && a_s < a_e) {
    float* rv = vals + base;
    int t0 = 0;  // index of the chunk's first product within the row
    for (int c0 = a_s; c0 < a_e; c0 += kC) {
      float av[kC], vv[kC][kU];
      int bsv[kC], bev[kC];
#pragma unroll
      for (int u = 0; u < kC; ++u) {
        meta(min(c0 + u, a_e - 1), &av[u], &bsv[u], &bev[u]);
        if (c0 + u >= a_e) bev[u] = bsv[u];
      }
#pragma unroll
      for (int u = 0; u < kC; ++u)
#pragma unroll
        for (int j = 0; j < kU; ++j) {
          const int p = bsv[u] + l + j * G;
          vv[u][j] = p < bev[u] ? B.val[p] : 0.0f;
        }
#pragma unroll
      for (int u = 0; u < kC; ++u) {
        const int len = bev[u] - bsv[u];
#pragma unroll
        for (int j = 0; j < kU; ++j) {
          const int t = l + j * G;
          if (t < len) {
            const int pos = rm[t0 + t];
            rv[pos] = acc_step(rv[pos], av[u], vv[u][j]);
          }
        }
        for (int t = kU * G + l; t < len; t += G) {
          const int pos = rm[t0 + t];
          rv[pos] = acc_step(rv[pos], av[u], B.val[bsv[u] + t]);
        }
        t0 += len;
      }
    }
  }
  }  // VAR != 2

  if (seg) {
    __syncthreads();
    for (int i = threadIdx.x; i < s1 - s0; i += kBlock) c_store<G>(vals[i], c_val + s0 + i);
  } else {
    for (int j = l; j < n; j += G) c_store<G>(vals[base + j], c_val + c_s + j);
  }
}

template <int G>
static hipError_t launch_numeric_map_g(const BinSpec& b, int width, const int* rows, int nrows,
                                       int row_base, const DevCsr& A, const DevCsr& B,
                                       const int* c_rp, const long long* poff, const void* map,
                                       float* c_val, const int* brow, const unsigned short* blen,
                                       hipStream_t s) {
  const int groups = kBlock / G;
  const int grid = (nrows + groups - 1) / groups;
#define RVSP_MAP_LAUNCH(T, SA)                                                              \
  hipLaunchKernelGGL((k_numeric_map<G, T, SA>), dim3(grid), dim3(kBlock), b.lds_bytes, s, rows, \
                     nrows, row_base, b.cap, A, B, c_rp, poff, (const T*)map, c_val, brow, blen)
  if (b.variant < 0 || b.variant > 2) return hipErrorInvalidValue;
  if (width == 1) {
    if (b.variant == 0) RVSP_MAP_LAUNCH(unsigned char, 0);
    else if (b.variant == 1) RVSP_MAP_LAUNCH(unsigned char, 1);
    else RVSP_MAP_LAUNCH(unsigned char, 2);
  } else {
    if (b.variant == 0) RVSP_MAP_LAUNCH(unsigned short, 0);
    else if (b.variant == 1) RVSP_MAP_LAUNCH(unsigned short, 1);
    else RVSP_MAP_LAUNCH(unsigned short, 2);
  }
#undef RVSP_MAP_LAUNCH
  return hipGetLastError();
}

hipError_t launch_numeric_map_bin(const BinSpec& b, const MapView& mv, const int* rows, int nrows,
                                  const DevCsr& A, const DevCsr& B, const int* c_rp,
                                  float* c_val, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  if (!mv.map || !mv.poff || (mv.width != 1 && mv.width != 2)) return hipErrorInvalidValue;
  switch (b.group) {
    case 1: return launch_numeric_map_g<1>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    case 2: return launch_numeric_map_g<2>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    case 4: return launch_numeric_map_g<4>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    case 8: return launch_numeric_map_g<8>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    case 16: return launch_numeric_map_g<16>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    case 32: return launch_numeric_map_g<32>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    case 64: return launch_numeric_map_g<64>(b, mv.width, rows, nrows, mv.row_base, A, B, c_rp, mv.poff, mv.map, c_val, mv.brow, mv.blen, s);
    default: return hipErrorInvalidValue;
  }
}

// First index in the ascending b_ci[s, e) whose value is >= key.
__device__ __forceinline__ int gm_lower_bound(const int* __restrict__ b_ci, int s,
                                              int e, int key) {
  while (s < e) {
    const int m = (s + e) >> 1;
    if (b_ci[m] < key) s = m + 1; else e = m;
  }
  return s;
}

// One workgroup per row, each of its four waves owning a contiguous quarter
// of the row's output positions (hence a contiguous column range). Every wave
// walks all of A's row in order and applies only the products that land in
// its own columns, so no two waves ever touch the same C(i,j) and no
// barrier is needed between A entries. GLOBAL_ACC == false keeps the
// row's columns and accumulators in LDS (n <= cap); GLOBAL_ACC == true
// accumulates in this workgroup's dense slot of global memory, indexed by
// column.
template <bool GLOBAL_ACC>
__global__ __launch_bounds__(kBlock) void k_numeric_rowblock(
    const int* __restrict__ rows, int nrows, int cap, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, const int* __restrict__ c_ci,
    float* __restrict__ c_val, float* __restrict__ gacc, size_t slot_floats) {
  extern __shared__ int lds[];
  constexpr int kWaves = kBlock / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  int* cols = lds;
  float* vals = reinterpret_cast<float*>(lds + cap);
  float* acc = GLOBAL_ACC ? gacc + (size_t)blockIdx.x * slot_floats : nullptr;
  for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
    const int row = rows[r];
    const int c_s = c_rp[row];
    const int n = c_rp[row + 1] - c_s;
    // This wave's output positions [p0, p1) and their column range.
    const int p0 = (int)(((long long)n * wid) / kWaves);
    const int p1 = (int)(((long long)n * (wid + 1)) / kWaves);
    if (p1 > p0) {
      for (int j = p0 + lane; j < p1; j += kWave) {
        const int col = c_ci[c_s + j];
        if (GLOBAL_ACC) {
          acc[col] = 0.0f;
        } else {
          cols[j] = col;
          vals[j] = 0.0f;
        }
      }
      const int clo = c_ci[c_s + p0];
      const int chi = c_ci[c_s + p1 - 1];
      const int a_e = A.rp[row + 1];
      for (int ap = A.rp[row]; ap < a_e; ++ap) {
        const int k = A.ci[ap];
        const float a = A.val[ap];
        int bs = B.rp[k], be = B.rp[k + 1];
        if (be - bs > 4 * kWave) {
          // Long B row: narrow it to this wave's columns first.
          bs = gm_lower_bound(B.ci, bs, be, clo);
          be = gm_lower_bound(B.ci, bs, be, chi + 1);
        }
        for (int p = bs + lane; p < be; p += kWave) {
          const int col = B.ci[p];
          if (col < clo || col > chi) continue;
          if (GLOBAL_ACC) {
            acc[col] = acc_step(acc[col], a, B.val[p]);
          } else {
            const int pos = p0 + lds_find(cols + p0, p1 - p0, col);
            vals[pos] = acc_step(vals[pos], a, B.val[p]);
          }
        }
      }
      for (int j = p0 + lane; j < p1; j += kWave)
        __builtin_nontemporal_store(GLOBAL_ACC ? acc[c_ci[c_s + j]] : vals[j], c_val + c_s + j);
    }
    __syncthreads();  // LDS / slot reuse by the next row
  }
}

// First index in [lo, hi] of the ascending b_ci[lo, hi) whose value is >= key,
// searched forward from lo: doubling steps, then halving (cheap when the
// answer is near lo).
// gm_lower_bound over the B row [b0, be) through its split-table entries
// `sp` (meta = sp[0..1]), for an answer known to be >= from.
__device__ __forceinline__ int split_lower_bound(const int* __restrict__ b_ci,
                                                 const int* __restrict__ sp, int2 meta, int b0,
                                                 int be, int from, int key) {
  const int g = meta.y & 31, nb = meta.y >> 5;
  if (key <= meta.x) return max(b0, from);
  const int j = (int)min((long long)nb, ((long long)key - meta.x) >> g);
  // Granule j: positions [bound(j), bound(j + 1)), bound(0) = b0, bound(nb) = be.
  const int s = j <= 0 ? b0 : (j >= nb ? be : sp[1 + j]);
  const int e = j + 1 >= nb ? be : sp[2 + j];
  return gm_lower_bound(b_ci, max(s, from), max(e, from), key);
}

__device__ __forceinline__ int gm_gallop(const int* __restrict__ b_ci, int lo, int hi,
                                         int key) {
  if (lo >= hi || b_ci[lo] >= key) return lo;
  int bound = 1;
  while (lo + bound < hi && b_ci[lo + bound] < key) bound <<= 1;
  return gm_lower_bound(b_ci, lo + (bound >> 1) + 1, min(lo + bound, hi), key);
}

// Long rows: one workgroup per (row, run of consecutive output positions).
// The run is cut into tiles of at most T positions whose columns span fewer
// than 32 * (T / 4) columns; a tile's accumulators live in LDS and its
// positions are found through a bitmap index of its column span (per 32
// columns the set bits and the positions before them). Where the span allows
// fewer than T / 2 positions, the tile holds T / 2 positions instead, whose
// columns replace the index and are binary-searched. The row's products
// are streamed in reference order (A entries ascending, then B positions
// ascending) in groups of kTiledItems products per lane: every lane finds its
// products' A entries, gathers the B elements and looks their columns up in
// the tile, all lanes in parallel. Only the accumulate is ordered. Each wave
// owns every fourth 32-position stripe of the tile; a group's products are
// routed to the owner's list in product order (a stable split through LDS),
// and each wave applies its list 64 products at a time: products of one batch
// that land on the same position are applied lowest lane first (per round an
// LDS atomicMin of (position, lane) on a small per-wave hash of positions
// picks, for every position, its lowest pending lane; one wave's LDS
// operations execute in program order, so no barrier is needed). Every C(i,j)
// thus sees exactly the reference's sequence of separately rounded adds. A
// tile that is not the whole row narrows every B row to the tile's columns;
// along a workgroup's run the narrowed range of a B row starts where the
// previous tile's ended (a per-lane cursor, for A rows of up to kCurChunks *
// kBlock entries), and split tables bound the searches of long B rows.
constexpr int kTiledItems = 2;             // products per lane per group
constexpr int kTiledGroup = kTiledItems * kBlock;
constexpr int kTiledWaves = kBlock / kWave;
constexpr int kTagSlots = 256;             // per-wave hash of pending positions
constexpr int kEntrySlots = 1024;          // entry table of a chunk
constexpr int kPrePad = 8;                 // entries probed past a table hit

__host__ __device__ constexpr int tiled_index_words(int T) { return T / 4 > 64 ? T / 4 : 64; }

struct TiledLds {
  int* cnt;         // [owner][item][wave] products of a group
  float* vals;      // T accumulators
  int2* stage;      // kTiledGroup (position, product bits), grouped by owner wave
  uint2* idx;       // tiled_index_words(T) words of the tile's bitmap index
  int* tags;        // kTiledWaves x kTagSlots
  int* etab;        // kEntrySlots: entry of product m << shift
  int* pre;         // kBlock + kPrePad exclusive prefix of the chunk's product counts
  int* bstart;      // kBlock B-row starts
  float* aval;      // kBlock A values
};

__device__ __forceinline__ TiledLds tiled_lds(int* lds, int T) {
  TiledLds L;
  L.cnt = lds;  // 16-byte aligned; every later array stays 8-byte aligned
  L.stage = reinterpret_cast<int2*>(lds + kTiledItems * kTiledWaves * kTiledWaves);
  L.idx = reinterpret_cast<uint2*>(L.stage + kTiledGroup);
  L.vals = reinterpret_cast<float*>(L.idx + tiled_index_words(T));
  L.tags = reinterpret_cast<int*>(L.vals + ((T + 1) & ~1));
  L.etab = L.tags + kTiledWaves * kTagSlots;
  L.pre = L.etab + kEntrySlots;
  L.bstart = L.pre + kBlock + kPrePad;
  L.aval = reinterpret_cast<float*>(L.bstart + kBlock);
  return L;
}

constexpr int kCurChunks = 16;
constexpr int kProbe = 8;  // B columns read together past a narrowed range's start

// The A entry of product q of a chunk (the last e with pre[e] <= q): the
// entry table gives the entry of the slot's first product, the kPrePad next
// prefixes are compared at once, and galloping finishes the rare longer step.
__device__ __forceinline__ int tiled_entry(const TiledLds& L, int shift, int q) {
  int e = L.etab[q >> shift];
  int c = 0;
#pragma unroll
  for (int j = 1; j <= kPrePad; ++j) c += L.pre[e + j] <= q ? 1 : 0;
  e += c;
  if (c == kPrePad) e = advance_entry(L.pre, e, q);
  return e;
}

// Position of column col (clo <= col <= chi, present) in the tile.
__device__ __forceinline__ int tiled_index_pos(const uint2* idx, int clo, int col) {
  const int d = col - clo;
  const uint2 v = idx[d >> 5];
  return (int)v.y + __popc(v.x & ((1u << (d & 31)) - 1u));
}

// The LDS set-up every tiled workgroup does once, before its first row.
__device__ __forceinline__ void tiled_init(const TiledLds& L) {
  const int tid = threadIdx.x;
  if (tid < kPrePad) L.pre[kBlock + tid] = INT_MAX;  // probes past the chunk end
  for (int i = tid; i < kTiledWaves * kTagSlots; i += kBlock) L.tags[i] = INT_MAX;
  __syncthreads();
}

// One workgroup's part of a long row: run blockIdx.y of the row's runs
// (block-uniform arguments; every thread of the block must call it).
__device__ __forceinline__ void tiled_row(const TiledLds& L, int* scan_tmp, int row, int T,
                                          const DevCsr& A, const DevCsr& B,
                                          const int* __restrict__ c_rp,
                                          const int* __restrict__ c_ci,
                                          float* __restrict__ c_val, const SplitTable& split) {
  const int span = 32 * tiled_index_words(T);  // columns a tile may span
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wid = tid >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  int* my_tags = L.tags + wid * kTagSlots;
  {
    const int c_s = c_rp[row];
    const int n = c_rp[row + 1] - c_s;
    // This workgroup's run: one of up to gridDim.y equal parts of the row.
    // (32-bit arithmetic only: 64-bit division expands into fused float
    // operations, which no device code here may contain.)
    const int runs = min((int)gridDim.y, n / T + (n % T != 0));
    if ((int)blockIdx.y >= runs) return;
    const int rq = n / runs, rr = n % runs, y = blockIdx.y;
    const int P0 = rq * y + rr * y / runs;                 // floor(n * y / runs)
    const int P1 = rq * (y + 1) + rr * (y + 1) / runs;
    const int a_s = A.rp[row], a_e = A.rp[row + 1];
    const bool cursors = a_e - a_s <= kCurChunks * kBlock;
    int cur[kCurChunks];  // per chunk: where this lane's B row continues
    bool first = true;
    for (int p0 = P0; p0 < P1;) {
      const int clo = c_ci[c_s + p0];
      // The tile: up to T positions whose columns stay within the span of
      // the index, or, where that holds fewer positions, up to T / 2
      // positions whose columns are kept in the index's place and searched.
      int pe = min(P1, p0 + T);
      if (c_ci[c_s + pe - 1] - clo >= span) {
        // First position at or past clo + span: a probe every `step`
        // positions, then every position of the step that holds it.
        const int len = pe - p0;
        const int step = (len + kBlock - 1) / kBlock;
        const int c1 = __syncthreads_count(c_ci[c_s + p0 + min(len - 1, tid * step)] - clo < span);
        const int b1 = p0 + (c1 - 1) * step;  // last probe inside the span
        const int q2 = b1 + 1 + tid;
        const int c2 = __syncthreads_count(tid < step - 1 && q2 < p0 + len &&
                                           c_ci[c_s + q2] - clo < span);
        pe = b1 + 1 + c2;
      }
      const bool use_idx = pe - p0 >= min(P1 - p0, T / 2);  // uniform
      if (!use_idx) pe = min(P1, p0 + T / 2);
      const int tn = pe - p0;
      const int chi = c_ci[c_s + pe - 1];
      const bool narrow = !(p0 == 0 && pe == n);
      int* cols = reinterpret_cast<int*>(L.idx);  // T / 2 columns in cols mode
      const int iw = ((chi - clo) >> 5) + 1;
      if (use_idx)
        for (int w = tid; w < iw; w += kBlock) L.idx[w] = make_uint2(0u, 0u);
      for (int j = tid; j < tn; j += kBlock) L.vals[j] = 0.0f;
      __syncthreads();
      if (use_idx) {
        for (int j = tid; j < tn; j += kBlock) {
          const int d = c_ci[c_s + p0 + j] - clo;
          atomicOr(&L.idx[d >> 5].x, 1u << (d & 31));
        }
      } else {
        for (int j = tid; j < tn; j += kBlock) cols[j] = c_ci[c_s + p0 + j];
      }
      __syncthreads();
      if (use_idx) {
        // Positions before each word: per thread a run of consecutive words.
        const int per_t = (iw + kBlock - 1) / kBlock;
        const int w_beg = min(iw, tid * per_t), w_end = min(iw, w_beg + per_t);
        int cnt_t = 0;
        for (int w = w_beg; w < w_end; ++w) cnt_t += __popc(L.idx[w].x);
        int tot_w;
        int run = block_excl_scan(cnt_t, scan_tmp, &tot_w);
        for (int w = w_beg; w < w_end; ++w) {
          L.idx[w].y = (unsigned)run;
          run += __popc(L.idx[w].x);
        }
      }
      __syncthreads();
      int chunk = 0;
      for (int cb = a_s; cb < a_e; cb += kBlock, ++chunk) {
        // One A entry per lane: its scale and the part of its B row that
        // can land in this tile.
        const int ap = cb + tid;
        int bs = 0, len = 0;
        float a = 0.0f;
        if (ap < a_e) {
          const int k = A.ci[ap];
          a = A.val[ap];
          int be = B.rp[k + 1];
          if (narrow) {
            const int b0 = B.rp[k];
            // A long B row's split table bounds both searches to one granule
            // (read only when a search needs it).
            int so = -1;
            int2 meta = make_int2(0, 0);
            auto load_split = [&] {
              if (split.off != nullptr && be - b0 >= kSplitMin && split.off[k + 1] > split.off[k]) {
                so = split.off[k];
                meta = *reinterpret_cast<const int2*>(split.sp + so);
              }
            };
            const bool resume = cursors && !first;
            int lo = 0;
            if (resume) {
#pragma unroll
              for (int c = 0; c < kCurChunks; ++c)
                if (c == chunk) lo = cur[c];
            } else {
              load_split();
              lo = so >= 0 ? split_lower_bound(B.ci, split.sp + so, meta, b0, be, b0, clo)
                           : gm_lower_bound(B.ci, b0, be, clo);
            }
            bs = lo;
            // Most narrowed ranges hold a product or two: the columns at lo ..
            // lo + kProbe - 1, read together (clamped to the row), end the
            // range without a search unless all of them lie in the tile.
            int n_in = 0;
            if (lo < be) {
#pragma unroll
              for (int j = 0; j < kProbe; ++j) {
                const int cj = B.ci[min(lo + j, be - 1)];
                n_in += (lo + j < be && cj <= chi) ? 1 : 0;
              }
            }
            if (n_in < kProbe) {
              be = lo + n_in;
            } else {
              if (resume) load_split();
              be = so >= 0 ? split_lower_bound(B.ci, split.sp + so, meta, b0, be, lo + kProbe,
                                               chi + 1)
                           : gm_gallop(B.ci, lo + kProbe, be, chi + 1);
            }
            if (cursors) {
#pragma unroll
              for (int c = 0; c < kCurChunks; ++c)
                if (c == chunk) cur[c] = be;
            }
          } else {
            bs = B.rp[k];
          }
          len = be - bs;
        }
        int total;
        const int excl = block_excl_scan(len, scan_tmp, &total);
        L.pre[tid] = excl;
        L.bstart[tid] = bs;
        L.aval[tid] = a;
        // Entry table: slot m holds the entry of product m << shift.
        int shift = 5;
        while ((total >> shift) >= kEntrySlots) ++shift;
        {
          const int step = 1 << shift;
          for (int m = (excl + step - 1) >> shift; (m << shift) < excl + len; ++m)
            L.etab[m] = tid;
        }
        __syncthreads();
        // Groups of kTiledGroup products, kTiledItems per lane (products
        // g kTiledGroup + i kBlock + tid), software-pipelined: the gathers of
        // the next two groups are in flight while a group's lookups and adds
        // run (occupancy is bounded by LDS here, not registers). Unrolled by
        // three so the groups' registers change roles without copies (a copy
        // of a load's destination waits for the load), with the loop's only
        // exit at its end (an exit between the groups would leave a path to
        // the loop's start with a group's gathers unused, where the compiler
        // then waits for every load in flight).
        // The gathers are unconditional (a lane past the chunk's last
        // product reads that product again and drops it), and every group's
        // loaded values are used on every path: a load skipped or left unused
        // on some path makes the compiler wait there for later ones.
        struct TGroup {
          int col[kTiledItems];
          float b[kTiledItems], a[kTiledItems];
          bool ok[kTiledItems];
        };
        auto gather = [&](int q, int& c, float& bv_, float& av_, bool& ok) {
          // The A entry of product q (entries without products share their
          // successor's prefix).
          ok = q < total;
          const int qq = ok ? q : total - 1;
          const int e = tiled_entry(L, shift, qq);
          const int bi = L.bstart[e] + (qq - L.pre[e]);
          c = B.ci[bi];
          bv_ = B.val[bi];
          av_ = L.aval[e];
        };
        auto issue = [&](int g) {
          TGroup t;
#pragma unroll
          for (int i = 0; i < kTiledItems; ++i)
            gather(g * kTiledGroup + i * kBlock + tid, t.col[i], t.b[i], t.a[i], t.ok[i]);
          return t;
        };
        auto process = [&](const TGroup& t) {
          int pos[kTiledItems], own[kTiledItems], rank[kTiledItems];
          float prod[kTiledItems];
#pragma unroll
          for (int i = 0; i < kTiledItems; ++i) {
            const bool in = t.ok[i] && t.col[i] >= clo && t.col[i] <= chi;
            prod[i] = in ? t.a[i] * t.b[i] : 0.0f;  // rounded here, added below
            pos[i] = -1;
            own[i] = kTiledWaves;  // no product
            if (in) {
              pos[i] = use_idx ? tiled_index_pos(L.idx, clo, t.col[i]) : lds_find(cols, tn, t.col[i]);
              own[i] = (pos[i] >> 5) & (kTiledWaves - 1);
            }
          }
          // Stable split by owner: per (item, wave) the count for every
          // owner, and each product's rank among its wave's products with
          // the same owner. Counts are stored owner-major, (item, wave) minor,
          // so their exclusive scan is every (owner, item, wave) run's offset
          // in the staging area: owner 0's list first, each list in (item,
          // wave, lane) order, which is product order.
          const unsigned long long has0 = __ballot(own[0] < kTiledWaves);
          const unsigned long long lo0 = __ballot(own[0] & 1), hi0 = __ballot(own[0] & 2);
          const unsigned long long has1 = __ballot(own[1] < kTiledWaves);
          const unsigned long long lo1 = __ballot(own[1] & 1), hi1 = __ballot(own[1] & 2);
#pragma unroll
          for (int i = 0; i < kTiledItems; ++i) {
            const unsigned long long has = i ? has1 : has0;
            const unsigned long long lo = i ? lo1 : lo0, hi = i ? hi1 : hi0;
            if (lane < kTiledWaves) {
              const unsigned long long m =
                  has & ((lane & 1) ? lo : ~lo) & ((lane & 2) ? hi : ~hi);
              L.cnt[lane * kTiledItems * kTiledWaves + i * kTiledWaves + wid] = __popcll(m);
            }
            const int o = own[i] & (kTiledWaves - 1);
            const unsigned long long mine = has & ((o & 1) ? lo : ~lo) & ((o & 2) ? hi : ~hi);
            rank[i] = __popcll(mine & below);
          }
          static_assert(kTiledItems == 2 && kTiledWaves == 4, "split is written for 2 x 4");
          __syncthreads();
          constexpr int kRuns = kTiledWaves * kTiledItems * kTiledWaves;  // 32
          const int cv = lane < kRuns ? L.cnt[lane] : 0;
          const int run_incl = group_incl_scan<kWave>(cv);
          const int run_off = run_incl - cv;
#pragma unroll
          for (int i = 0; i < kTiledItems; ++i) {
            const int o = own[i] & (kTiledWaves - 1);
            const int off = __shfl(run_off, o * kTiledItems * kTiledWaves + i * kTiledWaves + wid);
            if (own[i] < kTiledWaves)
              L.stage[off + rank[i]] = make_int2(pos[i], __float_as_int(prod[i]));
          }
          // This wave's list: positions in its stripes only, so no other
          // wave touches them; 64 products per batch, in order.
          const int my_base = __shfl(run_off, wid * kTiledItems * kTiledWaves);
          const int my_end = wid + 1 < kTiledWaves
                                 ? __shfl(run_off, (wid + 1) * kTiledItems * kTiledWaves)
                                 : __shfl(run_incl, kRuns - 1);
          const int my_n = my_end - my_base;
          __syncthreads();
          for (int k0 = 0; k0 < my_n; k0 += kWave) {
            const int j = k0 + lane;
            bool act = j < my_n;
            int p = 0, key = 0, slot = 0;
            float v = 0.0f;
            if (act) {
              const int2 sv = L.stage[my_base + j];
              p = sv.x;
              v = __int_as_float(sv.y);
              key = (p << 6) | lane;
              slot = (int)(((unsigned)p * 2654435761u) >> 24);  // kTagSlots = 256
            }
            // Per round, each hash slot keeps its smallest (position, lane):
            // for every position the lowest lane still pending is the only
            // one that can win, so its products are applied in lane order.
            while (__any(act)) {
              if (act) atomicMin(&my_tags[slot], key);
              if (act && my_tags[slot] == key) {
                L.vals[p] = L.vals[p] + v;
                my_tags[slot] = INT_MAX;
                act = false;
              }
            }
          }
        };
        const int ngroups = (total + kTiledGroup - 1) / kTiledGroup;  // block-uniform
        if (ngroups > 0) {
          TGroup ga = issue(0);
          TGroup gb = issue(1);
          TGroup gc;
          int g = 0;
          for (; g + 3 <= ngroups; g += 3) {
            gc = issue(g + 2);
            process(ga);
            ga = issue(g + 3);
            process(gb);
            gb = issue(g + 4);
            process(gc);
          }
          // The last one or two groups.
          for (int r = g; r < ngroups; ++r) process(r == g ? ga : gb);
        }
        __syncthreads();  // the chunk's metadata is replaced next
      }
      for (int j = tid; j < tn; j += kBlock)
        __builtin_nontemporal_store(L.vals[j], c_val + c_s + p0 + j);
      __syncthreads();  // the tile's LDS is reused by the next tile
      p0 = pe;
      first = false;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_numeric_tiled(
    const int* __restrict__ rows, int nrows, int T, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, const int* __restrict__ c_ci,
    float* __restrict__ c_val, SplitTable split) {
  extern __shared__ __attribute__((aligned(16))) int tiled_lds_raw[];
  __shared__ int scan_tmp[kBlock / kWave];
  const TiledLds L = tiled_lds(tiled_lds_raw, T);
  tiled_init(L);
  for (int r = blockIdx.x; r < nrows; r += gridDim.x)
    tiled_row(L, scan_tmp, rows[r], T, A, B, c_rp, c_ci, c_val, split);
}

// The tiled kernel over the long-row bins of several descriptors at once
// (the groups of a batch): segment i covers workgroup rows [start, start +
// nrows) of the fused grid, with its own row list, A row pointers (a row
// range of the stacked A) and C arrays; A's columns and values and all of
// B are shared. One launch replaces one per bin and group, so the groups'
// long rows share the GPU without a tail per launch.
__global__ __launch_bounds__(kBlock) void k_numeric_tiled_fused(TiledSegs segs, int T, DevCsr A,
                                                                DevCsr B, SplitTable split) {
  extern __shared__ __attribute__((aligned(16))) int tiled_lds_raw[];
  __shared__ int scan_tmp[kBlock / kWave];
  const TiledLds L = tiled_lds(tiled_lds_raw, T);
  tiled_init(L);
  for (int x = blockIdx.x; x < segs.total; x += gridDim.x) {
    int i = 0;
    while (i + 1 < segs.n && segs.seg[i + 1].start <= x) ++i;  // uniform
    const TiledSeg& sg = segs.seg[i];
    DevCsr Ai = A;
    Ai.rp = sg.a_rp;
    tiled_row(L, scan_tmp, sg.rows[x - sg.start], T, Ai, B, sg.c_rp, sg.c_ci, sg.c_val, split);
  }
}

// ---------------------------------------------------------------------------
// Long rows, wave-independent tiles (KIND_BLOCK variant 1).
//
// Every wave (one per 64-thread workgroup) owns a run of consecutive output
// positions of one row and walks it in tiles of up to Tw positions, its
// accumulators in LDS. For each tile it streams the row's products that
// land in the tile's columns in reference order (A entries ascending, each
// B row's positions ascending), 64 at a time: per chunk of 64 A entries
// every lane narrows its entry's B row to the tile's columns (cursor along
// the run, or a search through the split table), a wave scan lays the
// entries' products end to end, and the entry of every product of a step is
// the running maximum of the entries that start in the step (one LDS mark
// per entry). Products of one entry have distinct columns, so a step drawn
// from a single entry updates its accumulators in one go; a step that spans
// several entries applies products on the same position lowest lane first
// (per round an LDS atomicMin of (position, lane) on a small per-wave hash
// of positions). A wave's LDS operations execute in program order and no
// other wave touches its LDS, so the kernel has no barrier at all. A tile's
// positions are found through a bitmap index of its column span (32 Tw
// columns at most) or, for sparser tiles, an LDS hash of its columns.
constexpr int kWaveTags = 128;   // per-wave hash of pending positions
constexpr int kWaveItems = 4;    // 64-product sub-steps per step
constexpr int kWaveCur = 16;     // cursor chunks: A rows up to 1024 entries
constexpr int kWavePosBits = 11; // positions per tile < 2^11
constexpr int kWaveHashSpan = 1 << (31 - kWavePosBits);  // hash tiles' column span

int wave_tile_lds_bytes(int Tw) {
  // vals[Tw] | index (Tw uint2 words, or 2 Tw hash slots) | tags | marks
  return (3 * Tw + kWaveTags + kWave) * 4;
}

__global__ __launch_bounds__(kWave) void k_numeric_wave(
    const int* __restrict__ rows, int nrows, int Tw, DevCsr A, DevCsr B,
    const int* __restrict__ c_rp, const int* __restrict__ c_ci,
    float* __restrict__ c_val, SplitTable split) {
  extern __shared__ __attribute__((aligned(16))) int wave_lds[];
  float* vals = reinterpret_cast<float*>(wave_lds);
  uint2* idx = reinterpret_cast<uint2*>(wave_lds + Tw);  // Tw is even: 8-byte aligned
  int* hsh = wave_lds + Tw;                              // 2 Tw slots, same bytes as idx
  int* tags = wave_lds + 3 * Tw;
  int* mark = tags + kWaveTags;
  const int lane = threadIdx.x;
  const int hmask = 2 * Tw - 1;         // Tw is a power of two
  const int hbits = 31 - __clz(2 * Tw);  // 2 Tw == 1 << hbits
  for (int i = lane; i < kWaveTags; i += kWave) tags[i] = INT_MAX;
  for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
    const int row = rows[r];
    const int c_s = c_rp[row];
    const int n = c_rp[row + 1] - c_s;
    const int runs = min((int)gridDim.y, n / Tw + (n % Tw != 0));
    if ((int)blockIdx.y >= runs) continue;
    const int rq = n / runs, rr = n % runs, y = blockIdx.y;
    const int P0 = rq * y + rr * y / runs;
    const int P1 = rq * (y + 1) + rr * (y + 1) / runs;
    const int a_s = A.rp[row], a_e = A.rp[row + 1];
    const bool cursors = a_e - a_s <= kWaveCur * kWave;
    int cur[kWaveCur];
    bool first = true;
    for (int p0 = P0; p0 < P1;) {
      int pe = min(P1, p0 + Tw);
      const int clo = c_ci[c_s + p0];
      if (c_ci[c_s + pe - 1] - clo >= kWaveHashSpan) {
        // Too wide even for the hash's column offsets: fewer positions.
        int lo = p0 + 1, hi = pe;  // answer in [lo, hi]: first past the span
        while (lo < hi) {
          const int m = (lo + hi) >> 1;
          if (c_ci[c_s + m] - clo < kWaveHashSpan) lo = m + 1; else hi = m;
        }
        pe = lo;
      }
      const int tn = pe - p0;
      const int chi = c_ci[c_s + pe - 1];
      const bool use_bm = chi - clo < 32 * Tw;  // uniform
      const bool narrow = !(p0 == 0 && pe == n);
      const int iw = ((chi - clo) >> 5) + 1;
      if (use_bm) {
        for (int w = lane; w < iw; w += kWave) idx[w] = make_uint2(0u, 0u);
      } else {
        for (int w = lane; w <= hmask; w += kWave) hsh[w] = -1;
      }
      for (int j = lane; j < tn; j += kWave) vals[j] = 0.0f;
      if (use_bm) {
        for (int j = lane; j < tn; j += kWave) {
          const int d = c_ci[c_s + p0 + j] - clo;
          atomicOr(&idx[d >> 5].x, 1u << (d & 31));
        }
        // Positions before each word: per lane a run of consecutive words.
        const int per = (iw + kWave - 1) / kWave;
        const int w_beg = min(iw, lane * per), w_end = min(iw, w_beg + per);
        int cnt = 0;
        for (int w = w_beg; w < w_end; ++w) cnt += __popc(idx[w].x);
        int run = group_incl_scan<kWave>(cnt) - cnt;
        for (int w = w_beg; w < w_end; ++w) {
          idx[w].y = (unsigned)run;
          run += __popc(idx[w].x);
        }
      } else {
        for (int j = lane; j < tn; j += kWave) {
          const int d = c_ci[c_s + p0 + j] - clo;
          const int key = (d << kWavePosBits) | j;
          int slot = hash_slot(d, hbits);
          while (atomicCAS(&hsh[slot], -1, key) != -1) slot = (slot + 1) & hmask;
        }
      }
      int chunk = 0;
      for (int cb = a_s; cb < a_e; cb += kWave, ++chunk) {
        const int ap = cb + lane;
        int bs = 0, len = 0;
        float a = 0.0f;
        if (ap < a_e) {
          const int k = A.ci[ap];
          a = A.val[ap];
          int be = B.rp[k + 1];
          if (narrow) {
            const int b0 = B.rp[k];
            int so = -1;
            int2 meta = make_int2(0, 0);
            const bool resume = cursors && !first;
            int lo = 0;
            if (resume) {
#pragma unroll
              for (int c = 0; c < kWaveCur; ++c)
                if (c == chunk) lo = cur[c];
            } else {
              if (split.off != nullptr && be - b0 >= kSplitMin && split.off[k + 1] > split.off[k]) {
                so = split.off[k];
                meta = *reinterpret_cast<const int2*>(split.sp + so);
              }
              lo = so >= 0 ? split_lower_bound(B.ci, split.sp + so, meta, b0, be, b0, clo)
                           : gm_lower_bound(B.ci, b0, be, clo);
            }
            bs = lo;
            // The columns at lo .. lo + kProbe - 1, read together, end most
            // ranges without a search.
            int n_in = 0;
            if (lo < be) {
#pragma unroll
              for (int j = 0; j < kProbe; ++j) {
                const int cj = B.ci[min(lo + j, be - 1)];
                n_in += (lo + j < be && cj <= chi) ? 1 : 0;
              }
            }
            if (n_in < kProbe)
              be = lo + n_in;
            else if (so >= 0)
              be = split_lower_bound(B.ci, split.sp + so, meta, b0, be, lo + kProbe, chi + 1);
            else
              be = gm_gallop(B.ci, lo + kProbe, be, chi + 1);
            if (cursors) {
#pragma unroll
              for (int c = 0; c < kWaveCur; ++c)
                if (c == chunk) cur[c] = be;
            }
          } else {
            bs = B.rp[k];
          }
          len = be - bs;
        }
        const int incl = group_incl_scan<kWave>(len);
        const int excl = incl - len;
        const int total = __shfl(incl, kWave - 1);
        int carry = -1;  // entry of the previous step's last product
        // The entry, B position and scale of product q0 + lane.
        auto locate = [&](int q0, int& e, int& bi, float& av) {
          mark[lane] = -1;
          if (len > 0 && excl >= q0 && excl < q0 + kWave) mark[excl - q0] = lane;
          wave_lds_order();  // other lanes' marks are read below
          const int m = mark[lane];
          wave_lds_order();  // ... before the next step clears them
          e = max(group_incl_max<kWave>(m), carry);
          carry = __shfl(e, kWave - 1);
          bi = __shfl(bs, e) + (q0 + lane - __shfl(excl, e));
          av = __shfl(a, e);
        };
        // Steps of kWaveItems x 64 products; the next step's gathers are
        // issued before this step's lookups and adds, so up to 2 kWaveItems
        // products per lane are in flight.
        int e[kWaveItems], col[kWaveItems], ne[kWaveItems], ncol[kWaveItems];
        float av[kWaveItems], bv[kWaveItems], nav[kWaveItems], nbv[kWaveItems];
        auto fetch = [&](int q0, int* fe, int* fc, float* fa, float* fb) {
#pragma unroll
          for (int i = 0; i < kWaveItems; ++i) {
            const int q = q0 + i * kWave;
            int bi = 0;
            fe[i] = carry;
            fc[i] = -1;
            fa[i] = fb[i] = 0.0f;
            if (q < total) {  // uniform
              locate(q, fe[i], bi, fa[i]);
              if (q + lane < total) {
                fc[i] = B.ci[bi];
                fb[i] = B.val[bi];
              }
            }
          }
        };
        fetch(0, e, col, av, bv);
        for (int q0 = 0; q0 < total; q0 += kWaveItems * kWave) {
          if (q0 + kWaveItems * kWave < total) fetch(q0 + kWaveItems * kWave, ne, ncol, nav, nbv);
#pragma unroll
          for (int i = 0; i < kWaveItems; ++i) {
            const int qs = q0 + i * kWave;
            if (qs >= total) break;  // uniform
            const bool valid = qs + lane < total;
            int pos = -1;
            float prod = 0.0f;
            if (valid) {
              const int d = col[i] - clo;
              if (use_bm) {
                const uint2 v = idx[d >> 5];
                pos = (int)v.y + __popc(v.x & ((1u << (d & 31)) - 1u));
              } else {
                int slot = hash_slot(d, hbits);
                int v = hsh[slot];
                while ((v >> kWavePosBits) != d) {
                  slot = (slot + 1) & hmask;
                  v = hsh[slot];
                }
                pos = v & ((1 << kWavePosBits) - 1);
              }
              prod = av[i] * bv[i];  // rounded here, added below
            }
            const int last = min(kWave, total - qs) - 1;
            if (__shfl(e[i], 0) == __shfl(e[i], last)) {
              // One entry: distinct positions.
              if (valid) vals[pos] = vals[pos] + prod;
            } else {
              bool act = valid;
              const int key = (pos << 6) | lane;
              const int slot = (int)(((unsigned)pos * 2654435761u) >> 25);  // kWaveTags = 128
              while (__any(act)) {
                if (act) atomicMin(&tags[slot], key);
                if (act && tags[slot] == key) {
                  vals[pos] = vals[pos] + prod;
                  tags[slot] = INT_MAX;
                  act = false;
                }
              }
            }
          }
#pragma unroll
          for (int i = 0; i < kWaveItems; ++i) {
            e[i] = ne[i];
            col[i] = ncol[i];
            av[i] = nav[i];
            bv[i] = nbv[i];
          }
        }
      }
      for (int j = lane; j < tn; j += kWave)
        __builtin_nontemporal_store(vals[j], c_val + c_s + p0 + j);
      p0 = pe;
      first = false;
    }
  }
}

int tiled_lds_bytes(int T) {
  return kTiledItems * kTiledWaves * kTiledWaves * 4 + kTiledGroup * 8 +
         tiled_index_words(T) * 8 + ((T + 1) & ~1) * 4 + kTiledWaves * kTagSlots * 4 +
         (kEntrySlots + 3 * kBlock + kPrePad) * 4;
}

template <int G>
static hipError_t launch_numeric_g(const BinSpec& b, const int* rows, int nrows,
                                   const DevCsr& A, const DevCsr& B,
                                   const int* c_rp, const int* c_ci,
                                   float* c_val, hipStream_t s) {
  const int groups = kBlock / G;
  const int grid = (nrows + groups - 1) / groups;
  if (b.kind == KIND_DIRECT) {
    hipLaunchKernelGGL(k_numeric_direct<G>, dim3(grid), dim3(kBlock), 0, s,
                       rows, nrows, A, B, c_rp, c_val);
  } else {
    hipLaunchKernelGGL(k_numeric_search<G>, dim3(grid), dim3(kBlock),
                       b.lds_bytes, s, rows, nrows, b.cap, A, B, c_rp, c_ci,
                       c_val);
  }
  return hipGetLastError();
}

hipError_t launch_numeric_tiled_fused(const TiledSegs& segs, int T, int max_row, const DevCsr& A,
                                      const DevCsr& B, const SplitTable& split, int lds_bytes,
                                      hipStream_t s) {
  if (segs.n <= 0 || segs.total <= 0) return hipSuccess;
  const long long most = max_row > 0 ? max_row : 1;
  const int ytiles = (int)std::min<long long>(kTiledRuns, (most + T - 1) / T);
  const int gx = (std::min(segs.total, 1 << 20) + 7) / 8 * 8;
  hipLaunchKernelGGL(k_numeric_tiled_fused, dim3(gx, ytiles), dim3(kBlock), lds_bytes, s, segs,
                     T, A, B, split);
  return hipGetLastError();
}

hipError_t launch_numeric_bin(const BinSpec& b, const int* rows, int nrows,
                              const DevCsr& A, const DevCsr& B, RowStats st,
                              const int* c_rp, const int* c_ci, float* c_val,
                              const GlobalSlots& gs, int max_row, const SplitTable& split,
                              hipStream_t s) {
  (void)st;
  if (nrows <= 0) return hipSuccess;
  switch (b.kind) {
    case KIND_DIRECT:
    case KIND_SEARCH:
      switch (b.group) {
        case 1: return launch_numeric_g<1>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        case 2: return launch_numeric_g<2>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        case 4: return launch_numeric_g<4>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        case 8: return launch_numeric_g<8>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        case 16: return launch_numeric_g<16>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        case 32: return launch_numeric_g<32>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        case 64: return launch_numeric_g<64>(b, rows, nrows, A, B, c_rp, c_ci, c_val, s);
        default: return hipErrorInvalidValue;
      }
    case KIND_BLOCK: {
      // Tiles of one row get the same x, so they share an XCD (and its L2)
      // when the x extent is a multiple of the XCD count.
      const long long most = max_row > 0 ? max_row : 1;
      const int ytiles = (int)std::min<long long>(kTiledRuns, (most + b.cap - 1) / b.cap);
      const int gx = (std::min(nrows, 1 << 20) + 7) / 8 * 8;
      if (b.variant == 1) {
        // Wave tiles: up to kWaveRuns runs per row, one wave each.
        const int wy = (int)std::min<long long>(kWaveRuns, (most + b.cap - 1) / b.cap);
        hipLaunchKernelGGL(k_numeric_wave, dim3(gx, wy), dim3(kWave), b.lds_bytes, s,
                           rows, nrows, b.cap, A, B, c_rp, c_ci, c_val, split);
        return hipGetLastError();
      }
      hipLaunchKernelGGL(k_numeric_tiled, dim3(gx, ytiles), dim3(kBlock), b.lds_bytes, s,
                         rows, nrows, b.cap, A, B, c_rp, c_ci, c_val, split);
      return hipGetLastError();
    }
    case KIND_GLOBAL:
      if (!gs.base || gs.nslots <= 0) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_numeric_rowblock<true>, dim3(min(nrows, gs.nslots)),
                         dim3(kBlock), 0, s, rows, nrows, 0, A, B, c_rp, c_ci,
                         c_val, (float*)gs.base, gs.slot_bytes / sizeof(float));
      return hipGetLastError();
    default:
      return hipErrorInvalidValue;
  }
}

__global__ __launch_bounds__(kBlock) void k_split_counts(DevCsr B, int* __restrict__ cnt) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= B.rows) return;
  const int len = B.rp[k + 1] - B.rp[k];
  cnt[k] = len >= kSplitMin ? (2 + split_granules(len) + 1) & ~1 : 0;
}

hipError_t launch_split_counts(const DevCsr& B, int* cnt, hipStream_t s) {
  if (B.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_counts, dim3((B.rows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     B, cnt);
  return hipGetLastError();
}

// One wave per row with a table: every boundary by its own binary search.
__global__ __launch_bounds__(kBlock) void k_split_build(DevCsr B, SplitTable t,
                                                        int* __restrict__ sp) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int nwaves = gridDim.x * (kBlock / kWave);
  for (int k = wave; k < B.rows; k += nwaves) {
    const int so = t.off[k];
    if (t.off[k + 1] == so) continue;
    const int s = B.rp[k], e = B.rp[k + 1];
    const int nb = split_granules(e - s);
    const int lo = B.ci[s], hi = B.ci[e - 1];
    int g = 0;
    while (((hi - lo) >> g) >= nb) ++g;
    if (lane == 0) {
      sp[so] = lo;
      sp[so + 1] = g | (nb << 5);
    }
    for (int j = 1 + lane; j <= nb; j += kWave) {
      const long long key = (long long)lo + ((long long)j << g);
      sp[so + 1 + j] = key > INT_MAX ? e : gm_lower_bound(B.ci, s, e, (int)key);
    }
  }
}

hipError_t launch_split_build(const DevCsr& B, const SplitTable& t, int* sp, hipStream_t s) {
  if (B.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_build, dim3(2048), dim3(kBlock), 0, s, B, t, sp);
  return hipGetLastError();
}

// Fill of rows whose sorted columns the count phase kept: one workgroup per
// row copies them to C.
__global__ __launch_bounds__(kBlock) void k_copy_kept(const int* __restrict__ rows, int nrows,
                                                      const int* __restrict__ keep,
                                                      const long long* __restrict__ koff,
                                                      const int* __restrict__ c_rp,
                                                      int* __restrict__ c_ci) {
  // Kept rows are long: gridDim.y workgroups share each row.
  for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
    const int row = rows[r];
    const int c_s = c_rp[row];
    const int n = c_rp[row + 1] - c_s;
    const int* src = keep + koff[row];
    const int per = (n + gridDim.y - 1) / gridDim.y;
    const int j0 = blockIdx.y * per, j1 = min(n, j0 + per);
    // Four loads in flight per thread before their stores.
    constexpr int kU = 4;
    int j = j0 + threadIdx.x;
    for (; j + (kU - 1) * kBlock < j1; j += kU * kBlock) {
      int v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = src[j + u * kBlock];
#pragma unroll
      for (int u = 0; u < kU; ++u) c_ci[c_s + j + u * kBlock] = v[u];
    }
    for (; j < j1; j += kBlock) c_ci[c_s + j] = src[j];
  }
}

hipError_t launch_copy_kept(const int* rows, int nrows, const int* keep,
                            const long long* koff, const int* c_rp, int* c_ci,
                            hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_kept, dim3(row_grid(nrows), 8), dim3(kBlock), 0, s, rows, nrows,
                     keep, koff, c_rp, c_ci);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------

template <typename K>
static hipError_t allow_lds(K kernel, int bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

hipError_t init_kernel_attributes(int lds_bytes) {
  hipError_t e = hipSuccess;
#define RVSP_ALLOW(k) \
  if ((e = allow_lds(k, lds_bytes)) != hipSuccess) return e;
  RVSP_ALLOW(k_count_hash<1>) RVSP_ALLOW(k_count_hash<2>)
  RVSP_ALLOW(k_count_hash<4>) RVSP_ALLOW(k_count_hash<8>)
  RVSP_ALLOW(k_count_hash<16>) RVSP_ALLOW(k_count_hash<32>)
  RVSP_ALLOW(k_count_hash<64>)
  RVSP_ALLOW(k_fill_hash<1>) RVSP_ALLOW(k_fill_hash<2>)
  RVSP_ALLOW(k_fill_hash<4>) RVSP_ALLOW(k_fill_hash<8>)
  RVSP_ALLOW(k_fill_hash<16>) RVSP_ALLOW(k_fill_hash<32>)
  RVSP_ALLOW(k_fill_hash<64>)
  RVSP_ALLOW(k_numeric_search<1>) RVSP_ALLOW(k_numeric_search<2>)
  RVSP_ALLOW(k_numeric_search<4>) RVSP_ALLOW(k_numeric_search<8>)
  RVSP_ALLOW(k_numeric_search<16>) RVSP_ALLOW(k_numeric_search<32>)
  RVSP_ALLOW(k_numeric_search<64>)
  RVSP_ALLOW((k_bitmap_row<false, kBitmapCount>))
  RVSP_ALLOW((k_bitmap_row<false, kBitmapFill>))
  RVSP_ALLOW((k_bitmap_row<false, kBitmapCountKeep>))
  RVSP_ALLOW((k_bitmap_row<false, kBitmapCount, kWideBlock>))
  RVSP_ALLOW((k_bitmap_row<false, kBitmapFill, kWideBlock>))
  RVSP_ALLOW((k_bitmap_row<false, kBitmapCountKeep, kWideBlock>))
  RVSP_ALLOW(k_numeric_rowblock<false>) RVSP_ALLOW(k_numeric_tiled) RVSP_ALLOW(k_numeric_wave)
  RVSP_ALLOW(k_numeric_tiled_fused)
#define RVSP_ALLOW_MAP(G)                                                     \
  RVSP_ALLOW((k_numeric_map<G, unsigned char, 0>))                             \
  RVSP_ALLOW((k_numeric_map<G, unsigned char, 1>))                             \
  RVSP_ALLOW((k_numeric_map<G, unsigned char, 2>))                             \
  RVSP_ALLOW((k_numeric_map<G, unsigned short, 0>))                            \
  RVSP_ALLOW((k_numeric_map<G, unsigned short, 1>))                            \
  RVSP_ALLOW((k_numeric_map<G, unsigned short, 2>))
  RVSP_ALLOW_MAP(1) RVSP_ALLOW_MAP(2) RVSP_ALLOW_MAP(4) RVSP_ALLOW_MAP(8)
  RVSP_ALLOW_MAP(16) RVSP_ALLOW_MAP(32) RVSP_ALLOW_MAP(64)
#undef RVSP_ALLOW_MAP
#undef RVSP_ALLOW
  return e;
}

}  // namespace rvsp
