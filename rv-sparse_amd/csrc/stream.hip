// stream.hip — persistent row-stream numeric pass.
//
// One launch serves the numeric bins of one or several descriptors (all
// groups of a batch). A bin is a segment: its rows in list order, whose
// products, laid end to end, form the segment's stream. The streams are cut
// into queue items of kStreamItem products; every wave of the launch takes
// items from one queue until none is left, so the work is balanced by
// products across all bins and groups, and the launch has no per-bin tail.
//
// A wave walks the rows of its item in sub-batches (up to 64 rows whose C
// rows fit the wave's LDS accumulators) and their products 64 at a time,
// across A entries and rows: lane l of a step takes product 64 w + l of the
// stream (window w), finds its A entry by the entry marks of the step and a
// running maximum, and gathers one B value. Its output position comes from
// the segment's map (built once per structure by the same traversal, with a
// binary search of the product's column in the C row's sorted columns).
//
// Order. The reference adds the products of C(i,j) in A order, one rounded
// add at a time (gustavson_core_f32.inc:18-64, accum_scalar_f32.c:67-70). A
// step holds products of several A entries, so two of its lanes may hit the
// same C entry. Every map entry carries its order among the earlier products
// of the same C entry in the same window (computed once, k_stream_rounds);
// a step applies its lanes in rounds of that order, which is one round when
// no C entry repeats in the window. Lanes of one round have distinct
// positions; windows follow each other in stream order, and a row belongs
// to one wave. The result is bit-identical to every other numeric kernel.
#include <climits>

#include "common.hpp"
#include "kernels.hpp"

namespace rvsp {

namespace {

constexpr unsigned kPosMask = (1u << kStreamPosBits) - 1u;
constexpr unsigned kRowMark = 1u << kStreamPosBits;
constexpr int kOrderShift = kStreamPosBits + 1;
constexpr unsigned kLaneOrder = 7;  // order value: apply in lane order
static_assert(kOrderShift + 3 == 16, "the order takes the entry's top three bits");

__global__ __launch_bounds__(kBlock) void k_stream_weights(const int* __restrict__ rows,
                                                           int row_base, int nrows,
                                                           const long long* __restrict__ ops,
                                                           long long* __restrict__ w) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nrows) return;
  w[i] = ops[rows ? rows[i] : row_base + i];
}

// first[k] = the first list position whose stream offset is >= k * item,
// for k in [0, items]; first[items] = nrows.
__global__ __launch_bounds__(kBlock) void k_stream_items(const long long* __restrict__ soff,
                                                         int nrows, int items,
                                                         int* __restrict__ first) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i > nrows) return;
  const long long lo = i > 0 ? soff[i - 1] : -1;
  const long long k0 = lo < 0 ? 0 : lo / kStreamItem + 1;
  long long k1 = i == nrows ? items : soff[i] / kStreamItem;
  if (k1 > items) k1 = items;
  for (long long k = k0; k <= k1; ++k) first[k] = i;
}

__global__ __launch_bounds__(kBlock) void k_stream_mark(const int* __restrict__ rows,
                                                        int row_base, int nrows,
                                                        int* __restrict__ flags) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nrows) flags[rows ? rows[i] : row_base + i] = 1;
}

__global__ __launch_bounds__(kBlock) void k_stream_compact(const int* __restrict__ flags,
                                                           const int* __restrict__ pos, int M,
                                                           int* __restrict__ list) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < M && flags[i]) list[pos[i]] = i;
}

// Wave-wide inclusive scans on DPP lane moves (no LDS round trips): row
// shifts by 1, 2, 4 and 8 scan each row of 16 lanes, then the last lane of
// rows 0 and 2 is broadcast into rows 1 and 3, and lane 31 into rows 2 and 3.
// Lanes whose source lies outside the row keep the operation's identity.
__device__ __forceinline__ int wave_incl_max(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x111, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x112, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x114, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x118, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xc, 0xf, false));
  return v;
}

__device__ __forceinline__ int wave_incl_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ int wave_last(int v) { return __builtin_amdgcn_readlane(v, kWave - 1); }

// Global-memory views of the segment arrays (read from the segment table,
// so the compiler cannot tell their address space): global loads and stores
// count only against the vector-memory counter, so waits for LDS results do
// not also wait for them.
template <typename T>
using gp = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gp<const T> gload(const T* p) {
  return (gp<const T>)p;
}
template <typename T>
__device__ __forceinline__ gp<T> gstore(T* p) {
  return (gp<T>)p;
}

// Lanes with `has` whose `start` lies in [base, base + 64) mark that slot
// with their lane id; every lane returns the last mark at or before its own
// slot, or `carry` (the owner of the previous window's last slot). `mark`
// holds 128 slots: lanes that mark nothing write into the upper half, so the
// store needs no branch.
__device__ __forceinline__ int wave_owner(int* mark, bool has, int start, int base, int carry) {
  const int lane = threadIdx.x & (kWave - 1);
  mark[lane] = -1;
  wave_lds_order();
  const unsigned rel = (unsigned)(start - base);
  mark[has && rel < (unsigned)kWave ? (int)rel : kWave + lane] = lane;
  wave_lds_order();
  const int m = mark[lane];
  wave_lds_order();
  return max(wave_incl_max(m), carry);
}

// One step's adds into the accumulators, in the rounds of the map's order
// bits (lanes of one round have distinct positions). `prod` is the rounded
// product a * b, taken by every lane before any branch so that the loads it
// reads are consumed on every path (a load left pending past the loop's end
// makes the compiler wait for all loads in flight there).
__device__ __forceinline__ void stream_apply(float* vals, bool pv, int at, unsigned ord,
                                             float prod) {
#pragma clang fp contract(off)
  if (__ballot(pv && ord != 0) == 0) {
    if (pv) vals[at] = vals[at] + prod;
    return;
  }
#pragma unroll 1
  for (unsigned r = 0; r < kLaneOrder; ++r) {
    if (__ballot(pv && ord >= r) == 0) return;
    if (pv && ord == r) vals[at] = vals[at] + prod;
    wave_lds_order();
  }
  const int lane = threadIdx.x & (kWave - 1);
  unsigned long long left = __ballot(pv && ord == kLaneOrder);
  while (left) {
    const int l = __ffsll((long long)left) - 1;
    if (lane == l) vals[at] = vals[at] + prod;
    wave_lds_order();
    left &= left - 1;
  }
}

// MODE kStreamMapPass: positions from the map. kStreamBuildPass: writes the
// map (positions by binary search in the rows' C columns). kStreamSearchPass:
// positions by the same search, adds ordered at run time (no map needed).
// The search pass's adds: products of one window that hit the same
// accumulator apply in lane order (= product order). Each round, every
// pending lane proposes its lane id to its position's tag (an LDS atomic
// minimum); the lowest proposer of each position adds and empties the tag.
// One round when no position repeats in the window.
__device__ __forceinline__ void search_apply(float* vals, int* tags, bool use, int at, float a,
                                             float b) {
  const int lane = threadIdx.x & (kWave - 1);
  bool act = use;
  while (__ballot(act)) {
    if (act) atomicMin(&tags[at], lane);
    wave_lds_order();
    const bool won = act && tags[at] == lane;
    if (won) {
      vals[at] = acc_step(vals[at], a, b);
      tags[at] = kWave;
    }
    wave_lds_order();
    act = act && !won;
  }
}

template <int MODE>
__global__ __launch_bounds__(kWave) void k_stream(const StreamSeg* __restrict__ segs, int nseg,
                                                  int items, DevCsr A, DevCsr B,
                                                  StreamQueue* __restrict__ q, int capw) {
  constexpr bool BUILD = MODE == kStreamBuildPass;
  constexpr bool SEARCH = MODE == kStreamSearchPass;
  extern __shared__ __attribute__((aligned(16))) int st_lds[];
  const int cap4 = (capw + 3) & ~3;
  // Numeric passes: accumulators of the sub-batch; the build: its rows' C
  // columns; the search pass: accumulators, columns and one order tag per
  // position.
  float* vals = reinterpret_cast<float*>(st_lds);
  int* cols = BUILD ? st_lds : st_lds + cap4;
  int* tags = st_lds + 2 * cap4;
  int* mark = st_lds + (BUILD ? 1 : SEARCH ? 3 : 1) * cap4;  // 2 kWave slots
  // Per A entry of the current chunk: B position of its product at chunk
  // offset 0, scale, and accumulator (or C) offset of its row.
  int4* etab = reinterpret_cast<int4*>(mark + 2 * kWave);
  const int lane = threadIdx.x;
  constexpr int kSegLanes = kStreamMaxSegs / kWave;
  unsigned bad = 0;  // kStreamBad* bits of refused writes
  int seg_i0[kSegLanes];
#pragma unroll
  for (int j = 0; j < kSegLanes; ++j) {
    const int sg = lane + j * kWave;
    seg_i0[j] = sg < nseg ? segs[sg].item0 : INT_MAX;
  }
  const int x0 = blockIdx.x & (kStreamQueues - 1);
  int tried = 0;  // queues found empty (own XCD's first)
  for (;;) {
    int item = -1;
    while (tried < kStreamQueues) {
      const int x = (x0 + tried) & (kStreamQueues - 1);
      const int lo = (int)((long long)items * x / kStreamQueues);
      const int hi = (int)((long long)items * (x + 1) / kStreamQueues);
      int t = 0;
      if (lane == 0) t = (int)atomicAdd(&q->next[x], 1u);
      t = __builtin_amdgcn_readfirstlane(t);
      if (t < hi - lo) {
        item = lo + t;
        break;
      }
      ++tried;
    }
    if (item < 0) break;
    int sidx = -1;
#pragma unroll
    for (int j = 0; j < kSegLanes; ++j) sidx += __popcll(__ballot(seg_i0[j] <= item));
    sidx = __builtin_amdgcn_readfirstlane(sidx);
    const StreamSeg& S = segs[sidx];
    const bool ident = S.ident != 0;
    if (BUILD && ident) continue;
    const gp<const int> rows = gload(S.rows);
    const bool listed = S.rows != nullptr;
    const int row_base = S.row_base;
    const gp<const int> a_rp = gload(S.a_rp);
    const gp<const int> c_rp = gload(S.c_rp);
    const gp<const long long> soff = gload(S.soff);
    const gp<unsigned short> map = gstore(S.map);
    const gp<float> c_val = gstore(S.c_val);
    const long long total = S.total;
    const int c_nnz = S.c_nnz;
    const int k = item - S.item0;
    int i = gload(S.item_first)[k];
    const int i_end = gload(S.item_first)[k + 1];
    while (i < i_end) {
      // One lane per row of the sub-batch.
      const int ii = i + lane;
      const bool rv = ii < i_end;
      int row = 0, cs = 0, n = 0, as = 0, ae = 0;
      long long rq = 0;
      if (rv) {
        row = listed ? rows[ii] : row_base + ii;
        cs = c_rp[row];
        n = c_rp[row + 1] - cs;
        as = a_rp[row];
        ae = a_rp[row + 1];
        rq = soff[ii];
      }
      const int incl_n = wave_incl_sum(n);
      // Rows of the sub-batch: as many as the accumulators hold. A row longer
      // than the accumulators (numeric pass only; the build gets room for
      // every row) is a sub-batch of its own, walked once per range of capw
      // positions, each walk applying only the products of its range.
      int R = ident ? min(kWave, i_end - i) : __popcll(__ballot(rv && incl_n <= capw));
      R = __builtin_amdgcn_readfirstlane(R);
      int npass = 1;
      if (R == 0) {
        if (BUILD) {  // never planned: skipped rather than overrun
          ++i;
          continue;
        }
        R = 1;
        npass = (__builtin_amdgcn_readfirstlane(n) + capw - 1) / capw;
      }
      const bool rl = lane < R;
      if (!rl) {
        n = 0;
        as = ae = 0;
      }
      const int abase = incl_n - n;  // accumulator offset of the row (rows < R)
      const int totn = ident ? 0 : __shfl(incl_n, R - 1);
      const int ne = ae - as;
      const int incl_e = wave_incl_sum(ne);
      const int ebase = incl_e - ne;
      const int E = wave_last(incl_e);
      for (int pass = 0; pass < npass; ++pass) {
        const int p0 = pass * capw;  // first position of this walk
        const int pn = npass > 1 ? min(capw, totn - p0) : totn;  // accumulators in use
        int clo = 0, chi = 0;  // search pass over one long row: columns of this walk
        if (!ident) {
          // Accumulators cleared (numeric passes), the rows' columns of this
          // walk staged (build, search) and the order tags emptied (search).
          int rc = 0;
          for (int j0 = 0; j0 < pn; j0 += kWave) {
            const int j = j0 + lane;
            if (!BUILD && j < pn) vals[j] = 0.0f;
            if (SEARCH && j < pn) tags[j] = kWave;
            if (!BUILD && !SEARCH) continue;
            const int r = wave_owner(mark, rl && n > 0, abase, j0, rc);
            rc = wave_last(r);
            const int src = __shfl(cs, r) + p0 + j - __shfl(abase, r);
            if (j < pn) cols[j] = gload(S.c_ci)[src];
          }
          wave_lds_order();
          if (SEARCH && npass > 1) {
            clo = cols[0];
            chi = cols[pn - 1];
          }
        }
        // Stream index of the chunk's first product (uniform: lane 0's row).
        long long qc = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane(
                                        (int)((unsigned long long)rq >> 32))) << 32 |
                                    (unsigned)__builtin_amdgcn_readfirstlane((int)rq));
        int ec = 0;
        for (int c0 = 0; c0 < E; c0 += kWave) {
          // One lane per A entry of the chunk.
          const int er = wave_owner(mark, rl && ne > 0, ebase, c0, ec);
          ec = wave_last(er);
          const int eidx = c0 + lane;
          const bool ev = eidx < E;
          const int ap = __shfl(as, er) + eidx - __shfl(ebase, er);
          const int e_acc = __shfl(abase, er);
          const int e_n = __shfl(n, er);
          const int e_cs = __shfl(cs, er);
          const long long e_rq = __shfl(rq, er);
          int bs = 0, len = 0;
          float a = 0.0f;
          if (ev) {
            const int kk = A.ci[ap];
            a = A.val[ap];
            bs = B.rp[kk];
            len = B.rp[kk + 1] - bs;
          }
          const int incl_p = wave_incl_sum(len);
          const int excl_p = incl_p - len;
          const int boff = bs - excl_p;  // B position of the entry's product at chunk offset t: boff + t
          const int T = wave_last(incl_p);
          if (T > 0) {
            const long long w0 = qc >> 6, w1 = (qc + T - 1) >> 6;
            int pc = 0;
            // Entry, B position and operands of every lane of window w.
            auto locate = [&](long long w, bool& pv, int& bi, float& av, int& p_acc, int& p_n,
                              int& p_cs, long long& p_rq) {
              const int off = (int)((w << 6) - qc);
              const int e = wave_owner(mark, len > 0, excl_p, off, pc);
              pc = wave_last(e);
              const int t = off + lane;
              pv = t >= 0 && t < T;
              bi = __shfl(boff, e) + t;
              av = __shfl(a, e);
              p_acc = __shfl(e_acc, e);
              p_n = __shfl(e_n, e);
              p_cs = __shfl(e_cs, e);
              p_rq = __shfl(e_rq, e);
            };
            if (BUILD) {
              for (long long w = w0; w <= w1; ++w) {
                bool pv;
                int bi, p_acc, p_n, p_cs;
                float av;
                long long p_rq;
                locate(w, pv, bi, av, p_acc, p_n, p_cs, p_rq);
                const long long qq = (w << 6) + lane;
                if (pv) {
                  const int col = B.ci[bi];
                  const unsigned pos = (unsigned)lds_find(cols + p_acc, p_n, col);
                  if (cols[p_acc + pos] != col || pos > kPosMask) bad |= kStreamBadColumn;
                  if (qq < 0 || qq >= total) bad |= kStreamBadMap;
                  else map[qq] = (unsigned short)(pos | (qq == p_rq ? kRowMark : 0u));
                }
              }
            } else {
              // Windows k = 0.. of the chunk: window k starts `lead` products
              // before chunk offset 64 k. Each entry's operands come from the
              // entry table with one LDS read; for map segments the next
              // window's map entry and B value are in flight while this window
              // is applied.
              const long long qa = qc & ~63LL;
              const int lead = (int)(qc - qa);
              const int nwin = (lead + T + 63) >> 6;
              // Map reads cover whole windows, up to three past the chunk's
              // last (the map has 256 entries of slack past the stream); lanes
              // without a product read B position 0.
              const gp<const unsigned short> mc = ident ? nullptr : (gp<const unsigned short>)(map + qa);
            const gp<const char> mc_g = (gp<const char>)mc;
            const gp<const char> b_val_g = (gp<const char>)B.val;
            const gp<const char> b_ci_g = (gp<const char>)B.ci;
              // Identity rows: C position of the product at chunk offset t is
              // base + t, base = the row's C start minus the chunk offset of
              // the row's first product.
              const int e_base = ident ? e_cs - (int)(e_rq - qc) : e_acc;
              // Columns the search pass looks a product up in: its row's
              // (or, walking one long row, this walk's).
              const int e_len = npass > 1 ? pn : e_n;
              etab[lane] = make_int4(boff, __float_as_int(a), e_base, e_len);
              wave_lds_order();
              auto step = [&](int k, bool& pv, int& t, int& bi, float& av, int& base,
                              int& nlen) {
                const int off = (k << 6) - lead;
                const int e = wave_owner(mark, len > 0, excl_p, off, pc);
                pc = wave_last(e);
                t = off + lane;
                pv = t >= 0 && t < T;
                const int4 et = etab[e];
                bi = et.x + t;
                av = __int_as_float(et.y);
                base = et.z;
                nlen = et.w;
              };
              if (ident) {
                for (int k = 0; k < nwin; ++k) {
                  bool pv;
                  int t, bi, base, nlen;
                  float av;
                  step(k, pv, t, bi, av, base, nlen);
                  const int dst = base + t;
                  if (pv && (dst < 0 || dst >= c_nnz)) bad |= kStreamBadValue;
                  else if (pv) __builtin_nontemporal_store(acc_step(0.0f, av, B.val[bi]), &c_val[dst]);
                }
              } else {
                // Windows k + 1 and k + 2 are in flight while window k is
                // applied. Their loads are unconditional, also past the last
                // window (no products there; map entries in the slack), so no
                // branch makes the compiler wait for them early.
                // x: the aligned map dword holding the product's entry (map
                // pass; the entry is its half msel) or B column (search).
                struct Win {
                  bool pv;
                  int acc, nlen;
                  float av, bv;
                  unsigned x;
                };
                // Loads address a uniform base plus a 32-bit unsigned byte
                // offset (global loads with a scalar base, no 64-bit adds).
                // Byte offset of the lane's map dword within a window, and
                // the shift of its entry in that dword.
                const unsigned mlane = (unsigned)(lane & ~1) * 2u;
                const unsigned msel = (unsigned)(lane & 1) * 16u;
                auto issue = [&](int k) {
                  Win w;
                  int t, bi;
                  step(k, w.pv, t, bi, w.av, w.acc, w.nlen);
                  const unsigned bq = w.pv ? (unsigned)bi * 4u : 0u;
                  // The map streams through once per compute: a non-temporal
                  // load, so it does not evict the B rows the L2 reuses. Two
                  // lanes read the dword of their two entries: a 32-bit value
                  // carried to the window's apply as loaded (a 16-bit one
                  // is widened where the loop repeats, which waits for it).
                  w.x = SEARCH ? (unsigned)*(gp<const int>)(b_ci_g + bq)
                               : __builtin_nontemporal_load((gp<const unsigned>)(
                                     mc_g + ((unsigned)k << 7) + mlane));
                  w.bv = *(gp<const float>)(b_val_g + bq);
                  return w;
                };
                // Positions beyond the walk's range (or, in one walk, past its
                // accumulators: a map that does not match) are skipped.
                const unsigned plim = (unsigned)pn;
                auto apply = [&](const Win& w) {
#pragma clang fp contract(off)
                  if (SEARCH) {
                    const int col = (int)w.x;
                    const bool use = w.pv && (npass == 1 || (col >= clo && col <= chi));
                    const int pos = lds_find(cols + w.acc, use ? w.nlen : 1, col);
                    const int at = w.acc + pos;
                    if (use && (at >= pn || cols[at] != col)) bad |= kStreamBadColumn;
                    search_apply(vals, tags, use && at < pn, at, w.av, w.bv);
                  } else {
                    const unsigned ent = w.x >> msel;
                    const int at = w.acc + (int)(ent & kPosMask) - p0;
                    const bool use = w.pv && (unsigned)at < plim;
                    stream_apply(vals, use, at, (ent >> kOrderShift) & kLaneOrder, w.av * w.bv);
                  }
                };
                // Unrolled by three so the windows' registers change roles
                // without copies (a copy of a load's destination would wait
                // for the load). The loop has one exit, at its end: with
                // exits between the windows, the compiler sees a path to the
                // loop's start with a window's loads unused and waits there
                // for every load in flight.
                Win wa = issue(0);
                Win wb = issue(1);
                Win wc;
                int k = 0;
                for (; k + 3 <= nwin; k += 3) {
                  wc = issue(k + 2);
                  apply(wa);
                  wa = issue(k + 3);
                  apply(wb);
                  wb = issue(k + 4);
                  apply(wc);
                }
                // Windows k and k + 1 are in flight; one past the chunk's
                // last has no products and adds nothing.
                apply(wa);
                apply(wb);
              }
              wave_lds_order();  // the entry table is rewritten by the next chunk
            }
          }
          qc += T;
        }
        if (!BUILD && !ident) {
          // Accumulators to C in runs of rows that follow each other in C
          // (their accumulators do too), the run's bounds uniform; a walk of
          // one long row writes its range. Every store covers one 64-entry
          // aligned block of C, so only a run's two ends write part of a
          // cache line.
          wave_lds_order();
          for (int r = 0; r < R;) {
            const int r_acc = __builtin_amdgcn_readlane(abase, r);
            const int r_dst = __builtin_amdgcn_readlane(cs, r) + p0;
            int r_n = npass > 1 ? pn : __builtin_amdgcn_readlane(n, r);
            for (++r; r < R && __builtin_amdgcn_readlane(cs, r) == r_dst + r_n; ++r)
              r_n += __builtin_amdgcn_readlane(n, r);
            for (int j = lane - (r_dst & (kWave - 1)); j < r_n; j += kWave) {
              if (j < 0) continue;
              const int dst = r_dst + j;
              if (dst < 0 || dst >= c_nnz) bad |= kStreamBadValue;
              else __builtin_nontemporal_store(vals[r_acc + j], &c_val[dst]);  // written once
            }
          }
          wave_lds_order();
        }
      }  // walks
      i += R;
    }
  }
  unsigned all = bad;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) all |= __shfl_xor(all, d);
  if (lane == 0 && all) atomicOr(&q->bad, all);
  // The last wave out leaves the queue at zero for the next launch.
  if (lane == 0) {
    const unsigned d = atomicAdd(&q->done, 1u);
    if (d == gridDim.x - 1) {
      for (int x = 0; x < kStreamQueues; ++x) atomicExch(&q->next[x], 0u);
      atomicExch(&q->done, 0u);
    }
  }
}

// Order bits of one segment's map: one wave per window of 64 stream
// positions; rows within the window are told apart by the row marks.
__global__ __launch_bounds__(kBlock) void k_stream_rounds(unsigned short* __restrict__ map,
                                                          long long total) {
  const long long qq = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool v = qq < total;
  const unsigned ent = v ? map[qq] : 0u;
  const int rowid = group_incl_scan<kWave>((int)((ent >> kStreamPosBits) & 1u));
  const int key = v ? (int)(((unsigned)rowid << kStreamPosBits) | (ent & kPosMask)) : -1 - lane;
  int before = 0;
#pragma unroll 8
  for (int d = 1; d < kWave; ++d) {
    const int o = __shfl_up(key, d);
    if (lane >= d && o == key) ++before;
  }
  const unsigned ord = before < (int)kLaneOrder ? (unsigned)before : kLaneOrder;
  if (v) map[qq] = (unsigned short)((ent & (kPosMask | kRowMark)) | (ord << kOrderShift));
}

// LDS of one wave: capw positions (rounded to 16 bytes) of accumulators
// (map pass), columns (build) or accumulators, columns and tags (search),
// then the marks and the entry table.
int stream_lds_bytes(int mode, int capw) {
  const int arrays = mode == kStreamSearchPass ? 3 : 1;
  return arrays * ((capw + 3) & ~3) * 4 + 2 * kWave * 4 + kWave * 16;
}

int stream_waves_per_cu(int mode, int lds) {
  int nb = 0;
  hipError_t e;
  if (mode == kStreamBuildPass)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_stream<kStreamBuildPass>, kWave, lds);
  else if (mode == kStreamSearchPass)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_stream<kStreamSearchPass>, kWave, lds);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_stream<kStreamMapPass>, kWave, lds);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 8;
  }
  return nb > 0 ? nb : 1;
}

}  // namespace

hipError_t launch_stream_weights(const int* rows, int row_base, int nrows, const long long* ops,
                                 long long* w, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream_weights, dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     rows, row_base, nrows, ops, w);
  return hipGetLastError();
}

hipError_t launch_stream_items(const long long* soff, int nrows, int items, int* item_first,
                               hipStream_t s) {
  if (nrows < 0 || items <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream_items, dim3(nrows / kBlock + 1), dim3(kBlock), 0, s, soff, nrows,
                     items, item_first);
  return hipGetLastError();
}

int stream_grid(int mode, int capw, int items) {
  if (items <= 0 || capw < 0 || capw > kStreamMaxRow) return 0;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    cus = 256;
  }
  const long long fit = (long long)stream_waves_per_cu(mode, stream_lds_bytes(mode, capw)) * cus;
  return (int)(fit < items ? fit : items);
}

hipError_t launch_stream_mark(const int* rows, int row_base, int nrows, int* flags, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream_mark, dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rows,
                     row_base, nrows, flags);
  return hipGetLastError();
}

hipError_t launch_stream_compact(const int* flags, const int* pos, int M, int* list,
                                 hipStream_t s) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream_compact, dim3((M + kBlock - 1) / kBlock), dim3(kBlock), 0, s, flags,
                     pos, M, list);
  return hipGetLastError();
}

hipError_t launch_stream(int mode, const StreamSeg* segs, int nseg, int items, const DevCsr& A,
                         const DevCsr& B, StreamQueue* q, int capw, int grid, hipStream_t s) {
  if (items <= 0 || nseg <= 0) return hipSuccess;
  if (nseg > kStreamMaxSegs || capw < 0 || capw > kStreamMaxRow || grid <= 0)
    return hipErrorInvalidValue;
  const int lds = stream_lds_bytes(mode, capw);
  if (mode == kStreamBuildPass)
    hipLaunchKernelGGL(k_stream<kStreamBuildPass>, dim3(grid), dim3(kWave), lds, s, segs, nseg,
                       items, A, B, q, capw);
  else if (mode == kStreamSearchPass)
    hipLaunchKernelGGL(k_stream<kStreamSearchPass>, dim3(grid), dim3(kWave), lds, s, segs, nseg,
                       items, A, B, q, capw);
  else
    hipLaunchKernelGGL(k_stream<kStreamMapPass>, dim3(grid), dim3(kWave), lds, s, segs, nseg,
                       items, A, B, q, capw);
  return hipGetLastError();
}

hipError_t launch_stream_rounds(unsigned short* map, long long total, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  const long long blocks = (total + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_stream_rounds, dim3((unsigned)blocks), dim3(kBlock), 0, s, map, total);
  return hipGetLastError();
}

}  // namespace rvsp
