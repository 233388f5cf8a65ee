// plan.cpp — the fixed per-bin strategy table and the rules that keep any
// override launchable.
#include "plan.hpp"

#include <algorithm>

#include "common.hpp"

#include <climits>

namespace rvsp {

namespace {


int g_lds_limit = kMaxLdsBytes - kStaticLdsReserve;  // dynamic part, set per device

bool pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

void set_bin(PhaseTable& t, int phase, int b, int kind, int group, int cap,
             int bound, int span_max) {
  t.rule.bound[b] = bound;
  t.rule.span_max[b] = span_max;
  t.spec[b].kind = kind;
  t.spec[b].group = group;
  t.spec[b].cap = cap;
  t.spec[b].lds_bytes = lds_bytes_for(phase, kind, group, cap);
}

}  // namespace

void set_lds_limit(int bytes) { g_lds_limit = bytes; }
int lds_limit() { return g_lds_limit; }

const char* kind_name(int kind) {
  switch (kind) {
    case KIND_NONE: return "none";
    case KIND_DIRECT: return "direct";
    case KIND_HASH: return "lds_hash";
    case KIND_SEARCH: return "lds_search";
    case KIND_BITMAP: return "lds_bitmap";
    case KIND_BLOCK: return "lds_block";
    case KIND_GLOBAL: return "global_dense";
    case KIND_MAP: return "lds_map";
    default: return "?";
  }
}

int lds_bytes_for(int phase, int kind, int group, int cap, int variant) {
  const bool sub_wave = pow2(group) && group <= kWave;
  long long bytes = -1;
  switch (kind) {
    case KIND_NONE:
    case KIND_GLOBAL:
      bytes = 0;
      break;
    case KIND_DIRECT:
      bytes = sub_wave ? 0 : -1;
      break;
    case KIND_HASH:
      if (!sub_wave || !pow2(cap) || cap < 2 || cap > 8192) break;
      if (phase == PHASE_COUNT)
        bytes = (long long)(kBlock / group) * cap * 4;
      else if (phase == PHASE_FILL)
        bytes = (long long)(kBlock / group) * (cap + cap / 2 + 4) * 4;
      break;
    case KIND_SEARCH:
      if (phase != PHASE_NUMERIC || !sub_wave || cap < 1) break;
      // C columns + accumulators, the segment's C and A row offsets, and
      // the staged A-entry metadata (scale, B-row start, B-row end).
      bytes = (long long)(kBlock / group) * cap * 8 + 2LL * (kBlock / group + 1) * 4 +
              (long long)(kBlock / group) * kSegAPerRow * 12;
      break;
    case KIND_MAP:
      if (phase != PHASE_NUMERIC || !sub_wave || cap < 1 || cap > 65536) break;
      // Accumulators, a 16-byte spare slot, and the staged slice of the map
      // (up to 2-byte entries).
      bytes = (long long)(kBlock / group) * cap * 4 + 16 + map_stage_bytes(kBlock / group, cap);
      if (variant == 1) bytes += (long long)(kBlock / group) * map_stage_a_per_row(cap) * 12;
      else if (variant != 0 && variant != 2) bytes = -1;
      break;
    case KIND_BITMAP:
      if (phase == PHASE_NUMERIC || cap < 32 || (cap & 31)) break;
      bytes = cap / 8;
      break;
    case KIND_BLOCK:
      // Tiles of up to `cap` positions: accumulators, the column index,
      // staging of one group of products, per-wave order tags and one chunk
      // of A-entry metadata.
      if (phase != PHASE_NUMERIC) break;
      if (variant == 1) {
        // Wave tiles: accumulators, index or hash, tags, entry marks.
        if (cap < 256 || cap > 2048 || (cap & (cap - 1))) break;
        bytes = wave_tile_lds_bytes(cap);
      } else if (variant == 0 && cap >= kBlock) {
        bytes = tiled_lds_bytes(cap);
      }
      break;
    default:
      break;
  }
  if (bytes > kMaxLdsBytes - kStaticLdsReserve) return -1;
  return (int)bytes;
}

Strategy fixed_strategy() {
  Strategy s;
  // ---- count: metric = min(Op_i, span_i), an upper bound of the row's nnz.
  {
    PhaseTable& t = s.count;
    int b = 0;
    set_bin(t, PHASE_COUNT, b++, KIND_NONE, 1, 0, 0, INT_MAX);
    set_bin(t, PHASE_COUNT, b++, KIND_DIRECT, 16, 0, 0, INT_MAX);
    const int T[] = {32, 64, 128, 256, 512, 1024, 2048, 4096};
    const int G[] = {4, 8, 16, 32, 64, 64, 64, 64};
    for (int i = 0; i < 8; ++i)
      set_bin(t, PHASE_COUNT, b++, KIND_HASH, G[i], T[i], T[i] / 2, INT_MAX);
    const int span[] = {1 << 15, 1 << 18, 1 << 20};
    for (int i = 0; i < 3; ++i)
      set_bin(t, PHASE_COUNT, b++, KIND_BITMAP, kBlock, span[i], INT_MAX, span[i]);
    set_bin(t, PHASE_COUNT, b++, KIND_GLOBAL, kBlock, 0, INT_MAX, INT_MAX);
    t.rule.nbins = b;
  }
  // ---- fill: metric = exact nnz of the C row.
  {
    PhaseTable& t = s.fill;
    int b = 0;
    set_bin(t, PHASE_FILL, b++, KIND_NONE, 1, 0, 0, INT_MAX);
    set_bin(t, PHASE_FILL, b++, KIND_DIRECT, 16, 0, 0, INT_MAX);
    const int T[] = {32, 64, 128, 256, 512};
    const int G[] = {4, 8, 16, 32, 64};
    for (int i = 0; i < 5; ++i)
      set_bin(t, PHASE_FILL, b++, KIND_HASH, G[i], T[i], T[i] / 2, INT_MAX);
    const int span[] = {1 << 15, 1 << 18, 1 << 20};
    for (int i = 0; i < 3; ++i)
      set_bin(t, PHASE_FILL, b++, KIND_BITMAP, kBlock, span[i], INT_MAX, span[i]);
    set_bin(t, PHASE_FILL, b++, KIND_GLOBAL, kBlock, 0, INT_MAX, INT_MAX);
    t.rule.nbins = b;
  }
  // ---- numeric: metric = exact nnz of the C row.
  {
    PhaseTable& t = s.numeric;
    int b = 0;
    set_bin(t, PHASE_NUMERIC, b++, KIND_NONE, 1, 0, 0, INT_MAX);
    set_bin(t, PHASE_NUMERIC, b++, KIND_DIRECT, 16, 0, 0, INT_MAX);
    const int C[] = {16, 32, 64, 128, 256, 512, 1024, 2048};
    const int G[] = {4, 8, 16, 32, 64, 64, 64, 64};
    for (int i = 0; i < 8; ++i)
      set_bin(t, PHASE_NUMERIC, b++, KIND_SEARCH, G[i], C[i], C[i], INT_MAX);
    // Longer rows: tiles of kTiledPositions positions, runs of tiles per
    // workgroup.
    set_bin(t, PHASE_NUMERIC, b++, KIND_BLOCK, kBlock, kTiledPositions, 4096, INT_MAX);
    set_bin(t, PHASE_NUMERIC, b++, KIND_BLOCK, kBlock, kTiledPositions, 16384, INT_MAX);
    set_bin(t, PHASE_NUMERIC, b++, KIND_BLOCK, kBlock, kTiledPositions, INT_MAX, INT_MAX);
    t.rule.nbins = b;
  }
  return s;
}

std::vector<int> rule_params(int phase) {
  switch (phase) {
    case PHASE_COUNT: return {4, 2, 3, 5, 6};
    case PHASE_FILL: return {4, 2, 3};  // the fill list holds half the slots
    default: return {2048, 1024, 512, 256};
  }
}

void apply_rule(PhaseTable& t, int phase, int param) {
  if (phase == PHASE_NUMERIC) {
    for (int b = 2; b < t.rule.nbins; ++b)
      if (t.spec[b].kind == KIND_SEARCH && t.spec[b].cap > param)
        t.rule.bound[b] = t.rule.bound[b - 1];
    return;
  }
  int last = -1;  // the largest hash bin
  for (int b = 2; b < t.rule.nbins; ++b)
    if (t.spec[b].kind == KIND_HASH) last = b;
  for (int b = 2; b < t.rule.nbins; ++b) {
    if (t.spec[b].kind != KIND_HASH) continue;
    if (phase == PHASE_COUNT && b == last) continue;
    t.rule.bound[b] = std::max(1, (int)((long long)t.spec[b].cap * param / 8));
  }
}

bool apply_choice(PhaseTable& t, int phase, int b, const BinChoice& c,
                  int max_metric) {
  BinSpec cand = t.spec[b];
  if (c.kind) {
    cand.kind = c.kind;
    cand.variant = c.variant;
  }
  if (c.group) cand.group = c.group;
  if (c.cap) cand.cap = c.cap;
  bool ok = true;
  if (b == 1) {
    // The direct bin needs no table; only its lane count may change.
    ok = cand.kind == KIND_DIRECT;
  } else {
    switch (cand.kind) {
      case KIND_HASH:
        ok = phase != PHASE_NUMERIC && max_metric <= cand.cap / 2;
        break;
      case KIND_SEARCH:
      case KIND_MAP:
        ok = phase == PHASE_NUMERIC && max_metric <= cand.cap;
        break;
      case KIND_BLOCK:
        // Tiled: any row length.
        ok = phase == PHASE_NUMERIC;
        if (ok) cand.group = kBlock;
        break;
      case KIND_BITMAP:
        // Bitmap bins are selected by span in the rule, so the kind may not
        // be moved onto a bin whose rows were not span-checked.
        ok = t.spec[b].kind == KIND_BITMAP && cand.cap >= t.rule.span_max[b];
        if (ok) cand.group = kBlock;
        break;
      case KIND_GLOBAL:
        cand.group = kBlock;
        cand.cap = 0;
        break;
      default:
        ok = false;
    }
  }
  if (ok) {
    cand.lds_bytes = lds_bytes_for(phase, cand.kind, cand.group, cand.cap, cand.variant);
    ok = cand.lds_bytes >= 0;
  }
  if (ok) {
    t.spec[b] = cand;
    return true;
  }
  if (b != 1) {
    t.spec[b].kind = KIND_GLOBAL;
    t.spec[b].group = kBlock;
    t.spec[b].cap = 0;
    t.spec[b].lds_bytes = 0;
    t.spec[b].variant = 0;
  }
  return false;
}

}  // namespace rvsp
