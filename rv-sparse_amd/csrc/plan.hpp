// plan.hpp — strategy tables: how rows are binned and which accumulator
// kernel (kind, lanes per row, LDS capacity) serves each bin, per phase.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "kernels.hpp"

namespace rvsp {

struct PhaseTable {
  BinRule rule;
  BinSpec spec[kMaxBins];
};

struct Strategy {
  PhaseTable count, fill, numeric;
};

// LDS bytes a workgroup needs for a bin of this kind/group/capacity in the
// given phase; -1 if the combination is not launchable.
int lds_bytes_for(int phase, int kind, int group, int cap, int variant = 0);

// The plain fixed-strategy table (RVSP_SPGEMM_ALGO_DEFAULT / HIP_FIXED).
Strategy fixed_strategy();

// Per-bin overrides of lanes-per-row and kind, as found by the autotuner or
// chosen by the learned selector. kind == 0 keeps the table's kind.
struct BinChoice {
  int kind = 0;
  int group = 0;
  int cap = 0;
  int variant = 0;  // applied together with `kind`
};

// Applies `choice` to bin b of `t` if the result is launchable and holds the
// bin's largest row (`max_metric`); otherwise the bin becomes GLOBAL (the
// exact count vetoes a table that is too small). Returns true when the
// choice was taken as given.
bool apply_choice(PhaseTable& t, int phase, int b, const BinChoice& c,
                  int max_metric);

const char* kind_name(int kind);

// Bin thresholds (the autotuner's second dimension besides kernel shapes).
// A phase's rule variant is one integer: COUNT and FILL, the hash bins'
// admitted load in eighths of their table size (the fixed table's is 4:
// rows up to half the slots; COUNT keeps its largest hash bin's bound, so
// the rows reaching the bitmap bins, whose sorted columns the count phase
// keeps, do not change); NUMERIC, the longest row the searching bins serve
// (the fixed table's 2048; longer rows go to the long-row bins). The first
// candidate is the fixed table's.
std::vector<int> rule_params(int phase);
// Rewrites the bounds of table t (a fixed-table phase) for param; bins that
// the variant leaves empty get their predecessor's bound.
void apply_rule(PhaseTable& t, int phase, int param);

// Largest dynamic LDS per workgroup the current device grants (set once at
// device initialisation).
void set_lds_limit(int bytes);
int lds_limit();

}  // namespace rvsp
