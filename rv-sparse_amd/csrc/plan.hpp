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

// Largest dynamic LDS per workgroup the current device grants (set once at
// device initialisation).
void set_lds_limit(int bytes);
int lds_limit();

}  // namespace rvsp
