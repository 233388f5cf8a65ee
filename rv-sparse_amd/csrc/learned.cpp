// learned.cpp — the learned per-bin selector (RVSP_SPGEMM_ALGO_HIP_LEARNED).
#include "selector.hpp"

namespace rvsp {

rvsp_status_t select_learned(PhaseTable& t, int phase, const std::vector<BinFeatures>& f,
                             hipStream_t s) {
  (void)t;
  (void)phase;
  (void)f;
  (void)s;
  return RVSP_SUCCESS;
}

}  // namespace rvsp
