// selector.hpp — per-bin features and the tuned / learned selectors.
#pragma once

#include <vector>

#include "descr_impl.hpp"
#include "rv_sparse_learned.h"

namespace rvsp {

// Aggregates of one bin (see launch_bin_stats), as doubles.
struct BinFeatures {
  double rows = 0, nnza = 0, ops = 0, metric = 0, max_metric = 0, span = 0,
         max_maxb = 0, max_ops = 0;
};

rvsp_status_t bin_features(rvsp_spgemm_descr* d, const PhasePlan& p, int offsets_slot,
                           const int* a_rp, const int* c_rp, hipStream_t s,
                           std::vector<BinFeatures>* out);

// floor(log2(mean B-row length reached)), clamped to [0, 10]: the table key
// that separates e.g. stencils (short B rows) from power-law products.
int len_class(const BinFeatures& f);

// HIP_TUNED: applies the tuning table's entry of every non-empty bin.
void select_tuned(PhaseTable& t, int phase, const std::vector<BinFeatures>& f);

// HIP_LEARNED: the MLP selector's choice for every non-empty bin.
// The model is read under the library's lock; the prediction runs on the
// descriptor's scratch (learn_feat) with no lock held.
// reuse: selection for numeric-only computes of an analysed structure; it
// re-shapes only the bins already on the position map, among map classes.
rvsp_status_t select_learned(rvsp_spgemm_descr* d, PhaseTable& t, int phase,
                             const std::vector<BinFeatures>& f, hipStream_t s,
                             bool reuse = false);

// Output classes of the selector network (rv_sparse_learned.h). A class is a
// kernel shape relative to the fixed table's bin: its kind, lanes per row and
// a multiple of the bin's base table size.
// The class the fixed table uses for bin b (-1 for bin 0).
int fixed_class(int phase, int b);
// Classes that may serve bin b (bit c = class c); map_ok admits the MAP
// classes of a numeric bin.
uint64_t candidate_classes(int phase, int b, bool map_ok);
// The choice class `cls` stands for in bin b, given the bin's largest row.
BinChoice class_choice(int cls, int phase, int b, int max_metric);
// The network's raw input for bin b.
void bin_feature_vector(int phase, int b, int bound, const BinFeatures& f, float* out);
// Appends a timed bin to the process-wide sample store.
void record_sample(const rvsp_learned_sample_t& s);

// Times every launchable shape of every non-empty bin of the descriptor's
// plan and keeps the fastest (also recorded in the tuning table). c_ci and
// c_val are C's device arrays; the fill and numeric phases are re-run at the
// end so they hold the product.
rvsp_status_t autotune_plan(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                            int* c_ci, float* c_val, int reps, hipStream_t s);

int tuning_table_save(const char* path);
int tuning_table_load(const char* path);
void tuning_table_clear();
int tuning_table_size();

}  // namespace rvsp
