// selector.cpp — per-bin strategy selection beyond the fixed table:
//   * per-bin features from the device (k_bin_stats),
//   * the autotuner, which times every launchable shape of every non-empty
//     bin on the descriptor's own operands and records the fastest in a
//     process-wide tuning table,
//   * RVSP_SPGEMM_ALGO_HIP_TUNED, which looks bins up in that table.
// Both only change which kernel shape serves a bin. Every choice goes
// through apply_choice, where the exact symbolic count vetoes a table that
// cannot hold the bin's rows, so output bits never change.
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "descr_impl.hpp"
#include "selector.hpp"

namespace rvsp {

// ---------------------------------------------------------------------------
// Features.

rvsp_status_t bin_features(rvsp_spgemm_descr* d, const PhasePlan& p, int offsets_slot,
                           const int* a_rp, const int* c_rp, hipStream_t s,
                           std::vector<BinFeatures>* out) {
  RVSP_TRY(ensure(d->bin_stats, sizeof(unsigned long long) * kMaxBins * kBinStats));
  const int nb = p.table.rule.nbins;
  const int total = p.offsets.empty() ? 0 : p.offsets[nb];
  RVSP_HIP_TRY(launch_bin_stats((const int*)p.rows.p,
                                (const int*)d->bin_offsets.p + offsets_slot * (kMaxBins + 1),
                                nb, total, a_rp, row_stats(d), c_rp,
                                (unsigned long long*)d->bin_stats.p, s));
  std::vector<unsigned long long> h((size_t)kMaxBins * kBinStats);
  RVSP_HIP_TRY(hipMemcpyAsync(h.data(), d->bin_stats.p, h.size() * 8,
                              hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  out->assign(nb, BinFeatures());
  for (int b = 0; b < nb; ++b) {
    const unsigned long long* v = &h[(size_t)b * kBinStats];
    BinFeatures& f = (*out)[b];
    f.rows = (double)v[0];
    f.nnza = (double)v[1];
    f.ops = (double)v[2];
    f.metric = (double)v[3];
    f.max_metric = (double)v[4];
    f.span = (double)v[5];
    f.max_maxb = (double)v[6];
    f.max_ops = (double)v[7];
  }
  return RVSP_SUCCESS;
}

int len_class(const BinFeatures& f) {
  if (f.rows <= 0 || f.nnza <= 0) return 0;
  const double mean = f.ops / f.nnza;  // mean B-row length reached
  int c = (int)std::floor(std::log2(std::max(1.0, mean)));
  return c < 0 ? 0 : (c > 10 ? 10 : c);
}

// ---------------------------------------------------------------------------
// Tuning table.

namespace {

struct TuneEntry {
  BinChoice choice;
  float us = 0.f;
};

using TuneKey = std::tuple<int, int, int>;  // phase, bin, len class

std::mutex g_table_mu;
std::map<TuneKey, TuneEntry> g_table;
bool g_table_loaded = false;

std::string default_table_path() {
  if (const char* env = getenv("RVSP_TUNING_TABLE")) return env;
  Dl_info info;
  if (dladdr((void*)&default_table_path, &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) +
           "/tuning_table.txt";
  }
  return "tuning_table.txt";
}

int load_locked(const std::string& path) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return -1;
  char line[256];
  int n = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int ph, b, lc, kind, group, cap, variant = 0;
    float us = 0.f;
    if (sscanf(line, "%d %d %d %d %d %d %f %d", &ph, &b, &lc, &kind, &group, &cap, &us,
               &variant) >= 6) {
      TuneEntry e;
      e.choice.kind = kind;
      e.choice.group = group;
      e.choice.cap = cap;
      e.choice.variant = variant;
      e.us = us;
      g_table[TuneKey(ph, b, lc)] = e;
      ++n;
    }
  }
  fclose(f);
  return n;
}

void ensure_loaded_locked() {
  if (g_table_loaded) return;
  g_table_loaded = true;
  load_locked(default_table_path());
}

}  // namespace

int tuned_rule(int phase, int key) {
  std::lock_guard<std::mutex> lock(g_table_mu);
  ensure_loaded_locked();
  auto it = g_table.find(TuneKey(phase, 0, key));  // bin 0 (empty rows) holds the rule
  return it == g_table.end() ? -1 : it->second.choice.cap;
}

void select_tuned(PhaseTable& t, int phase, const std::vector<BinFeatures>& f) {
  std::lock_guard<std::mutex> lock(g_table_mu);
  ensure_loaded_locked();
  for (int b = 1; b < t.rule.nbins && b < (int)f.size(); ++b) {
    if (f[b].rows <= 0) continue;
    auto it = g_table.find(TuneKey(phase, b, len_class(f[b])));
    if (it == g_table.end()) continue;
    apply_choice(t, phase, b, it->second.choice, (int)f[b].max_metric);
  }
}

// ---------------------------------------------------------------------------
// Autotuner.

namespace {

// Launches one bin's kernel of `phase` with shape `spec`.
hipError_t launch_bin(rvsp_spgemm_descr* d, int phase, const BinSpec& spec, const int* rows,
                      int n, const DevCsr& A, const DevCsr& B, int* c_ci, float* c_val,
                      int max_row, hipStream_t s) {
  const RowStats st = row_stats(d);
  switch (phase) {
    case PHASE_COUNT:
      return launch_count_bin(spec, rows, n, A, B, st, (int*)d->cnt.p, d->gs_bits, s);
    case PHASE_FILL:
      return launch_fill_bin(spec, rows, n, A, B, st, (const int*)d->c_rp.p, c_ci,
                             d->gs_bits, s);
    default:
      if (spec.kind == KIND_MAP) {
        MapView mv = map_view(d);
        mv.row_base = numeric_row_base(d, rows);
        return launch_numeric_map_bin(spec, mv, rows, n, A, B, (const int*)d->c_rp.p, c_val, s);
      }
      if (spec.kind == KIND_BLOCK && ensure_split(d, B, s) != RVSP_SUCCESS)
        return hipErrorOutOfMemory;
      return launch_numeric_bin(spec, rows, n, A, B, st, (const int*)d->c_rp.p, c_ci, c_val,
                                d->gs_dense, max_row, d->split, s);
  }
}

}  // namespace

rvsp_status_t autotune_plan(rvsp_spgemm_descr* d, const DevCsr& A, const DevCsr& B,
                            int* c_ci, float* c_val, int reps, hipStream_t s) {
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  RVSP_HIP_TRY(hipEventCreate(&e0));
  RVSP_HIP_TRY(hipEventCreate(&e1));
  PhasePlan* plans[] = {&d->count, &d->fill, &d->numeric};
  rvsp_status_t status = RVSP_SUCCESS;
  // Bin thresholds first: each rule variant of each phase re-bins the rows
  // and times all of the phase's bins with the fixed table's shapes; the
  // fastest variant stays (recorded under bin 0 of the tuning table) and
  // the per-bin shapes are tuned on its bins below. Count runs first and
  // numeric last, so the row classification left behind is the numeric
  // phase's, which the position map below is planned from.
  {
    const Strategy fixed = fixed_strategy();
    const PhaseTable* ftab[] = {&fixed.count, &fixed.fill, &fixed.numeric};
    const int key = rule_len_class(B.rows, B.nnz);
    for (int ph = 0; ph < 3 && status == RVSP_SUCCESS; ++ph) {
      PhasePlan& p = *plans[ph];
      float best_us = -1.f;
      int best = 0;
      for (int param : rule_params(ph)) {
        PhaseTable t = *ftab[ph];
        apply_rule(t, ph, param);
        status = repartition_phase(d, ph, A, B, t, s);
        if (status != RVSP_SUCCESS) break;
        auto run_all = [&]() -> hipError_t {
          for (int b = ph == PHASE_COUNT ? 2 : 1; b < p.table.rule.nbins; ++b) {
            const int n = p.offsets[b + 1] - p.offsets[b];
            if (n == 0) continue;
            const hipError_t e = launch_bin(d, ph, p.table.spec[b],
                                            (const int*)p.rows.p + p.offsets[b], n, A, B,
                                            c_ci, c_val, p.table.rule.bound[b], s);
            if (e != hipSuccess) return e;
          }
          return hipSuccess;
        };
        hipError_t e = run_all();
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = run_all();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          status = from_hip(e);
          break;
        }
        const float us = ms * 1000.f / reps;
        if (getenv("RVSP_TUNE_VERBOSE"))
          fprintf(stderr, "rvsp tune: phase %d rule %d  %9.2f us\n", ph, param, us);
        if (best_us < 0.f || us < best_us) {
          best_us = us;
          best = param;
        }
      }
      if (status != RVSP_SUCCESS) break;
      PhaseTable t = *ftab[ph];
      apply_rule(t, ph, best);
      status = repartition_phase(d, ph, A, B, t, s);
      TuneEntry te;
      te.choice.cap = best;
      te.us = best_us;
      std::lock_guard<std::mutex> lock(g_table_mu);
      ensure_loaded_locked();
      g_table[TuneKey(ph, 0, key)] = te;
    }
  }
  d->map_promoted = true;  // the measured choice stands for later computes
  // Give every searching numeric bin a position map when the memory budget
  // allows, so searching and the map can be timed against each other.
  {
    PhaseTable& nt = d->numeric.table;
    bool any = false;
    for (int b = 2; b < nt.rule.nbins; ++b) {
      const int n = d->numeric.offsets[b + 1] - d->numeric.offsets[b];
      if (n == 0 || nt.spec[b].kind != KIND_SEARCH) continue;
      BinChoice c;
      c.kind = KIND_MAP;
      PhaseTable trial = nt;  // a refused choice would turn the bin GLOBAL
      if (apply_choice(trial, PHASE_NUMERIC, b, c, nt.rule.bound[b])) {
        nt.spec[b] = trial.spec[b];
        any = true;
      }
    }
    if (any) status = plan_map(d, s);
    if (status == RVSP_SUCCESS && d->map_width) status = ensure_map_built(d, A, B, c_ci, s);
  }
  for (int ph = 0; ph < 3 && status == RVSP_SUCCESS; ++ph) {
    PhasePlan& p = *plans[ph];
    std::vector<BinFeatures> feats;
    status = bin_features(d, p, ph, (const int*)A.rp, ph == PHASE_COUNT ? nullptr
                          : (const int*)d->c_rp.p, s, &feats);
    for (int b = 1; b < p.table.rule.nbins && status == RVSP_SUCCESS; ++b) {
      const int n = p.offsets[b + 1] - p.offsets[b];
      if (n == 0) continue;
      const int* rows = (const int*)p.rows.p + p.offsets[b];
      float best_us = 0.f;
      BinSpec best = p.table.spec[b];
      bool have = false;
      const bool map_ok = ph == PHASE_NUMERIC && ((d->map_bins >> b) & 1u) && d->map_valid;
      const int max_metric = (int)std::min(feats[b].max_metric, 2147483647.0);
      // Every candidate class is timed and recorded as one training sample
      // of the learned selector.
      rvsp_learned_sample_t smp;
      memset(&smp, 0, sizeof smp);
      smp.phase = ph;
      smp.bin = b;
      smp.fixed_class = fixed_class(ph, b);
      for (int c = 0; c < RVSP_MLP_CLASSES; ++c) smp.us[c] = -1.f;
      bin_feature_vector(ph, b, p.table.rule.bound[b], feats[b], smp.features);
      const uint64_t cands = candidate_classes(ph, b, map_ok);
      for (int cls = 0; cls < RVSP_MLP_CLASSES; ++cls) {
        if (!((cands >> cls) & 1ull)) continue;
        PhaseTable trial = p.table;
        if (!apply_choice(trial, ph, b, class_choice(cls, ph, b, max_metric), max_metric))
          continue;
        const BinSpec& spec = trial.spec[b];
        if (spec.lds_bytes > lds_limit()) continue;
        const GlobalSlots& gs = ph == PHASE_NUMERIC ? d->gs_dense : d->gs_bits;
        if (spec.kind == KIND_GLOBAL && gs.nslots == 0) continue;  // no scratch sized
        // One untimed launch, then `reps` timed ones.
        const int max_row = p.table.rule.bound[b];
        hipError_t e = launch_bin(d, ph, spec, rows, n, A, B, c_ci, c_val, max_row, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; ++r)
          e = launch_bin(d, ph, spec, rows, n, A, B, c_ci, c_val, max_row, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          status = from_hip(e);
          break;
        }
        const float us = ms * 1000.f / reps;
        smp.us[cls] = us;
        ++smp.n_candidates;
        if (getenv("RVSP_TUNE_VERBOSE"))
          fprintf(stderr, "rvsp tune: phase %d bin %d rows %d  %-12s v%d G=%-3d cap=%-6d %9.2f us\n",
                  ph, b, n, kind_name(spec.kind), spec.variant, spec.group, spec.cap, us);
        if (!have || us < best_us) {
          have = true;
          best_us = us;
          best = spec;
        }
      }
      if (status != RVSP_SUCCESS || !have) break;
      record_sample(smp);
      p.table.spec[b] = best;
      TuneEntry te;
      te.choice.kind = best.kind;
      te.choice.group = best.group;
      te.choice.cap = best.cap;
      te.choice.variant = best.variant;
      te.us = best_us;
      std::lock_guard<std::mutex> lock(g_table_mu);
      ensure_loaded_locked();
      g_table[TuneKey(ph, b, len_class(feats[b]))] = te;
    }
  }
  // Leave the outputs as compute would: re-run this plan's fill and numeric.
  if (status == RVSP_SUCCESS) status = run_fill(d, A, B, c_ci, s);
  if (status == RVSP_SUCCESS) status = run_numeric(d, A, B, c_ci, c_val, s);
  if (status == RVSP_SUCCESS) status = from_hip(hipStreamSynchronize(s));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return status;
}

int tuning_table_save(const char* path) {
  std::lock_guard<std::mutex> lock(g_table_mu);
  ensure_loaded_locked();
  FILE* f = fopen(path ? path : default_table_path().c_str(), "w");
  if (!f) return -1;
  fprintf(f, "# rv-sparse gfx950 tuning table: phase bin len_class kind group cap us variant\n");
  for (const auto& kv : g_table)
    fprintf(f, "%d %d %d %d %d %d %.3f %d\n", std::get<0>(kv.first), std::get<1>(kv.first),
            std::get<2>(kv.first), kv.second.choice.kind, kv.second.choice.group,
            kv.second.choice.cap, kv.second.us, kv.second.choice.variant);
  fclose(f);
  return (int)g_table.size();
}

int tuning_table_load(const char* path) {
  std::lock_guard<std::mutex> lock(g_table_mu);
  g_table_loaded = true;
  return load_locked(path ? path : default_table_path());
}

void tuning_table_clear() {
  std::lock_guard<std::mutex> lock(g_table_mu);
  g_table.clear();
  g_table_loaded = true;
}

int tuning_table_size() {
  std::lock_guard<std::mutex> lock(g_table_mu);
  ensure_loaded_locked();
  return (int)g_table.size();
}

void destroy_graph(rvsp_spgemm_descr* d) {
  if (d->graph_exec) (void)hipGraphExecDestroy(d->graph_exec);
  if (d->graph) (void)hipGraphDestroy(d->graph);
  d->graph_exec = nullptr;
  d->graph = nullptr;
}

}  // namespace rvsp
