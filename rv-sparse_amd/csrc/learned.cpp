// learned.cpp — the learned per-bin selector (RVSP_SPGEMM_ALGO_HIP_LEARNED):
// the class table the network chooses from, per-bin features, the sample
// store the autotuner fills, training on the device, model persistence and
// the selection step of work_estimation.
//
// The network replaces the fixed table's per-bin choice of accumulator
// kernel; the row binning itself (bin bounds) stays the fixed rule, so the
// exact symbolic count still decides which rows share a bin, and every
// chosen shape passes apply_choice, which turns a bin GLOBAL when the chosen
// table cannot hold the bin's largest row. There is no counterpart in the
// reference; its ADAPTIVE accumulator (src/backend/rvv/
// accum_adaptive_f32.c:213-278) is the structural rule this generalises.
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "descr_impl.hpp"
#include "mlp.hpp"
#include "rv_sparse_learned.h"
#include "selector.hpp"

namespace rvsp {

// ---------------------------------------------------------------------------
// Classes.

namespace {

const int kGroups[7] = {1, 2, 4, 8, 16, 32, 64};

constexpr int kClsHash1 = 0;     // 7 lane counts, the bin's table size
constexpr int kClsHash2 = 7;     // 7 lane counts, twice the table size
constexpr int kClsBitmap = 14;
constexpr int kClsGlobal = 15;
constexpr int kClsSearch = 16;   // 7 lane counts
constexpr int kClsMap = 23;      // 3 variants x 7 lane counts
constexpr int kClsBlock = 44;    // workgroup tiles of kTiledPositions
constexpr int kClsDirect = 45;   // 7 lane counts
constexpr int kClsBlockMore = 52;  // 4 more long-row shapes (kBlockMore)
static_assert(kClsBlockMore + 4 == kMlpK, "class table size");

// Long-row classes beyond kClsBlock: (variant, positions per tile).
// Workgroup tiles of 4096 and 3072 positions fit four CUs' worth of
// workgroups more per CU than the default; wave tiles trade more narrowing
// for no barriers.
const int kBlockMore[4][2] = {{0, 4096}, {0, 3072}, {1, 512}, {1, 1024}};

// Index of the reuse flag among the network's inputs.
constexpr int kReuseFeature = 16;

rvsp_learned_class_t class_entry(int cls) {
  rvsp_learned_class_t c;
  c.cap_mult = 1;
  c.variant = 0;
  c.group = kBlock;
  c.cap = 0;
  if (cls >= kClsBlockMore) {
    c.kind = KIND_BLOCK;
    c.variant = kBlockMore[cls - kClsBlockMore][0];
    c.cap = kBlockMore[cls - kClsBlockMore][1];
  } else if (cls < kClsBitmap) {
    c.kind = KIND_HASH;
    c.group = kGroups[cls % 7];
    c.cap_mult = cls < kClsHash2 ? 1 : 2;
  } else if (cls == kClsBitmap) {
    c.kind = KIND_BITMAP;
  } else if (cls == kClsGlobal) {
    c.kind = KIND_GLOBAL;
  } else if (cls < kClsMap) {
    c.kind = KIND_SEARCH;
    c.group = kGroups[cls - kClsSearch];
  } else if (cls < kClsBlock) {
    c.kind = KIND_MAP;
    c.group = kGroups[(cls - kClsMap) % 7];
    c.variant = (cls - kClsMap) / 7;
  } else if (cls == kClsBlock) {
    c.kind = KIND_BLOCK;
    c.cap = kTiledPositions;
  } else {
    c.kind = KIND_DIRECT;
    c.group = kGroups[cls - kClsDirect];
  }
  return c;
}

int group_index(int g) {
  for (int i = 0; i < 7; ++i)
    if (kGroups[i] == g) return i;
  return -1;
}

const Strategy& fixed() {
  static const Strategy s = fixed_strategy();
  return s;
}

const PhaseTable& fixed_table(int phase) {
  return phase == PHASE_COUNT ? fixed().count
                              : phase == PHASE_FILL ? fixed().fill : fixed().numeric;
}

int next_pow2(long long x) {
  int p = 1;
  while (p < x && p < (1 << 30)) p <<= 1;
  return p;
}

}  // namespace

int fixed_class(int phase, int b) {
  const PhaseTable& t = fixed_table(phase);
  if (b < 1 || b >= t.rule.nbins) return -1;
  const BinSpec& s = t.spec[b];
  const int gi = group_index(s.group);
  switch (s.kind) {
    case KIND_DIRECT: return gi < 0 ? -1 : kClsDirect + gi;
    case KIND_HASH: return gi < 0 ? -1 : kClsHash1 + gi;
    case KIND_SEARCH: return gi < 0 ? -1 : kClsSearch + gi;
    case KIND_MAP: return gi < 0 ? -1 : kClsMap + 7 * s.variant + gi;
    case KIND_BITMAP: return kClsBitmap;
    case KIND_BLOCK:
      for (int i = 0; i < 4; ++i)
        if (s.variant == kBlockMore[i][0] && s.cap == kBlockMore[i][1]) return kClsBlockMore + i;
      return kClsBlock;
    case KIND_GLOBAL: return kClsGlobal;
    default: return -1;
  }
}

uint64_t candidate_classes(int phase, int b, bool map_ok);

uint64_t map_classes() {
  uint64_t m = 0;
  for (int c = kClsMap; c < kClsBlock; ++c) m |= 1ull << c;
  return m;
}

// The classes the network may pick for bin b (a subset of the tuner's).
uint64_t selectable_classes(int phase, int b, bool reuse) {
  // The position map pays off only over repeated numeric-only computes: its
  // classes are selectable only for reuse (the reuse input set), and then
  // they are the only ones, since the descriptor's first numeric-only
  // compute moves exactly the bins that can take a map. The global dense
  // accumulator stays the veto's fallback for the short-row bins (its cost
  // there grows with B's column count, which the features do not carry)
  // and a choice for the long-row ones.
  if (reuse) return phase == PHASE_NUMERIC ? candidate_classes(phase, b, true) & map_classes() : 0;
  uint64_t m = candidate_classes(phase, b, false);
  const int kind = fixed_table(phase).spec[b].kind;
  if (kind == KIND_HASH || kind == KIND_SEARCH || kind == KIND_DIRECT)
    m &= ~(1ull << kClsGlobal);
  return m;
}

uint64_t candidate_classes(int phase, int b, bool map_ok) {
  const PhaseTable& t = fixed_table(phase);
  if (b < 1 || b >= t.rule.nbins) return 0;
  uint64_t m = 0;
  auto add = [&](int first, int n) {
    for (int i = 0; i < n; ++i) m |= 1ull << (first + i);
  };
  if (b == 1) {
    add(kClsDirect, 7);
    return m;
  }
  const int kind = t.spec[b].kind;
  if (phase == PHASE_COUNT || phase == PHASE_FILL) {
    add(kClsHash1, 7);
    add(kClsHash2, 7);
    if (kind == KIND_BITMAP) add(kClsBitmap, 1);
  } else {
    add(kClsSearch, 7);
    if (map_ok) add(kClsMap, 21);
    add(kClsBlock, 1);
    add(kClsBlockMore, 4);
  }
  add(kClsGlobal, 1);
  return m;
}

BinChoice class_choice(int cls, int phase, int b, int max_metric) {
  const PhaseTable& t = fixed_table(phase);
  const BinSpec& fs = t.spec[b];
  const rvsp_learned_class_t c = class_entry(cls);
  BinChoice ch;
  ch.kind = c.kind;
  ch.group = c.group;
  ch.variant = c.variant;
  const long long need = std::max(1, max_metric);
  switch (c.kind) {
    case KIND_HASH:
      ch.cap = (fs.kind == KIND_HASH ? fs.cap : next_pow2(std::max(32LL, 2 * need))) *
               c.cap_mult;
      break;
    case KIND_SEARCH:
    case KIND_MAP:
      ch.cap = fs.kind == KIND_SEARCH ? fs.cap : next_pow2(std::max(16LL, need));
      break;
    case KIND_BLOCK:
      ch.cap = c.cap;  // positions per tile
      break;
    case KIND_BITMAP:
      ch.cap = fs.cap;
      break;
    default:
      ch.cap = 0;
      break;
  }
  return ch;
}

void bin_feature_vector(int phase, int b, int bound, const BinFeatures& f, float* out) {
  const double rows = std::max(1.0, f.rows);
  const double nnza = std::max(1.0, f.nnza);
  auto lg = [](double v) { return (float)log2(1.0 + std::max(0.0, v)); };
  for (int i = 0; i < kMlpF; ++i) out[i] = 0.f;
  out[phase] = 1.f;  // 0..2: phase one-hot
  out[3] = lg(f.rows);
  out[4] = lg(f.nnza / rows);          // mean A-row length
  out[5] = lg(f.ops / rows);           // mean products per row
  out[6] = lg(f.metric / rows);        // mean size metric (nnz bound / exact nnz)
  out[7] = lg(f.max_metric);
  out[8] = lg(f.span / rows);          // mean column span reached
  out[9] = lg(f.max_maxb);             // longest B row reached
  out[10] = lg(f.max_ops);
  out[11] = lg(f.ops / nnza);          // mean B-row length reached
  out[12] = lg((double)bound);
  out[13] = f.ops > 0 ? (float)(f.metric / f.ops) : 0.f;  // compression
  out[14] = b == 1 ? 1.f : 0.f;
  out[15] = lg(f.ops);
}

// ---------------------------------------------------------------------------
// Sample store and model.

namespace {

std::mutex g_mu;
std::vector<rvsp_learned_sample_t> g_samples;

struct Model {
  bool ready = false;
  bool tried_default = false;
  std::vector<float> params;
  float mu[kMlpF];
  float sigma[kMlpF];
  // Bumped whenever params/mu/sigma change, so descriptors know when their
  // device copy of the parameters is stale.
  unsigned long long version = 1;
};
Model g_model;

std::string default_model_path() {
  if (const char* env = getenv("RVSP_LEARNED_MODEL")) return env;
  Dl_info info;
  if (dladdr((void*)&default_model_path, &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) +
           "/learned_model.txt";
  }
  return "learned_model.txt";
}

int load_model_locked(const std::string& path) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return -1;
  char tag[32];
  int F = 0, H = 0, K = 0;
  Model m;
  m.params.resize(kMlpP);
  bool ok = fscanf(f, "%31s %d %d %d", tag, &F, &H, &K) == 4 && !strcmp(tag, "rvsp-mlp") &&
            F == kMlpF && H == kMlpH && K == kMlpK;
  ok = ok && fscanf(f, "%31s", tag) == 1 && !strcmp(tag, "mu");
  for (int i = 0; ok && i < kMlpF; ++i) ok = fscanf(f, "%f", &m.mu[i]) == 1;
  ok = ok && fscanf(f, "%31s", tag) == 1 && !strcmp(tag, "sigma");
  for (int i = 0; ok && i < kMlpF; ++i) ok = fscanf(f, "%f", &m.sigma[i]) == 1;
  ok = ok && fscanf(f, "%31s", tag) == 1 && !strcmp(tag, "params");
  for (int i = 0; ok && i < kMlpP; ++i) ok = fscanf(f, "%f", &m.params[i]) == 1;
  fclose(f);
  if (!ok) return -1;
  ++g_model.version;
  g_model.params = m.params;
  memcpy(g_model.mu, m.mu, sizeof m.mu);
  memcpy(g_model.sigma, m.sigma, sizeof m.sigma);
  g_model.ready = true;
  return 0;
}

bool model_ready_locked() {
  if (!g_model.ready && !g_model.tried_default) {
    g_model.tried_default = true;
    load_model_locked(default_model_path());
  }
  return g_model.ready;
}

void normalise(const float* raw, const Model& m, float* out) {
  for (int i = 0; i < kMlpF; ++i) out[i] = (raw[i] - m.mu[i]) / m.sigma[i];
}

// Device buffers of one call, freed on every exit path.
struct Scratch {
  std::vector<void*> bufs;
  ~Scratch() {
    for (void* p : bufs) (void)hipFree(p);
  }
  hipError_t alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e == hipSuccess) bufs.push_back(*p);
    return e;
  }
  template <typename T>
  hipError_t upload(T** p, const T* h, size_t n, hipStream_t s) {
    hipError_t e = alloc((void**)p, n * sizeof(T));
    if (e == hipSuccess && n) e = hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
    return e;
  }
};

// splitmix64: the seeded initialisation of the trainer.
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Glorot-uniform weights, zero biases.
void init_params(std::vector<float>& p, uint64_t seed) {
  p.assign(kMlpP, 0.f);
  uint64_t s = seed;
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double a = sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < fan_out * fan_in; ++i) {
      const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      p[off + i] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(kMlpW1, kMlpH, kMlpF);
  fill(kMlpW2, kMlpH, kMlpH);
  fill(kMlpW3, kMlpK, kMlpH);
}

// The trainer behind rvsp_mlp_train and rvsp_learned_fit: uploads the
// batch, captures one step (gradient, Adam) in a graph and replays it.
rvsp_status_t train_device(float* params, const float* x, const float* q,
                           const unsigned long long* mask, const float* w, int n, int epochs,
                           float lr, float* loss_out) {
  RVSP_TRY(device_ready());
  hipStream_t s;
  RVSP_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{s};
  Scratch sc;
  float *d_p, *d_x, *d_q, *d_w, *d_pg, *d_pl, *d_g, *d_m, *d_v, *d_loss;
  unsigned long long* d_mask;
  int* d_step;
  const int parts = mlp_parts(n);
  RVSP_HIP_TRY(sc.upload(&d_p, params, kMlpP, s));
  RVSP_HIP_TRY(sc.upload(&d_x, x, (size_t)n * kMlpF, s));
  RVSP_HIP_TRY(sc.upload(&d_q, q, (size_t)n * kMlpK, s));
  RVSP_HIP_TRY(sc.upload(&d_mask, mask, (size_t)n, s));
  RVSP_HIP_TRY(sc.upload(&d_w, w, (size_t)n, s));
  RVSP_HIP_TRY(sc.alloc((void**)&d_pg, (size_t)parts * kMlpP * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_pl, (size_t)parts * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_g, (size_t)kMlpP * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_m, (size_t)kMlpP * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_v, (size_t)kMlpP * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_loss, (size_t)std::max(1, epochs) * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_step, 4));
  // Moments and the step counter start at zero (set outside the graph).
  RVSP_HIP_TRY(hipMemsetAsync(d_m, 0, (size_t)kMlpP * 4, s));
  RVSP_HIP_TRY(hipMemsetAsync(d_v, 0, (size_t)kMlpP * 4, s));
  RVSP_HIP_TRY(hipMemsetAsync(d_step, 0, 4, s));
  if (epochs > 0) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    RVSP_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_mlp_gradient(d_p, d_x, d_q, d_mask, d_w, n, d_pg, d_pl, d_g, d_loss,
                                       d_step, s);
    if (e == hipSuccess) e = launch_mlp_adam(d_p, d_g, d_m, d_v, lr, d_step, s);
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) e = hipGraphLaunch(ge, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    RVSP_HIP_TRY(e);
  }
  RVSP_HIP_TRY(hipMemcpyAsync(params, d_p, (size_t)kMlpP * 4, hipMemcpyDeviceToHost, s));
  if (loss_out && epochs > 0)
    RVSP_HIP_TRY(hipMemcpyAsync(loss_out, d_loss, (size_t)epochs * 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  return RVSP_SUCCESS;
}

// Soft targets from measured times: q_c ~ exp(-ln(t_c / t_best) / tau) over
// the candidates of the sample that selection may pick (for a reuse row, the
// position-map classes; otherwise every other class, as in select_learned).
constexpr double kTau = 0.1;

bool reuse_row(const rvsp_learned_sample_t& smp) {
  return smp.features[kReuseFeature] > 0.5f;
}

void sample_targets(const rvsp_learned_sample_t& smp, float* q, unsigned long long* mask) {
  double best = 0;
  *mask = 0;
  const uint64_t sel = selectable_classes(smp.phase, smp.bin, reuse_row(smp));
  for (int c = 0; c < kMlpK; ++c)
    if (smp.us[c] >= 0.f && ((sel >> c) & 1ull)) {
      *mask |= 1ull << c;
      if (best == 0 || smp.us[c] < best) best = smp.us[c];
    }
  best = std::max(best, 1e-3);
  double z = 0;
  std::vector<double> e(kMlpK, 0.0);
  for (int c = 0; c < kMlpK; ++c)
    if ((*mask >> c) & 1ull) {
      e[c] = exp(-log(std::max((double)smp.us[c], 1e-3) / best) / kTau);
      z += e[c];
    }
  for (int c = 0; c < kMlpK; ++c) q[c] = (float)(z > 0 ? e[c] / z : 0.0);
}

}  // namespace

void record_sample(const rvsp_learned_sample_t& s) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_samples.push_back(s);
}

// ---------------------------------------------------------------------------
// Selection (work_estimation with RVSP_SPGEMM_ALGO_HIP_LEARNED).

rvsp_status_t select_learned(rvsp_spgemm_descr* d, PhaseTable& t, int phase,
                             const std::vector<BinFeatures>& f, hipStream_t s, bool reuse) {
  // The model is copied out under the lock; the device work below runs on
  // the descriptor's own scratch without it, so concurrent estimations (a
  // batch's lanes) never wait for each other's streams.
  float mu[kMlpF], sigma[kMlpF];
  std::vector<float> params;
  unsigned long long version = 0;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!model_ready_locked()) return RVSP_SUCCESS;  // no model: the fixed table
    memcpy(mu, g_model.mu, sizeof mu);
    memcpy(sigma, g_model.sigma, sizeof sigma);
    version = g_model.version;
    if (version != d->learn_version) params = g_model.params;
  }
  Model norm;
  memcpy(norm.mu, mu, sizeof mu);
  memcpy(norm.sigma, sigma, sizeof sigma);
  std::vector<int> bins;
  std::vector<float> x;
  std::vector<unsigned long long> mask;
  for (int b = 1; b < t.rule.nbins && b < (int)f.size(); ++b) {
    if (f[b].rows <= 0) continue;
    // Reuse selection only re-shapes the bins already moved to the map.
    if (reuse && t.spec[b].kind != KIND_MAP) continue;
    uint64_t m = selectable_classes(phase, b, reuse);
    for (int c = 0; c < kMlpK; ++c) {
      if (!((m >> c) & 1ull)) continue;
      PhaseTable trial = t;
      const int mm = (int)std::min(f[b].max_metric, 2147483647.0);
      if (!apply_choice(trial, phase, b, class_choice(c, phase, b, mm), mm) ||
          trial.spec[b].lds_bytes > lds_limit())
        m &= ~(1ull << c);
    }
    if (!m) continue;
    float raw[kMlpF];
    bin_feature_vector(phase, b, t.rule.bound[b], f[b], raw);
    raw[kReuseFeature] = reuse ? 1.f : 0.f;
    float nx[kMlpF];
    normalise(raw, norm, nx);
    bins.push_back(b);
    x.insert(x.end(), nx, nx + kMlpF);
    mask.push_back(m);
  }
  if (bins.empty()) return RVSP_SUCCESS;
  // Scratch layout: params | x (kMaxBins rows) | mask | cls.
  const size_t off_x = (size_t)kMlpP * 4;
  const size_t off_mask = off_x + (size_t)kMaxBins * kMlpF * 4;
  const size_t off_cls = off_mask + (size_t)kMaxBins * 8;
  const size_t bytes = off_cls + (size_t)kMaxBins * 4;
  if (d->learn_feat.bytes < bytes) d->learn_version = 0;  // reallocated: params lost
  RVSP_TRY(ensure(d->learn_feat, bytes));
  char* base = (char*)d->learn_feat.p;
  if (version != d->learn_version) {
    if (params.empty()) {
      std::lock_guard<std::mutex> lock(g_mu);
      if (!g_model.ready) return RVSP_SUCCESS;
      params = g_model.params;
      version = g_model.version;
    }
    RVSP_HIP_TRY(hipMemcpyAsync(base, params.data(), (size_t)kMlpP * 4,
                                hipMemcpyHostToDevice, s));
  }
  const int n = (int)bins.size();  // <= kMaxBins
  RVSP_HIP_TRY(hipMemcpyAsync(base + off_x, x.data(), x.size() * sizeof(float),
                              hipMemcpyHostToDevice, s));
  RVSP_HIP_TRY(hipMemcpyAsync(base + off_mask, mask.data(),
                              mask.size() * sizeof(unsigned long long),
                              hipMemcpyHostToDevice, s));
  RVSP_HIP_TRY(launch_mlp_forward((const float*)base, (const float*)(base + off_x),
                                  (const unsigned long long*)(base + off_mask), n, nullptr,
                                  (int*)(base + off_cls), s));
  std::vector<int> cls(n);
  RVSP_HIP_TRY(hipMemcpyAsync(cls.data(), base + off_cls, (size_t)n * 4, hipMemcpyDeviceToHost,
                              s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  d->learn_version = version;
  for (int i = 0; i < n; ++i) {
    if (cls[i] < 0) continue;
    const int b = bins[i];
    const int mm = (int)std::min(f[b].max_metric, 2147483647.0);
    apply_choice(t, phase, b, class_choice(cls[i], phase, b, mm), mm);
  }
  return RVSP_SUCCESS;
}

}  // namespace rvsp

using namespace rvsp;

extern "C" {

rvsp_status_t rvsp_learned_class_info(int32_t cls, rvsp_learned_class_t* out) {
  if (out == NULL) return RVSP_ERROR_NULL_POINTER;
  if (cls < 0 || cls >= kMlpK) return RVSP_ERROR_INVALID_ARGUMENT;
  *out = class_entry(cls);
  return RVSP_SUCCESS;
}

int32_t rvsp_learned_sample_count(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  return (int32_t)g_samples.size();
}

rvsp_status_t rvsp_learned_sample_get(int32_t i, rvsp_learned_sample_t* out) {
  if (out == NULL) return RVSP_ERROR_NULL_POINTER;
  std::lock_guard<std::mutex> lock(g_mu);
  if (i < 0 || i >= (int32_t)g_samples.size()) return RVSP_ERROR_INVALID_ARGUMENT;
  *out = g_samples[i];
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_learned_sample_add(const rvsp_learned_sample_t* s) {
  if (s == NULL) return RVSP_ERROR_NULL_POINTER;
  if (s->phase < 0 || s->phase > 2) return RVSP_ERROR_INVALID_ARGUMENT;
  record_sample(*s);
  return RVSP_SUCCESS;
}

void rvsp_learned_sample_clear(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_samples.clear();
}

rvsp_status_t rvsp_learned_fit(int32_t first, int32_t count, int32_t epochs, float lr,
                               uint64_t seed, float* final_loss) {
  if (epochs < 1 || !(lr > 0.f)) return RVSP_ERROR_INVALID_ARGUMENT;
  std::vector<rvsp_learned_sample_t> smp;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (first < 0 || count < 1 || (int64_t)first + count > (int64_t)g_samples.size())
      return RVSP_ERROR_INVALID_ARGUMENT;
    smp.assign(g_samples.begin() + first, g_samples.begin() + first + count);
  }
  // A numeric sample with timed position-map classes also teaches the reuse
  // choice: a copy with the reuse input set, whose targets are the map
  // classes' times (the tuner times each map class per numeric-only compute).
  const size_t base_n = smp.size();
  for (size_t i = 0; i < base_n; ++i) {
    if (smp[i].phase != PHASE_NUMERIC || reuse_row(smp[i])) continue;
    const uint64_t maps = map_classes();
    bool timed = false;
    for (int c = 0; c < kMlpK; ++c)
      if (((maps >> c) & 1ull) && smp[i].us[c] >= 0.f) timed = true;
    if (!timed) continue;
    rvsp_learned_sample_t r = smp[i];
    r.features[kReuseFeature] = 1.f;
    smp.push_back(r);
  }
  const int n = (int)smp.size();
  Model m;
  for (int i = 0; i < kMlpF; ++i) {
    double s1 = 0, s2 = 0;
    for (const auto& sm : smp) {
      s1 += sm.features[i];
      s2 += (double)sm.features[i] * sm.features[i];
    }
    const double mean = s1 / n;
    const double var = std::max(0.0, s2 / n - mean * mean);
    m.mu[i] = (float)mean;
    m.sigma[i] = var > 1e-6 ? (float)sqrt(var) : 1.f;
  }
  std::vector<float> x((size_t)n * kMlpF), q((size_t)n * kMlpK), w(n, 1.f / n);
  std::vector<unsigned long long> mask(n);
  for (int i = 0; i < n; ++i) {
    normalise(smp[i].features, m, &x[(size_t)i * kMlpF]);
    sample_targets(smp[i], &q[(size_t)i * kMlpK], &mask[i]);
  }
  init_params(m.params, seed);
  std::vector<float> loss(epochs);
  RVSP_TRY(train_device(m.params.data(), x.data(), q.data(), mask.data(), w.data(), n, epochs,
                        lr, loss.data()));
  if (final_loss) *final_loss = loss.back();
  std::lock_guard<std::mutex> lock(g_mu);
  ++g_model.version;
  g_model.params = m.params;
  memcpy(g_model.mu, m.mu, sizeof m.mu);
  memcpy(g_model.sigma, m.sigma, sizeof m.sigma);
  g_model.ready = true;
  g_model.tried_default = true;
  return RVSP_SUCCESS;
}

int32_t rvsp_learned_model_save(const char* path) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_model.ready) return -1;
  FILE* f = fopen(path ? path : default_model_path().c_str(), "w");
  if (!f) return -1;
  fprintf(f, "rvsp-mlp %d %d %d\nmu", kMlpF, kMlpH, kMlpK);
  for (int i = 0; i < kMlpF; ++i) fprintf(f, " %.9g", g_model.mu[i]);
  fprintf(f, "\nsigma");
  for (int i = 0; i < kMlpF; ++i) fprintf(f, " %.9g", g_model.sigma[i]);
  fprintf(f, "\nparams\n");
  for (int i = 0; i < kMlpP; ++i) fprintf(f, "%.9g%c", g_model.params[i], (i % 8 == 7) ? '\n' : ' ');
  fprintf(f, "\n");
  return fclose(f) == 0 ? 0 : -1;
}

int32_t rvsp_learned_model_load(const char* path) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_model.tried_default = true;
  return load_model_locked(path ? path : default_model_path());
}

void rvsp_learned_model_clear(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  ++g_model.version;
  g_model.ready = false;
  g_model.tried_default = true;
}

int32_t rvsp_learned_model_ready(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  return model_ready_locked() ? 1 : 0;
}

rvsp_status_t rvsp_learned_predict(const float* features, const uint64_t* mask, int32_t n,
                                   int32_t* cls) {
  if (n < 0) return RVSP_ERROR_INVALID_ARGUMENT;
  if (n == 0) return RVSP_SUCCESS;
  if (features == NULL || mask == NULL || cls == NULL) return RVSP_ERROR_NULL_POINTER;
  std::vector<float> x((size_t)n * kMlpF);
  std::vector<float> params;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!model_ready_locked()) return RVSP_ERROR_INVALID_ARGUMENT;
    for (int i = 0; i < n; ++i)
      normalise(features + (size_t)i * kMlpF, g_model, &x[(size_t)i * kMlpF]);
    params = g_model.params;
  }
  return rvsp_mlp_forward(params.data(), x.data(), mask, n, nullptr, cls);
}

rvsp_status_t rvsp_mlp_forward(const float* params, const float* x, const uint64_t* mask,
                               int32_t n, float* logits, int32_t* cls) {
  if (n < 0) return RVSP_ERROR_INVALID_ARGUMENT;
  if (n == 0) return RVSP_SUCCESS;
  if (params == NULL || x == NULL || (cls != NULL && mask == NULL))
    return RVSP_ERROR_NULL_POINTER;
  RVSP_TRY(device_ready());
  hipStream_t s = nullptr;  // legacy default stream: synchronous helper
  Scratch sc;
  float *d_p, *d_x, *d_logits = nullptr;
  unsigned long long* d_mask = nullptr;
  int* d_cls = nullptr;
  RVSP_HIP_TRY(sc.upload(&d_p, params, kMlpP, s));
  RVSP_HIP_TRY(sc.upload(&d_x, x, (size_t)n * kMlpF, s));
  if (cls) RVSP_HIP_TRY(sc.upload(&d_mask, (const unsigned long long*)mask, (size_t)n, s));
  if (cls) RVSP_HIP_TRY(sc.alloc((void**)&d_cls, (size_t)n * 4));
  if (logits) RVSP_HIP_TRY(sc.alloc((void**)&d_logits, (size_t)n * kMlpK * 4));
  RVSP_HIP_TRY(launch_mlp_forward(d_p, d_x, d_mask, n, d_logits, d_cls, s));
  if (logits)
    RVSP_HIP_TRY(hipMemcpyAsync(logits, d_logits, (size_t)n * kMlpK * 4, hipMemcpyDeviceToHost, s));
  if (cls) RVSP_HIP_TRY(hipMemcpyAsync(cls, d_cls, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_mlp_gradient(const float* params, const float* x, const float* q,
                                const uint64_t* mask, const float* w, int32_t n, float* grad,
                                float* loss) {
  if (n < 1) return RVSP_ERROR_INVALID_ARGUMENT;
  if (!params || !x || !q || !mask || !w || !grad) return RVSP_ERROR_NULL_POINTER;
  RVSP_TRY(device_ready());
  hipStream_t s = nullptr;
  Scratch sc;
  float *d_p, *d_x, *d_q, *d_w, *d_pg, *d_pl, *d_g, *d_loss;
  unsigned long long* d_mask;
  const int parts = mlp_parts(n);
  RVSP_HIP_TRY(sc.upload(&d_p, params, kMlpP, s));
  RVSP_HIP_TRY(sc.upload(&d_x, x, (size_t)n * kMlpF, s));
  RVSP_HIP_TRY(sc.upload(&d_q, q, (size_t)n * kMlpK, s));
  RVSP_HIP_TRY(sc.upload(&d_mask, (const unsigned long long*)mask, (size_t)n, s));
  RVSP_HIP_TRY(sc.upload(&d_w, w, (size_t)n, s));
  RVSP_HIP_TRY(sc.alloc((void**)&d_pg, (size_t)parts * kMlpP * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_pl, (size_t)parts * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_g, (size_t)kMlpP * 4));
  RVSP_HIP_TRY(sc.alloc((void**)&d_loss, 4));
  RVSP_HIP_TRY(launch_mlp_gradient(d_p, d_x, d_q, d_mask, d_w, n, d_pg, d_pl, d_g, d_loss,
                                   nullptr, s));
  RVSP_HIP_TRY(hipMemcpyAsync(grad, d_g, (size_t)kMlpP * 4, hipMemcpyDeviceToHost, s));
  float l = 0.f;
  RVSP_HIP_TRY(hipMemcpyAsync(&l, d_loss, 4, hipMemcpyDeviceToHost, s));
  RVSP_HIP_TRY(hipStreamSynchronize(s));
  if (loss) *loss = l;
  return RVSP_SUCCESS;
}

rvsp_status_t rvsp_mlp_train(float* params, const float* x, const float* q,
                             const uint64_t* mask, const float* w, int32_t n, int32_t epochs,
                             float lr, float* loss_out) {
  if (n < 1 || epochs < 0 || !(lr > 0.f)) return RVSP_ERROR_INVALID_ARGUMENT;
  if (!params || !x || !q || !mask || !w) return RVSP_ERROR_NULL_POINTER;
  return train_device(params, x, q, (const unsigned long long*)mask, w, n, epochs, lr,
                      loss_out);
}

}  // extern "C"
