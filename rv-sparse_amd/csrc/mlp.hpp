// mlp.hpp — launch interface of the selector network's gfx950 kernels.
//
// The network: RVSP_MLP_FEATURES inputs -> RVSP_MLP_HIDDEN (ReLU) ->
// RVSP_MLP_HIDDEN (ReLU) -> RVSP_MLP_CLASSES logits; a softmax over the
// classes a sample allows, cross-entropy against a target distribution.
// Parameters are one flat float array (layout in rv_sparse_learned.h).
#pragma once

#include <hip/hip_runtime.h>

#include "rv_sparse_learned.h"

namespace rvsp {

constexpr int kMlpF = RVSP_MLP_FEATURES;
constexpr int kMlpH = RVSP_MLP_HIDDEN;
constexpr int kMlpK = RVSP_MLP_CLASSES;
constexpr int kMlpP = RVSP_MLP_PARAMS;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpH * kMlpF;
constexpr int kMlpW2 = kMlpB1 + kMlpH;
constexpr int kMlpB2 = kMlpW2 + kMlpH * kMlpH;
constexpr int kMlpW3 = kMlpB2 + kMlpH;
constexpr int kMlpB3 = kMlpW3 + kMlpK * kMlpH;
static_assert(kMlpB3 + kMlpK == kMlpP, "parameter layout");
static_assert(kMlpK <= 64, "class masks are 64-bit");

constexpr int kMlpSamples = 32;  // samples per workgroup

inline int mlp_parts(int n) { return (n + kMlpSamples - 1) / kMlpSamples; }

// Forward pass: logits (n x K, may be null) and the masked argmax class
// (may be null; -1 when a sample allows no class).
hipError_t launch_mlp_forward(const float* params, const float* x,
                              const unsigned long long* mask, int n, float* logits,
                              int* cls, hipStream_t s);

// Forward + backward over all samples: per-workgroup partial gradients
// (mlp_parts(n) x P) and partial losses, then their ordered sum into grad and
// loss_hist[*step] (step may be null: index 0).
hipError_t launch_mlp_gradient(const float* params, const float* x, const float* q,
                               const unsigned long long* mask, const float* w, int n,
                               float* part_grad, float* part_loss, float* grad,
                               float* loss_hist, const int* step, hipStream_t s);

// One bias-corrected Adam update at step *step + 1, then *step += 1.
hipError_t launch_mlp_adam(float* params, const float* grad, float* m, float* v, float lr,
                           int* step, hipStream_t s);

}  // namespace rvsp
