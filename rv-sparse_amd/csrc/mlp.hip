// mlp.hip — gfx950 kernels of the learned selector's network: forward,
// backward with an ordered (deterministic) gradient reduction, and Adam.
//
// One workgroup of 256 lanes serves 32 samples. Activations of its samples
// stay in LDS from the forward pass through the backward pass; every weight
// gradient is then a 32-term dot product over LDS columns, written as the
// workgroup's partial. A second kernel sums the partials in workgroup order,
// so a gradient does not depend on scheduling. The network is far too small
// for matrix cores to matter (8.6k parameters); the kernels are sized so
// that a whole training step is three short launches, which the trainer
// captures once in a HIP graph and replays per epoch.
//
// Transcendentals use the hardware exp2/log2/rcp/sqrt instructions (about
// 1 ulp): the network's outputs choose kernel shapes, never output values.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "mlp.hpp"

namespace rvsp {

namespace {

constexpr int kThreads = 256;
constexpr int kS = kMlpSamples;

__device__ __forceinline__ float fast_exp(float x) {
  return __builtin_amdgcn_exp2f(x * 1.44269504f);
}
__device__ __forceinline__ float fast_log(float x) {
  return __builtin_amdgcn_logf(x) * 0.693147181f;
}

struct MlpLds {
  float xs[kS][kMlpF];
  float h1[kS][kMlpH];
  float h2[kS][kMlpH];
  float z[kS][kMlpK];   // logits, then dL/dlogits
  float d2[kS][kMlpH];  // dL/dh2 (through the ReLU)
  float d1[kS][kMlpH];  // dL/dh1 (through the ReLU)
  float loss[kS];
};

// Loads the workgroup's samples (zero rows past n) and runs the forward
// pass into lds.z.
__device__ void forward_chunk(const float* __restrict__ prm, const float* __restrict__ x,
                              int base, int ns, MlpLds& L) {
  const int t = threadIdx.x;
  for (int i = t; i < kS * kMlpF; i += kThreads) {
    const int s = i / kMlpF, f = i % kMlpF;
    L.xs[s][f] = s < ns ? x[(size_t)(base + s) * kMlpF + f] : 0.f;
  }
  __syncthreads();
  for (int i = t; i < kS * kMlpH; i += kThreads) {
    const int s = i / kMlpH, j = i % kMlpH;
    const float* wr = prm + kMlpW1 + j * kMlpF;
    float a = prm[kMlpB1 + j];
#pragma unroll
    for (int f = 0; f < kMlpF; ++f) a = a + wr[f] * L.xs[s][f];
    L.h1[s][j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  for (int i = t; i < kS * kMlpH; i += kThreads) {
    const int s = i / kMlpH, j = i % kMlpH;
    const float* wr = prm + kMlpW2 + j * kMlpH;
    float a = prm[kMlpB2 + j];
#pragma unroll 16
    for (int k = 0; k < kMlpH; ++k) a = a + wr[k] * L.h1[s][k];
    L.h2[s][j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  for (int i = t; i < kS * kMlpK; i += kThreads) {
    const int s = i / kMlpK, c = i % kMlpK;
    const float* wr = prm + kMlpW3 + c * kMlpH;
    float a = prm[kMlpB3 + c];
#pragma unroll 16
    for (int k = 0; k < kMlpH; ++k) a = a + wr[k] * L.h2[s][k];
    L.z[s][c] = a;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_mlp_forward(
    const float* __restrict__ prm, const float* __restrict__ x,
    const unsigned long long* __restrict__ mask, int n, float* __restrict__ logits,
    int* __restrict__ cls) {
  __shared__ MlpLds L;
  const int base = blockIdx.x * kS;
  const int ns = min(kS, n - base);
  forward_chunk(prm, x, base, ns, L);
  const int t = threadIdx.x;
  if (logits)
    for (int i = t; i < ns * kMlpK; i += kThreads)
      logits[(size_t)base * kMlpK + i] = L.z[i / kMlpK][i % kMlpK];
  if (cls && t < ns) {
    const unsigned long long m = mask[base + t];
    int best = -1;
    float bz = 0.f;
    for (int c = 0; c < kMlpK; ++c)
      if (((m >> c) & 1ull) && (best < 0 || L.z[t][c] > bz)) {
        best = c;
        bz = L.z[t][c];
      }
    cls[base + t] = best;
  }
}

__global__ __launch_bounds__(kThreads) void k_mlp_backward(
    const float* __restrict__ prm, const float* __restrict__ x, const float* __restrict__ q,
    const unsigned long long* __restrict__ mask, const float* __restrict__ w, int n,
    float* __restrict__ part_grad, float* __restrict__ part_loss) {
  __shared__ MlpLds L;
  const int base = blockIdx.x * kS;
  const int ns = min(kS, n - base);
  forward_chunk(prm, x, base, ns, L);
  const int t = threadIdx.x;

  // Masked softmax cross-entropy, one lane per sample: dL/dz = w (p - q) on
  // allowed classes, 0 elsewhere.
  if (t < kS) {
    float loss = 0.f;
    const unsigned long long m = t < ns ? mask[base + t] : 0ull;
    if (m) {
      float zmax = -3.0e38f;
      for (int c = 0; c < kMlpK; ++c)
        if ((m >> c) & 1ull) zmax = fmaxf(zmax, L.z[t][c]);
      float se = 0.f;
      for (int c = 0; c < kMlpK; ++c)
        if ((m >> c) & 1ull) se = se + fast_exp(L.z[t][c] - zmax);
      const float lse = zmax + fast_log(se);
      const float ws = w[base + t];
      const float* qs = q + (size_t)(base + t) * kMlpK;
      for (int c = 0; c < kMlpK; ++c) {
        if ((m >> c) & 1ull) {
          const float lp = L.z[t][c] - lse;
          const float qc = qs[c];
          loss = loss - qc * lp;
          L.z[t][c] = ws * (fast_exp(lp) - qc);
        } else {
          L.z[t][c] = 0.f;
        }
      }
      loss = ws * loss;
    } else {
      for (int c = 0; c < kMlpK; ++c) L.z[t][c] = 0.f;
    }
    L.loss[t] = loss;
  }
  __syncthreads();
  for (int i = t; i < kS * kMlpH; i += kThreads) {
    const int s = i / kMlpH, j = i % kMlpH;
    float a = 0.f;
    if (L.h2[s][j] > 0.f)
      for (int c = 0; c < kMlpK; ++c) a = a + prm[kMlpW3 + c * kMlpH + j] * L.z[s][c];
    L.d2[s][j] = a;
  }
  __syncthreads();
  for (int i = t; i < kS * kMlpH; i += kThreads) {
    const int s = i / kMlpH, j = i % kMlpH;
    float a = 0.f;
    if (L.h1[s][j] > 0.f)
      for (int k = 0; k < kMlpH; ++k) a = a + prm[kMlpW2 + k * kMlpH + j] * L.d2[s][k];
    L.d1[s][j] = a;
  }
  __syncthreads();

  // Parameter gradients of this workgroup's samples.
  float* out = part_grad + (size_t)blockIdx.x * kMlpP;
  for (int p = t; p < kMlpP; p += kThreads) {
    float g = 0.f;
    if (p < kMlpB1) {
      const int j = p / kMlpF, f = p % kMlpF;
      for (int s = 0; s < kS; ++s) g = g + L.d1[s][j] * L.xs[s][f];
    } else if (p < kMlpW2) {
      const int j = p - kMlpB1;
      for (int s = 0; s < kS; ++s) g = g + L.d1[s][j];
    } else if (p < kMlpB2) {
      const int r = p - kMlpW2, j = r / kMlpH, k = r % kMlpH;
      for (int s = 0; s < kS; ++s) g = g + L.d2[s][j] * L.h1[s][k];
    } else if (p < kMlpW3) {
      const int j = p - kMlpB2;
      for (int s = 0; s < kS; ++s) g = g + L.d2[s][j];
    } else if (p < kMlpB3) {
      const int r = p - kMlpW3, c = r / kMlpH, k = r % kMlpH;
      for (int s = 0; s < kS; ++s) g = g + L.z[s][c] * L.h2[s][k];
    } else {
      const int c = p - kMlpB3;
      for (int s = 0; s < kS; ++s) g = g + L.z[s][c];
    }
    out[p] = g;
  }
  if (t == 0) {
    float l = 0.f;
    for (int s = 0; s < kS; ++s) l = l + L.loss[s];
    part_loss[blockIdx.x] = l;
  }
}

// Ordered sum of the partials: one lane per parameter.
__global__ __launch_bounds__(kThreads) void k_mlp_reduce(
    const float* __restrict__ part_grad, int nparts, float* __restrict__ grad,
    const float* __restrict__ part_loss, float* __restrict__ loss_hist,
    const int* __restrict__ step) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p < kMlpP) {
    float g = 0.f;
    for (int i = 0; i < nparts; ++i) g = g + part_grad[(size_t)i * kMlpP + p];
    grad[p] = g;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && loss_hist) {
    float l = 0.f;
    for (int i = 0; i < nparts; ++i) l = l + part_loss[i];
    loss_hist[step ? *step : 0] = l;
  }
}

__global__ __launch_bounds__(kThreads) void k_mlp_adam(
    float* __restrict__ prm, const float* __restrict__ grad, float* __restrict__ m,
    float* __restrict__ v, float lr, const int* __restrict__ step) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= kMlpP) return;
  const float t = (float)(*step + 1);
  // beta^t = 2^(t log2 beta)
  const float c1 = 1.f - __builtin_amdgcn_exp2f(t * -0.152003093f);   // log2(0.9)
  const float c2 = 1.f - __builtin_amdgcn_exp2f(t * -0.00144341686f); // log2(0.999)
  const float g = grad[p];
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * (g * g);
  m[p] = mm;
  v[p] = vv;
  const float mhat = mm * __builtin_amdgcn_rcpf(c1);
  const float vhat = vv * __builtin_amdgcn_rcpf(c2);
  prm[p] = prm[p] - lr * mhat * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(vhat) + 1e-8f);
}

__global__ void k_mlp_tick(int* step) {
  if (threadIdx.x == 0) *step = *step + 1;
}

}  // namespace

hipError_t launch_mlp_forward(const float* params, const float* x,
                              const unsigned long long* mask, int n, float* logits,
                              int* cls, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_parts(n)), dim3(kThreads), 0, s, params, x,
                     mask, n, logits, cls);
  return hipGetLastError();
}

hipError_t launch_mlp_gradient(const float* params, const float* x, const float* q,
                               const unsigned long long* mask, const float* w, int n,
                               float* part_grad, float* part_loss, float* grad,
                               float* loss_hist, const int* step, hipStream_t s) {
  if (n <= 0) return hipErrorInvalidValue;
  const int parts = mlp_parts(n);
  hipLaunchKernelGGL(k_mlp_backward, dim3(parts), dim3(kThreads), 0, s, params, x, q, mask,
                     w, n, part_grad, part_loss);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpP + kThreads - 1) / kThreads), dim3(kThreads),
                     0, s, part_grad, parts, grad, part_loss, loss_hist, step);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* params, const float* grad, float* m, float* v, float lr,
                           int* step, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpP + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     s, params, grad, m, v, lr, step);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(64), 0, s, step);
  return hipGetLastError();
}

}  // namespace rvsp
