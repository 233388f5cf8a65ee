// common.hpp — shared device helpers and host error plumbing for the gfx950
// SpGEMM library. Everything here is written for 64-lane CDNA wavefronts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "rv_sparse.h"

namespace rvsp {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // threads per workgroup for row kernels
constexpr int kEmptyKey = -1;

// ---------------------------------------------------------------------------
// Host-side status mapping.

inline rvsp_status_t from_hip(hipError_t e) {
  if (e == hipSuccess) return RVSP_SUCCESS;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return RVSP_ERROR_ALLOCATION_FAILED;
  return RVSP_ERROR_DEVICE;
}

#define RVSP_HIP_TRY(expr)                                  \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ::rvsp::from_hip(_e);      \
  } while (0)

#define RVSP_TRY(expr)                                      \
  do {                                                      \
    rvsp_status_t _s = (expr);                              \
    if (_s != RVSP_SUCCESS) return _s;                      \
  } while (0)

// ---------------------------------------------------------------------------
// Device helpers.

// Fibonacci hashing of a column index into a power-of-two table.
__device__ __forceinline__ int hash_slot(int key, int log2_size) {
  return (int)(((uint32_t)key * 2654435761u) >> (32 - log2_size));
}

// Position of `key` in the strictly ascending array s[0..n), n >= 1, where
// key is known to be present. Branch-free halving, ceil(log2 n) steps.
__device__ __forceinline__ int lds_find(const int* s, int n, int key) {
  int lo = 0;
  while (n > 1) {
    int half = n >> 1;
    lo = (s[lo + half] <= key) ? lo + half : lo;
    n -= half;
  }
  return lo;
}

// Inclusive prefix sum over the G consecutive lanes of a sub-wave group
// (G a power of two <= 64), Hillis-Steele over cross-lane shuffles. Every
// lane of the group must be active.
template <int G, typename T>
__device__ __forceinline__ T group_incl_scan(T v) {
  const int l = threadIdx.x & (G - 1);
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    T t = __shfl_up(v, d, G);
    if (l >= d) v += t;
  }
  return v;
}

// Inclusive running maximum over the G consecutive lanes of a group.
template <int G>
__device__ __forceinline__ int group_incl_max(int v) {
  const int l = threadIdx.x & (G - 1);
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const int t = __shfl_up(v, d, G);
    if (l >= d) v = max(v, t);
  }
  return v;
}

// Keeps the compiler from moving LDS accesses across this point (the
// hardware already executes one wave's LDS operations in program order), so
// values written by other lanes of the wave are read from LDS.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int d = G >> 1; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
  return v;
}

template <int G>
__device__ __forceinline__ int group_max(int v) {
#pragma unroll
  for (int d = G >> 1; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, G));
  return v;
}

template <int G>
__device__ __forceinline__ int group_min(int v) {
#pragma unroll
  for (int d = G >> 1; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, G));
  return v;
}

// Block-wide exclusive scan for blockDim.x == NT: wave scans plus an LDS
// carry. `tmp` holds NT / kWave values. Returns the exclusive prefix and
// the block total in *total. Must be reached by every thread of the block.
template <typename T, int NT = kBlock>
__device__ __forceinline__ T block_excl_scan(T v, T* tmp, T* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  T incl = group_incl_scan<kWave>(v);
  if (lane == kWave - 1) tmp[wid] = incl;
  __syncthreads();
  T carry = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) {
    T s = tmp[w];
    carry += (w < wid) ? s : T(0);
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return carry + incl - v;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* tmp) {
  T total;
  (void)block_excl_scan(v, tmp, &total);
  return total;
}

// Exact FP32 accumulate step of the reference scalar kernel
// (accum_scalar_f32.c:67-70 compiled under -std=c11, i.e. without FMA
// contraction): the product is rounded, then the sum is rounded. In this
// toolchain __fmul_rn/__fadd_rn are plain operators that hipcc would fuse
// into v_fma_f32 under its default -ffp-contract=fast, so contraction is
// switched off here explicitly (and for the whole library in the Makefile).
__device__ __forceinline__ float acc_step(float acc, float a, float b) {
#pragma clang fp contract(off)
  const float prod = a * b;
  return acc + prod;
}

}  // namespace rvsp
