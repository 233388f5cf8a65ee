// legacy.hip — numeric phase of the FP64 and INT8 -> INT32 products.
//
// The reference serves these dtypes with its legacy scalar kernels
// (src/kernels/spgemm/legacy/csr_scalar_f64.c:48-268, csr_scalar_i8.c:23-192):
// per C row, acc[col] += a * b over A's entries in order and each B row's
// entries in order (FP64: product and sum rounded separately; INT8: int32
// products and sums), then the row's columns in ascending order, entries
// whose sum is exactly zero dropped. The structure (sorted columns of every
// row, before dropping) comes from the FP32 path's symbolic phase; these
// kernels add the values in the reference's order and compact the rows.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace rvsp {

namespace {

constexpr int kLegacyChunk = 1024;  // output positions per wave pass
constexpr int kLegacyWaves = kBlock / kWave;

// First index in [b, e) of the ascending array whose value is >= key.
__device__ __forceinline__ int seg_lower_bound(const int* __restrict__ a, int b, int e, int key) {
  while (b < e) {
    const int m = b + ((e - b) >> 1);
    if (a[m] < key)
      b = m + 1;
    else
      e = m;
  }
  return b;
}

// One wave per C row. The row's positions are taken kLegacyChunk at a
// time: the chunk's columns and accumulators sit in this wave's LDS, every
// A entry in order offers the part of its B row inside the chunk's column
// range, and each lane adds its product at the column's position. A
// canonical B row holds a column once, so the lanes of one step touch
// distinct positions; steps follow each other in the wave's program order,
// which is the reference's order of additions for every position.
template <typename TV, typename TA>
__global__ __launch_bounds__(kBlock) void k_numeric_legacy(
    int rows, const int* __restrict__ a_rp, const int* __restrict__ a_ci,
    const TV* __restrict__ a_val, const int* __restrict__ b_rp, const int* __restrict__ b_ci,
    const TV* __restrict__ b_val, const int* __restrict__ c_rp, const int* __restrict__ c_ci,
    TA* __restrict__ c_acc, int* __restrict__ nonzero) {
  __shared__ TA acc_all[kLegacyWaves * kLegacyChunk];
  __shared__ int col_all[kLegacyWaves * kLegacyChunk];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  TA* acc = acc_all + wid * kLegacyChunk;
  int* cols = col_all + wid * kLegacyChunk;
  for (int row = blockIdx.x * kLegacyWaves + wid; row < rows;
       row += gridDim.x * kLegacyWaves) {
    const int cs = c_rp[row];
    const int n = c_rp[row + 1] - cs;
    int nz = 0;
    for (int p0 = 0; p0 < n; p0 += kLegacyChunk) {
      const int tn = min(kLegacyChunk, n - p0);
      for (int j = lane; j < tn; j += kWave) {
        cols[j] = c_ci[cs + p0 + j];
        acc[j] = TA(0);
      }
      const int lo = c_ci[cs + p0], hi = c_ci[cs + p0 + tn - 1];
      const bool whole = tn == n;
      for (int ap = a_rp[row]; ap < a_rp[row + 1]; ++ap) {
        const int k = a_ci[ap];
        const TA a = (TA)a_val[ap];
        int b0 = b_rp[k], b1 = b_rp[k + 1];
        if (!whole) {
          b0 = seg_lower_bound(b_ci, b0, b1, lo);
          b1 = seg_lower_bound(b_ci, b0, b1, hi + 1);
        }
        for (int bp = b0 + lane; bp < b1; bp += kWave) {
          const int pos = seg_lower_bound(cols, 0, tn, b_ci[bp]);
          const TA prod = a * (TA)b_val[bp];
          acc[pos] = acc[pos] + prod;
        }
      }
      for (int j = lane; j < tn; j += kWave) {
        const TA v = acc[j];
        c_acc[cs + p0 + j] = v;
        nz += v != TA(0);
      }
    }
    nz = group_sum<kWave>(nz);
    if (lane == 0) nonzero[row] = nz;
  }
}

// One wave per row: the row's non-zero sums, in column order, to the
// compacted row at out_rp[row].
template <typename TA>
__global__ __launch_bounds__(kBlock) void k_compact_legacy(
    int rows, const int* __restrict__ c_rp, const int* __restrict__ c_ci,
    const TA* __restrict__ c_acc, const int* __restrict__ out_rp, int* __restrict__ out_ci,
    TA* __restrict__ out_val) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int row = blockIdx.x * kLegacyWaves + wid; row < rows;
       row += gridDim.x * kLegacyWaves) {
    const int cs = c_rp[row], ce = c_rp[row + 1];
    int dst = out_rp[row];
    for (int p = cs; p < ce; p += kWave) {
      const int q = p + lane;
      TA v = TA(0);
      int c = 0;
      if (q < ce) {
        v = c_acc[q];
        c = c_ci[q];
      }
      const bool keep = q < ce && v != TA(0);
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int o = dst + __popcll(m & below);
        out_ci[o] = c;
        out_val[o] = v;
      }
      dst += __popcll(m);
    }
  }
}

template <typename TV, typename TA>
hipError_t launch_legacy(int rows, const int* a_rp, const int* a_ci, const void* a_val,
                         const int* b_rp, const int* b_ci, const void* b_val, const int* c_rp,
                         const int* c_ci, void* c_acc, int* nonzero, hipStream_t s) {
  const int grid = std::min((rows + kLegacyWaves - 1) / kLegacyWaves, 65535);
  hipLaunchKernelGGL((k_numeric_legacy<TV, TA>), dim3(grid), dim3(kBlock), 0, s, rows, a_rp,
                     a_ci, (const TV*)a_val, b_rp, b_ci, (const TV*)b_val, c_rp, c_ci,
                     (TA*)c_acc, nonzero);
  return hipGetLastError();
}

template <typename TA>
hipError_t launch_compact(int rows, const int* c_rp, const int* c_ci, const void* c_acc,
                          const int* out_rp, int* out_ci, void* out_val, hipStream_t s) {
  const int grid = std::min((rows + kLegacyWaves - 1) / kLegacyWaves, 65535);
  hipLaunchKernelGGL(k_compact_legacy<TA>, dim3(grid), dim3(kBlock), 0, s, rows, c_rp, c_ci,
                     (const TA*)c_acc, out_rp, out_ci, (TA*)out_val);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_numeric_legacy(bool i8, int rows, const int* a_rp, const int* a_ci,
                                 const void* a_val, const int* b_rp, const int* b_ci,
                                 const void* b_val, const int* c_rp, const int* c_ci,
                                 void* c_acc, int* nonzero, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  return i8 ? launch_legacy<signed char, int>(rows, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp,
                                              c_ci, c_acc, nonzero, s)
            : launch_legacy<double, double>(rows, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp,
                                            c_ci, c_acc, nonzero, s);
}

hipError_t launch_compact_legacy(bool i8, int rows, const int* c_rp, const int* c_ci,
                                 const void* c_acc, const int* out_rp, int* out_ci,
                                 void* out_val, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  return i8 ? launch_compact<int>(rows, c_rp, c_ci, c_acc, out_rp, out_ci, out_val, s)
            : launch_compact<double>(rows, c_rp, c_ci, c_acc, out_rp, out_ci, out_val, s);
}

}  // namespace rvsp
