// batch.hip — kernels of the batched SpGEMM path: segmented copies that pack
// many operands into one block-diagonal operand (and unpack per-pair
// results), per-pair product counts, and a small gather.
//
// Every launch covers the whole batch: the host cuts every segment into
// chunks of kSegChunk elements and one workgroup moves one chunk, so a batch
// of thousands of pairs costs one launch, not thousands of copies.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "common.hpp"

namespace rvsp {

namespace {

__global__ __launch_bounds__(kBlock) void k_seg_copy(const SegCopy* __restrict__ segs,
                                                     const SegChunk* __restrict__ chunks,
                                                     int nchunks) {
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const SegChunk ch = chunks[c];
    const SegCopy sg = segs[ch.seg];
    const long long end = ch.begin + kSegChunk < sg.len ? ch.begin + kSegChunk : sg.len;
    const int base = sg.rebase ? sg.add - sg.src[0] : sg.add;
    for (long long i = ch.begin + threadIdx.x; i < end; i += kBlock)
      sg.dst[i] = sg.src[i] + base;
  }
}

// Sum over a chunk of one pair's A entries of the B-row lengths they reach,
// added into ops[pair].
__global__ __launch_bounds__(kBlock) void k_pair_ops(const OpsChunk* __restrict__ chunks,
                                                     int nchunks, const int* __restrict__ a_ci,
                                                     const int* __restrict__ b_rp,
                                                     unsigned long long* __restrict__ ops) {
  __shared__ unsigned long long part[kBlock / kWave];
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const OpsChunk ch = chunks[c];
    unsigned long long s = 0;
    for (long long j = ch.begin + threadIdx.x; j < ch.end; j += kBlock) {
      const int k = a_ci[j];
      s += (unsigned long long)(b_rp[k + 1] - b_rp[k]);
    }
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_down(s, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int w = 0; w < kBlock / kWave; ++w) t += part[w];
      atomicAdd(&ops[ch.pair], t);
    }
    __syncthreads();
  }
}

__global__ void k_gather_i32(const int* __restrict__ src, const long long* __restrict__ idx,
                             int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

int chunk_grid(int nchunks) { return nchunks < 65536 ? nchunks : 65536; }

}  // namespace

hipError_t launch_seg_copy(const SegCopy* segs, const SegChunk* chunks, int nchunks,
                           hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_copy, dim3(chunk_grid(nchunks)), dim3(kBlock), 0, s, segs, chunks,
                     nchunks);
  return hipGetLastError();
}

hipError_t launch_pair_ops(const OpsChunk* chunks, int nchunks, const int* a_ci,
                           const int* b_rp, unsigned long long* ops, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_ops, dim3(chunk_grid(nchunks)), dim3(kBlock), 0, s, chunks,
                     nchunks, a_ci, b_rp, ops);
  return hipGetLastError();
}

hipError_t launch_gather_i32(const int* src, const long long* idx, int n, int* out,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_i32, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, src,
                     idx, n, out);
  return hipGetLastError();
}

}  // namespace rvsp
