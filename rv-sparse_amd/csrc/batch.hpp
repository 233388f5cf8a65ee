// batch.hpp — device-side pieces of the batched SpGEMM path.
#pragma once

#include <hip/hip_runtime.h>

namespace rvsp {

// dst[i] = src[i] + add for i < len; with rebase, dst[i] = src[i] - src[0] + add
// (row pointers moved to a new base). Elements are 32-bit words, so float
// values are moved bit for bit with add = 0.
struct SegCopy {
  const int* src;
  int* dst;
  long long len;
  int add;
  int rebase;
};

constexpr long long kSegChunk = 8192;  // elements per workgroup

struct SegChunk {
  int seg;
  int pad;
  long long begin;
};

// A slice [begin, end) of the packed A's entries, all of one pair.
struct OpsChunk {
  int pair;
  int pad;
  long long begin, end;
};

hipError_t launch_seg_copy(const SegCopy* segs, const SegChunk* chunks, int nchunks,
                           hipStream_t s);
hipError_t launch_pair_ops(const OpsChunk* chunks, int nchunks, const int* a_ci,
                           const int* b_rp, unsigned long long* ops, hipStream_t s);
hipError_t launch_gather_i32(const int* src, const long long* idx, int n, int* out,
                             hipStream_t s);

}  // namespace rvsp
