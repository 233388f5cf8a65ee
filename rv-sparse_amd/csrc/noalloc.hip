// noalloc.hip — rvsp_csr_check_device_noalloc: the device canonical-CSR
// check for callers that may not allocate device memory (tight serving
// loops, harnesses that forbid hipMalloc). Its only scratch, the 64-bit
// atomic-minimum key of the check's kernels, lives in this file's static
// device array instead of a buffer from the library's cache.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "descr_impl.hpp"
#include "rv_sparse_hip.h"

namespace {

// 256 bytes per device, held until the process ends; the check uses the
// first 8 (one key), whatever the size of the matrix checked.
constexpr int kScratchWords = 32;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned long long* g_ptr[kMaxDevices] = {};  // device address of scratch, per device

// The device address of `scratch` on the current device, looked up once per
// device with that device current (each device holds its own copy). Host
// code passes this address to kernels and copies, never the array's name.
rvsp_status_t scratch_ptr(unsigned long long** out) {
  int dev = 0;
  RVSP_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return RVSP_ERROR_DEVICE;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      return e != hipSuccess ? rvsp::from_hip(e) : RVSP_ERROR_DEVICE;
    }
    g_ptr[dev] = (unsigned long long*)p;
  }
  *out = g_ptr[dev];
  return RVSP_SUCCESS;
}

}  // namespace

extern "C" {

rvsp_status_t rvsp_csr_check_device_noalloc(int32_t rows, int32_t cols, int32_t nnz,
                                            const int32_t* row_ptr, const int32_t* col_idx,
                                            rvsp_csr_status_t* status_out,
                                            int32_t* bad_row_out) {
  if (status_out == NULL) return RVSP_ERROR_NULL_POINTER;
  RVSP_TRY(rvsp::device_ready());
  unsigned long long* key = nullptr;
  RVSP_TRY(scratch_ptr(&key));
  int bad = -1;
  // The key is set before each pass, so the previous call's contents do not
  // matter; calls that overlap on one device share the key and race.
  const rvsp_status_t st = rvsp::csr_check_device_key(rows, cols, nnz, row_ptr, col_idx, key,
                                                      nullptr, status_out, &bad);
  if (bad_row_out) *bad_row_out = bad;
  return st;
}

}  // extern "C"
