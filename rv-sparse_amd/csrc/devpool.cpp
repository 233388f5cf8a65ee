// devpool.cpp — the library's device memory cache.
//
// Descriptors allocate their device state per estimation (row statistics,
// bins, kept columns, global scratch slots, position maps) and release it
// when they are destroyed or re-estimated. A fresh hipMalloc of a large
// buffer costs the driver time proportional to its size (pages are mapped
// and cleared), which for multi-GB buffers is hundreds of milliseconds, and
// hipFree waits for the whole device. So released buffers are kept here
// and handed to the next request that fits them, the way a caching
// allocator does; memory goes back to the device only past a cap, when an
// allocation runs out of memory, or on rvsp_hip_pool_release.
//
// Safety of reuse: a buffer released without a guarantee that the GPU is
// done with it (a descriptor dropping state its stream may still read) is
// "pending". Pending buffers are handed out again only after one
// hipDeviceSynchronize, which makes every pending buffer idle. Descriptor
// destruction synchronises its stream first and releases its buffers as
// idle directly.
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <map>
#include <mutex>

#include "common.hpp"
#include "descr_impl.hpp"
#include "rv_sparse_hip.h"

namespace rvsp {

namespace {

constexpr size_t kSmall = (size_t)2 << 20;  // below: power-of-two classes

struct Pool {
  std::mutex mu;
  std::multimap<size_t, void*> idle, pending;  // size -> block
  size_t cached = 0;                           // bytes held in both maps
  size_t cap = 0;                              // 0 = not yet known
};

Pool& pool() {
  static Pool* p = new Pool();  // never destroyed: buffers may outlive statics
  return *p;
}

size_t round_size(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes < kSmall) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  return (bytes + kSmall - 1) & ~(kSmall - 1);
}

// Largest block size acceptable for a request of `bytes` (rounded).
size_t fit_limit(size_t bytes) {
  return bytes < kSmall ? bytes : bytes + bytes / 4 + kSmall;
}

size_t cache_cap_locked(Pool& p) {
  if (p.cap == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      total_b = (size_t)64 << 30;
    }
    // Half the device: the working set of a batched product over many
    // groups (R-MAT 2^20 in 10 groups keeps about 100 GB of sorted columns
    // between its count and fill phases) must fit, or every run frees and
    // re-allocates its largest buffers.
    p.cap = total_b / 2;
    if (const char* env = getenv("RVSP_POOL_CAP_MB")) p.cap = (size_t)atoll(env) << 20;
  }
  return p.cap;
}

bool take_locked(std::multimap<size_t, void*>& m, size_t bytes, void** out, size_t* got) {
  auto it = m.lower_bound(bytes);
  if (it == m.end() || it->first > fit_limit(bytes)) return false;
  *out = it->second;
  *got = it->first;
  m.erase(it);
  return true;
}

void free_all_locked(Pool& p, std::multimap<size_t, void*>& m) {
  for (auto& kv : m) {
    (void)hipFree(kv.second);
    p.cached -= kv.first;
  }
  m.clear();
}

// RVSP_POOL_POISON=1 fills every buffer handed out with 0xFF bytes, so a
// kernel that wrongly relies on fresh (driver-cleared) memory fails the same
// way whether or not its buffer came from the cache. Test suites set it.
bool poison_enabled() {
  static const bool on = [] {
    const char* env = getenv("RVSP_POOL_POISON");
    return env && env[0] == '1';
  }();
  return on;
}

hipError_t poisoned(void* ptr, size_t bytes, hipError_t e) {
  if (e != hipSuccess || !poison_enabled()) return e;
  e = hipMemset(ptr, 0xFF, bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return e;
}

}  // namespace

hipError_t pool_alloc(void** out, size_t* bytes_io) {
  const size_t bytes = round_size(*bytes_io ? *bytes_io : 4);
  Pool& p = pool();
  {
    std::lock_guard<std::mutex> lock(p.mu);
    size_t got = 0;
    if (take_locked(p.idle, bytes, out, &got)) {
      p.cached -= got;
      *bytes_io = got;
      return poisoned(*out, got, hipSuccess);
    }
    if (take_locked(p.pending, bytes, out, &got)) {
      // The block may still be read by work queued before its release:
      // wait for the device once, after which every pending block is idle.
      const hipError_t e = hipDeviceSynchronize();
      if (e != hipSuccess) {
        p.pending.emplace(got, *out);
        *out = nullptr;
        return e;
      }
      p.idle.insert(p.pending.begin(), p.pending.end());
      p.pending.clear();
      p.cached -= got;
      *bytes_io = got;
      return poisoned(*out, got, hipSuccess);
    }
  }
  hipError_t e = hipMalloc(out, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    rvsp_hip_pool_release();
    e = hipMalloc(out, bytes);
  }
  if (e != hipSuccess) {
    *out = nullptr;
    return e;
  }
  *bytes_io = bytes;
  return poisoned(*out, bytes, hipSuccess);
}

void pool_free(void* ptr, size_t bytes, bool idle) {
  if (!ptr) return;
  Pool& p = pool();
  std::lock_guard<std::mutex> lock(p.mu);
  if (bytes > cache_cap_locked(p)) {
    (void)hipFree(ptr);
    return;
  }
  (idle ? p.idle : p.pending).emplace(bytes, ptr);
  p.cached += bytes;
  // Over the cap: return the largest idle blocks to the device first.
  while (p.cached > p.cap && !p.idle.empty()) {
    auto it = std::prev(p.idle.end());
    (void)hipFree(it->second);
    p.cached -= it->first;
    p.idle.erase(it);
  }
}

size_t pool_idle_bytes() {
  Pool& p = pool();
  std::lock_guard<std::mutex> lock(p.mu);
  size_t n = 0;
  for (auto& kv : p.idle) n += kv.first;
  for (auto& kv : p.pending) n += kv.first;
  return n;
}

}  // namespace rvsp

extern "C" {

rvsp_status_t rvsp_hip_pool_release(void) {
  rvsp::Pool& p = rvsp::pool();
  std::lock_guard<std::mutex> lock(p.mu);
  if (!p.pending.empty()) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return rvsp::from_hip(e);
  }
  rvsp::free_all_locked(p, p.idle);
  rvsp::free_all_locked(p, p.pending);
  return RVSP_SUCCESS;
}

size_t rvsp_hip_pool_cached_bytes(void) {
  rvsp::Pool& p = rvsp::pool();
  std::lock_guard<std::mutex> lock(p.mu);
  return p.cached;
}

}  // extern "C"
