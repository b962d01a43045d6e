// oa_attr.hip -- per-device, thread-safe caches of kernel attributes.
//
// Contexts are created from any host thread (the request loop's scheduler
// thread, callers of oa_create, oa_ctx_bytes from oa_server_submit), and a
// process may drive several devices.  The dynamic-LDS limit of a kernel is
// raised once per (kernel, device) under one mutex; the compute-unit count
// is read once per device.
#include <atomic>
#include <map>
#include <mutex>
#include <utility>

#include "oa_kernels.h"

namespace oa {

namespace {

constexpr size_t kDefaultLds = 65536;  // dynamic LDS a launch gets without the attribute
constexpr int kMaxDevices = 64;

std::mutex& grant_mutex() {
  static std::mutex mu;
  return mu;
}

std::map<std::pair<const void*, int>, size_t>& grant_table() {
  static std::map<std::pair<const void*, int>, size_t> t;
  return t;
}

std::atomic<int> g_ncu[kMaxDevices];

}  // namespace

hipError_t grant_lds(const void* fn, size_t bytes) {
  if (bytes <= kDefaultLds) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(grant_mutex());
  size_t& have = grant_table()[{fn, dev}];
  if (have >= bytes) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) have = bytes;
  return e;
}

int cu_count() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
  int v = g_ncu[dev].load(std::memory_order_acquire);
  if (v > 0) return v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      v <= 0)
    v = 256;
  g_ncu[dev].store(v, std::memory_order_release);
  return v;
}

}  // namespace oa
