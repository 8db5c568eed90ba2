// Device memory: hipMalloc / hipFree wrappers and the per-context block cache
// (DevPool) that the plans' buffers draw from.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "common.h"

namespace qlg {

namespace {
bool verbose_alloc() {
  static const bool on = [] {
    const char *v = std::getenv("QLG_VERBOSE");
    return v && v[0] != '0';
  }();
  return on;
}

// hipMalloc / hipFree calls slower than this are reported under QLG_VERBOSE
constexpr double kSlowMs = 20.0;

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Cached blocks are handed out for requests down to 2/3 of their size.
constexpr size_t kFitNum = 3, kFitDen = 2;
// Requests are rounded up to 1 MiB above 1 MiB, to 4 KiB below it.
size_t round_request(size_t n) {
  const size_t g = n > (size_t(1) << 20) ? (size_t(1) << 20) : 4096;
  return (n + g - 1) / g * g;
}
}  // namespace

void *dev_malloc(size_t n) {
  void *p = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  const hipError_t e = hipMalloc(&p, n);
  const double ms = ms_since(t0);
  if (verbose_alloc() && ms > kSlowMs) std::fprintf(stderr, "[qlg alloc] hipMalloc %zu bytes: %.1f ms\n", n, ms);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(QLG_ERR_ALLOC, "hipMalloc of " + std::to_string(n) + " bytes failed");
  }
  return p;
}

void dev_free(void *p) {
  if (!p) return;
  const auto t0 = std::chrono::steady_clock::now();
  (void)hipFree(p);
  const double ms = ms_since(t0);
  if (verbose_alloc() && ms > kSlowMs) std::fprintf(stderr, "[qlg alloc] hipFree: %.1f ms\n", ms);
}

DevPool::DevPool(int dev) : device(dev) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) total_b = size_t(64) << 30;
  cap = std::min(size_t(16) << 30, total_b / 8);
  if (const char *v = std::getenv("QLG_POOL_MB")) cap = size_t(std::atoll(v)) << 20;
}

DevPool::~DevPool() {
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (cur != device) (void)hipSetDevice(device);
  trim(0);
  if (verbose_alloc())
    std::fprintf(stderr, "[qlg pool] device %d: %lld hipMalloc, %lld reuses, %lld hipFree\n", device,
                 (long long)n_malloc, (long long)n_reuse, (long long)n_free);
  if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
}

void *DevPool::get(size_t n, size_t *got) {
  n = round_request(n);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = blocks.lower_bound(n);
    if (it != blocks.end() && it->first <= n / kFitDen * kFitNum) {
      void *p = it->second;
      *got = it->first;
      cached -= it->first;
      blocks.erase(it);
      ++n_reuse;
      return p;
    }
  }
  void *p = nullptr;
  try {
    p = dev_malloc(n);
  } catch (const Error &) {
    // the device is full: give the cached blocks back and try once more
    {
      std::lock_guard<std::mutex> lk(mu);
      trim(0);
    }
    p = dev_malloc(n);
  }
  std::lock_guard<std::mutex> lk(mu);
  ++n_malloc;
  *got = n;
  return p;
}

void DevPool::put(void *p, size_t n) {
  std::lock_guard<std::mutex> lk(mu);
  if (n > cap) {   // never cached
    dev_free(p);
    ++n_free;
    return;
  }
  blocks.emplace(n, p);
  cached += n;
  if (cached > cap) trim(cap / 2);
}

// Caller holds mu (or is the destructor).
void DevPool::trim(size_t keep) {
  while (cached > keep && !blocks.empty()) {
    auto it = std::prev(blocks.end());
    dev_free(it->second);
    cached -= it->first;
    blocks.erase(it);
    ++n_free;
  }
}

std::shared_ptr<DevPool> &current_pool() {
  thread_local std::shared_ptr<DevPool> p;
  return p;
}

}  // namespace qlg
