// The planners' host thread pool (ultradmrg_amd/csrc/host_pool.cpp), built
// on its own with g++: every item runs exactly once, the first exception
// reaches the caller and the pool stays usable, and nested or concurrent
// loops (which run serially on their callers) complete.
#include <atomic>
#include <string>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>
#include "host_pool.h"
int main() {
  using qlg::parallel_for;
  int bad = 0;
  for (int rep = 0; rep < 200; ++rep) {
    const int n = 1 + rep * 7 % 500;
    std::vector<int> hit(n, 0);
    parallel_for(n, [&](int64_t i) { hit[i] += 1; });
    for (int i = 0; i < n; ++i) bad += hit[i] != 1;
  }
  // exceptions propagate and the pool stays usable
  int caught = 0;
  for (int rep = 0; rep < 20; ++rep) {
    try {
      parallel_for(100, [&](int64_t i) { if (i == 37) throw std::runtime_error("item 37"); });
    } catch (const std::runtime_error &e) { caught += std::string(e.what()) == "item 37"; }
  }
  // nested and concurrent calls run (serially where the pool is busy)
  std::atomic<long> sum{0};
  parallel_for(8, [&](int64_t i) { parallel_for(10, [&](int64_t j) { sum += i * 10 + j; }); });
  std::vector<std::thread> ts;
  for (int t = 0; t < 4; ++t) ts.emplace_back([&] { for (int r = 0; r < 50; ++r) parallel_for(64, [&](int64_t i) { sum += 1; }); });
  for (auto &t : ts) t.join();
  std::printf("workers %d bad %d caught %d sum %ld\n", qlg::host_workers(), bad, caught, sum.load());
  return (bad == 0 && caught == 20 && sum.load() == 3160 + 4 * 50 * 64) ? 0 : 1;
}
