// Host-side parallel loop for the planners' per-channel work.
#pragma once

#include <cstdint>
#include <functional>

namespace qlg {

// Runs f(i) for every i in [0, n) on the calling thread and up to
// host_workers() pool threads, and returns when all are done.  The items
// must be independent.  The first exception thrown by any f(i) is rethrown
// here after the loop has drained.  Calls made while another parallel_for is
// running (from another thread, or nested) run serially on their caller.
void parallel_for(int64_t n, const std::function<void(int64_t)> &f);

// Pool threads besides the caller: QLG_PLAN_THREADS - 1 if set, else
// min(hardware threads, 16) - 1.  0 means serial.
int host_workers();

}  // namespace qlg
