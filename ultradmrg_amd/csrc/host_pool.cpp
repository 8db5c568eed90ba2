// A small persistent pool of host threads for the planners.  The workers are
// started on first use and live until the process ends: the pool object is
// never destroyed, so no destructor joins threads at exit.  One loop runs at
// a time; items are handed out through an atomic counter.
#include "host_pool.h"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace qlg {

namespace {

// set while this thread runs items of a loop: a loop started from inside
// one runs serially, without touching the pool's locks
thread_local bool t_in_loop = false;

struct InLoop {
  bool prev;
  InLoop() : prev(t_in_loop) { t_in_loop = true; }
  ~InLoop() { t_in_loop = prev; }
};

struct Pool {
  std::mutex run_mu;                 // one loop at a time
  std::mutex mu;
  std::condition_variable wake, done;
  const std::function<void(int64_t)> *job = nullptr;
  int64_t n = 0;
  std::atomic<int64_t> next{0};
  int busy = 0;                      // workers inside the current loop
  uint64_t generation = 0;
  std::exception_ptr error;
  int workers = 0;

  explicit Pool(int nw) : workers(nw) {
    for (int i = 0; i < nw; ++i) std::thread([this] { loop(); }).detach();
  }

  // take items until none are left; record the first exception
  void drain(const std::function<void(int64_t)> &f, int64_t count) {
    InLoop mark;
    for (int64_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
      try {
        f(i);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!error) error = std::current_exception();
        next.store(count);   // stop handing out items
      }
    }
  }

  void loop() {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int64_t)> *f;
      int64_t count;
      {
        std::unique_lock<std::mutex> lk(mu);
        wake.wait(lk, [&] { return generation != seen; });
        seen = generation;
        f = job;
        count = n;
        if (!f) continue;   // woke after that loop had finished
        ++busy;
      }
      drain(*f, count);
      {
        std::lock_guard<std::mutex> lk(mu);
        if (--busy == 0) done.notify_all();
      }
    }
  }

  void run(int64_t count, const std::function<void(int64_t)> &f) {
    {
      std::lock_guard<std::mutex> lk(mu);
      job = &f;
      n = count;
      next.store(0);
      error = nullptr;
      ++generation;
    }
    wake.notify_all();
    drain(f, count);
    std::exception_ptr e;
    {
      // every worker that woke for this loop has left it (a worker still
      // asleep will find no items left and leave at once)
      std::unique_lock<std::mutex> lk(mu);
      done.wait(lk, [&] { return busy == 0; });
      e = error;
      job = nullptr;
    }
    if (e) std::rethrow_exception(e);
  }
};

int workers_wanted() {
  if (const char *v = std::getenv("QLG_PLAN_THREADS")) return std::max(0, std::atoi(v) - 1);
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(0, (int)std::min(hw == 0 ? 1u : hw, 16u) - 1);
}

Pool *pool() {
  static Pool *p = new Pool(workers_wanted());   // never deleted (see above)
  return p;
}

}  // namespace

int host_workers() { return pool()->workers; }

void parallel_for(int64_t n, const std::function<void(int64_t)> &f) {
  if (n <= 0) return;
  Pool *p = (n > 1 && !t_in_loop) ? pool() : nullptr;
  if (!p || p->workers == 0) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::unique_lock<std::mutex> lk(p->run_mu, std::try_to_lock);
  if (!lk.owns_lock()) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  p->run(n, f);
}

}  // namespace qlg
