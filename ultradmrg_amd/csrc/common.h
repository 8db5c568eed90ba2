// Internal shared definitions of libqlgpu: context, errors, device buffers and
// the host-side view of a block-sparse tensor descriptor.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/qlgpu.h"

namespace qlg {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define QLG_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::qlg::Error(QLG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define QLG_CHECK(cond, msg)                                                       \
  do {                                                                             \
    if (!(cond)) throw ::qlg::Error(QLG_ERR_INVALID, msg);                         \
  } while (0)

// Device memory of the library.  Allocations made inside a call on a context
// come from that context's block cache (DevPool): freed blocks are kept and
// handed to later requests of a similar size instead of going back to the
// driver, so the plans a sweep builds and destroys at every site do not call
// hipMalloc / hipFree (a device-wide synchronisation) each time.  Blocks are
// reused in the order of the context's stream: a block returned while work on
// that stream still reads it is only handed to work queued after it on the
// same stream (other streams fork from it with an event).  QLG_POOL=0 turns the
// cache off; QLG_POOL_MB caps the bytes it holds (default: min(16 GiB, 1/8 of
// the device)).
void *dev_malloc(size_t n);     // hipMalloc, throwing QLG_ERR_ALLOC on failure
void dev_free(void *p);

struct DevPool {
  int device = 0;
  size_t cap = 0, cached = 0;
  std::multimap<size_t, void *> blocks;   // cached blocks by size
  int64_t n_malloc = 0, n_reuse = 0, n_free = 0;
  std::mutex mu;
  explicit DevPool(int dev);
  ~DevPool();
  void *get(size_t n, size_t *got);       // a block of at least n bytes
  void put(void *p, size_t n);
  void trim(size_t keep);                 // free cached blocks, largest first, down to keep bytes
};

// The pool that DevBuf::ensure draws from on this thread (set by the
// entry-point guard for the call's context; null: plain hipMalloc).
std::shared_ptr<DevPool> &current_pool();

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  std::shared_ptr<DevPool> pool;   // where p came from (null: hipMalloc)
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      if (pool)
        pool->put(p, bytes);
      else
        dev_free(p);
    }
    p = nullptr;
    bytes = 0;
    pool.reset();
  }
  void ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    if (n == 0) n = 256;
    const std::shared_ptr<DevPool> &cp = current_pool();
    if (cp) {
      p = cp->get(n, &bytes);
      pool = cp;
    } else {
      p = dev_malloc(n);
      bytes = n;
    }
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct EighPool;
void destroy_eigh_pool(EighPool *p);

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  DevBuf partials;   // reduction scratch
  DevBuf scalars;    // device-side scalar results
  double *host_scalars = nullptr;  // pinned, 64 doubles
  EighPool *eig = nullptr;         // lanes of the batched eigensolver, created on first use
  std::shared_ptr<DevPool> pool;   // block cache of the plans' device buffers (null: off)
};

// ---------------------------------------------------------------------------
// Host view of a block-sparse tensor
// ---------------------------------------------------------------------------
using QN = std::array<int32_t, QLG_MAX_QN_LEN>;

inline QN qn_add(const QN &a, const QN &b) { return {a[0] + b[0], a[1] + b[1]}; }
inline QN qn_sub(const QN &a, const QN &b) { return {a[0] - b[0], a[1] - b[1]}; }

using Coords = std::array<int32_t, QLG_MAX_RANK>;

struct CoordsHash {
  size_t operator()(const Coords &c) const {
    uint64_t h = 1469598103934665603ull;
    for (int32_t v : c) { h ^= uint64_t(uint32_t(v)); h *= 1099511628211ull; }
    return size_t(h);
  }
};

struct Axis {
  std::vector<QN> qn;
  std::vector<int64_t> deg;
  int dir = 0;
  bool same_sectors(const Axis &o) const { return qn == o.qn && deg == o.deg; }
  int64_t dim() const { int64_t s = 0; for (auto d : deg) s += d; return s; }
};

struct Ten {
  int rank = 0;
  int qn_len = 1;
  int dtype = QLG_F64;
  std::vector<Axis> ax;
  std::vector<Coords> blk;
  std::vector<int64_t> off;
  std::unordered_map<Coords, int64_t, CoordsHash> find;
  const void *data = nullptr;

  int64_t blk_elems(int64_t b) const {
    int64_t n = 1;
    for (int i = 0; i < rank; ++i) n *= ax[i].deg[blk[b][i]];
    return n;
  }
  int64_t lookup(const Coords &c) const {
    auto it = find.find(c);
    return it == find.end() ? -1 : it->second;
  }
  QN flow(const Coords &c) const {
    QN q{0, 0};
    for (int i = 0; i < rank; ++i) {
      const QN &s = ax[i].qn[c[i]];
      q = ax[i].dir > 0 ? qn_add(q, s) : qn_sub(q, s);
    }
    return q;
  }
};

Ten parse_tensor(const qlg_bsten *d, int expect_rank, const char *name);

// element size in doubles
inline int elem_doubles(int dtype) { return dtype == QLG_C128 ? 2 : 1; }

}  // namespace qlg
