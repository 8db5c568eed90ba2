// Internal shared definitions of libqlgpu: context, errors, device buffers and
// the host-side view of a block-sparse tensor descriptor.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <map>
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

// Device allocation owned by the library (plan workspace).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    if (n == 0) n = 256;
    if (hipMalloc(&p, n) != hipSuccess) {
      p = nullptr;
      throw Error(QLG_ERR_ALLOC, "hipMalloc of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
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
