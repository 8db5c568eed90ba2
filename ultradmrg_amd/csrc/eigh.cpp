// Batched eigen-decomposition of Hermitian sector matrices for the block SVD.
//
// The block SVD of a two-site state (qlten SVD at
// two_site_update_finite_vmps_impl.h:302-307) is computed here from the
// eigen-decompositions of each sector's Gram matrix, grouped by padded size.
// rocSOLVER's divide-and-conquer solver (syevd / heevd) does the work; its
// Householder tridiagonal reduction issues several launches per column and is
// bound by host-side launch latency, not by the device.  Groups are therefore
// spread over kLanes lanes, each a host thread with its own stream and rocBLAS
// handle, forked from and joined back into the context stream with events.
//
// Initialisation is serial.  rocBLAS loads its Tensile kernel libraries
// lazily, on a handle's first GEMM; four lanes making their first call at the
// same moment from four threads ran that initialisation concurrently, and
// under a profiler (which slows code-object loading) it failed in every lane
// ("Could not initialize Tensile host: unordered_map::at"), after which the
// lanes' dispatches were malformed and the queue aborted.  So the pool calls
// rocblas_initialize() once per process, and each lane's handle makes its
// first GEMM and its first syevd on its own stream, one lane after another,
// before any lane thread exists.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <climits>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>

#include "common.h"
#include "eigh.h"

namespace qlg {

namespace {
constexpr int kLanes = 4;

struct Lane {
  hipStream_t stream = nullptr;
  rocblas_handle handle = nullptr;
  hipEvent_t done = nullptr;
};
}  // namespace

struct EighPool {
  int device = 0;
  Lane lane[kLanes];
  hipEvent_t fork = nullptr;
  DevBuf ws;                 // off-diagonal workspace (doubles), then info (ints)
  int *host_info = nullptr;  // pinned copy of the info words
  size_t host_info_n = 0;
};

void destroy_eigh_pool(EighPool *p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  for (auto &l : p->lane) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    if (l.handle) (void)rocblas_destroy_handle(l.handle);
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.stream) (void)hipStreamDestroy(l.stream);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->host_info) (void)hipHostFree(p->host_info);
  delete p;
}

// First use of every lane's handle, serially: one small GEMM (loads the
// Tensile kernels for the handle) and one small syevd on the lane's stream,
// each lane drained before the next starts.
static void warm_up(EighPool &p) {
  static std::once_flag once;
  std::call_once(once, [] { rocblas_initialize(); });
  constexpr int n = 64;
  DevBuf buf;
  buf.ensure((size_t)(4 * n * n + 2 * n) * sizeof(double) + (size_t)sizeof(int) + 256);
  double *A = buf.as<double>(), *B = A + n * n, *C = B + n * n, *S = C + n * n, *w = S + n * n, *e = w + n;
  int *info = reinterpret_cast<int *>(e + n);
  std::vector<double> h(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) h[i * n + j] = (i == j) ? 2.0 + i : 1.0 / (1.0 + i + j);
  QLG_HIP(hipMemcpy(A, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  QLG_HIP(hipMemcpy(B, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  const double one = 1.0, zero = 0.0;
  for (auto &l : p.lane) {
    if (rocblas_dgemm(l.handle, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, A, n, B, n, &zero, C, n) !=
        rocblas_status_success)
      throw Error(QLG_ERR_HIP, "eigensolver lane warm-up: rocblas_dgemm failed");
    QLG_HIP(hipMemcpyAsync(S, A, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, l.stream));
    if (rocsolver_dsyevd(l.handle, rocblas_evect_original, rocblas_fill_lower, n, S, n, w, e, info) !=
        rocblas_status_success)
      throw Error(QLG_ERR_HIP, "eigensolver lane warm-up: rocsolver_dsyevd failed");
    QLG_HIP(hipStreamSynchronize(l.stream));
  }
}

static EighPool *pool(Ctx &c) {
  if (c.eig) return c.eig;
  auto *p = new EighPool();
  p->device = c.device;
  try {
    for (auto &l : p->lane) {
      QLG_HIP(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
      QLG_HIP(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
      if (rocblas_create_handle(&l.handle) != rocblas_status_success)
        throw Error(QLG_ERR_HIP, "rocblas_create_handle failed");
      if (rocblas_set_stream(l.handle, l.stream) != rocblas_status_success)
        throw Error(QLG_ERR_HIP, "rocblas_set_stream failed");
    }
    QLG_HIP(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
    warm_up(*p);
  } catch (...) {
    destroy_eigh_pool(p);
    throw;
  }
  c.eig = p;
  return p;
}

static void solve_group(rocblas_handle h, int dtype, const qlg_eigh_group &g, double *e, int *info) {
  const rocblas_int n = (rocblas_int)g.n, b = (rocblas_int)g.batch;
  const rocblas_stride sa = (rocblas_stride)g.n * g.n, sv = g.n;
  rocblas_status st;
  // Row-major input read as column-major is the transpose, i.e. conj(A) for a
  // Hermitian A; eigenvector j then lands in row j of the buffer (conjugated
  // for complex data).  The caller undoes both.
  if (dtype == QLG_C128)
    st = rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_lower, n,
                                          static_cast<rocblas_double_complex *>(g.a), n, sa, g.w, sv, e, sv,
                                          info, b);
  else
    st = rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_lower, n,
                                          static_cast<double *>(g.a), n, sa, g.w, sv, e, sv, info, b);
  if (st != rocblas_status_success)
    throw Error(QLG_ERR_HIP, "rocSOLVER syevd/heevd returned status " + std::to_string((int)st));
}

void eigh_batched(Ctx &c, int dtype, int64_t ngroups, const qlg_eigh_group *groups) {
  QLG_CHECK(dtype == QLG_F64 || dtype == QLG_C128, "eigh: dtype must be F64 or C128");
  QLG_CHECK(ngroups >= 0 && (ngroups == 0 || groups), "eigh: bad group list");
  if (ngroups == 0) return;
  std::vector<int64_t> e_off(ngroups), i_off(ngroups);
  int64_t ne = 0, ni = 0;
  for (int64_t k = 0; k < ngroups; ++k) {
    const auto &g = groups[k];
    QLG_CHECK(g.n >= 1 && g.batch >= 1 && g.a && g.w, "eigh: group needs n >= 1, batch >= 1 and buffers");
    QLG_CHECK(g.n <= 46340 && g.batch <= INT_MAX, "eigh: group too large");
    e_off[k] = ne;
    i_off[k] = ni;
    ne += g.n * g.batch;
    ni += g.batch;
  }
  EighPool *p = pool(c);
  const size_t ebytes = (size_t)ne * sizeof(double);
  p->ws.ensure(ebytes + (size_t)ni * sizeof(int) + 256);
  double *E = p->ws.as<double>();
  int *info = reinterpret_cast<int *>(p->ws.as<char>() + ebytes);
  if (p->host_info_n < (size_t)ni) {
    if (p->host_info) (void)hipHostFree(p->host_info);
    p->host_info = nullptr;
    p->host_info_n = 0;
    QLG_HIP(hipHostMalloc((void **)&p->host_info, (size_t)ni * sizeof(int), hipHostMallocDefault));
    p->host_info_n = (size_t)ni;
  }
  // longest-processing-time assignment; a group costs ~ its order (launch
  // latency per column) with a mild batch term
  std::vector<int64_t> order(ngroups);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int64_t k) { return (double)groups[k].n * (1.0 + 0.05 * (double)groups[k].batch); };
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
  std::vector<std::vector<int64_t>> work(kLanes);
  double load[kLanes] = {0, 0, 0, 0};
  for (int64_t k : order) {
    int j = int(std::min_element(load, load + kLanes) - load);
    work[j].push_back(k);
    load[j] += cost(k);
  }
  // QLG_SVD_PROFILE=1: each lane drains its stream and the lanes' finishing
  // times (from the fork) are printed
  const char *prf = std::getenv("QLG_SVD_PROFILE");
  const bool profile = prf && prf[0] == '1';
  double lane_ms[kLanes] = {0, 0, 0, 0};
  const auto t0 = std::chrono::steady_clock::now();
  QLG_HIP(hipEventRecord(p->fork, c.stream));
  std::exception_ptr err[kLanes];
  std::vector<std::thread> th;
  for (int j = 0; j < kLanes; ++j) {
    if (work[j].empty()) continue;
    th.emplace_back([&, j] {
      try {
        QLG_HIP(hipSetDevice(c.device));
        Lane &l = p->lane[j];
        QLG_HIP(hipStreamWaitEvent(l.stream, p->fork, 0));
        for (int64_t k : work[j]) solve_group(l.handle, dtype, groups[k], E + e_off[k], info + i_off[k]);
        QLG_HIP(hipEventRecord(l.done, l.stream));
        if (profile) {
          QLG_HIP(hipStreamSynchronize(l.stream));
          lane_ms[j] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
      } catch (...) {
        err[j] = std::current_exception();
      }
    });
  }
  for (auto &t : th) t.join();
  if (profile) {
    std::string line;
    char buf[96];
    for (int j = 0; j < kLanes; ++j) {
      if (work[j].empty()) continue;
      std::snprintf(buf, sizeof buf, " lane %d [", j);
      line += buf;
      for (int64_t k : work[j]) {
        std::snprintf(buf, sizeof buf, " %lldx%lld", (long long)groups[k].n, (long long)groups[k].batch);
        line += buf;
      }
      std::snprintf(buf, sizeof buf, " ] %.2f ms", lane_ms[j]);
      line += buf;
    }
    std::fprintf(stderr, "[qlg eigh prof]%s\n", line.c_str());
  }
  for (int j = 0; j < kLanes; ++j)
    if (err[j]) {
      // work already queued on the other lanes writes the caller's buffers and
      // the pool's workspace: drain every lane before reporting the failure
      for (int i = 0; i < kLanes; ++i)
        if (!work[i].empty()) (void)hipStreamSynchronize(p->lane[i].stream);
      std::rethrow_exception(err[j]);
    }
  for (int j = 0; j < kLanes; ++j)
    if (!work[j].empty()) QLG_HIP(hipStreamWaitEvent(c.stream, p->lane[j].done, 0));
  QLG_HIP(hipMemcpyAsync(p->host_info, info, (size_t)ni * sizeof(int), hipMemcpyDeviceToHost, c.stream));
  QLG_HIP(hipStreamSynchronize(c.stream));
  for (int64_t k = 0; k < ngroups; ++k)
    for (int64_t i = 0; i < groups[k].batch; ++i)
      if (p->host_info[i_off[k] + i] != 0)
        throw Error(QLG_ERR_NUMERIC, "eigh: divide-and-conquer did not converge (group " + std::to_string(k) +
                                         ", matrix " + std::to_string(i) + ", info " +
                                         std::to_string(p->host_info[i_off[k] + i]) + ")");
}

}  // namespace qlg
