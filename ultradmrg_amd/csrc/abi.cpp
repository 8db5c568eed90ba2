// extern "C" entry points of libqlgpu (declared in include/qlgpu.h).  Every
// call converts library exceptions into status codes and records the message
// on the context.
#include <cstdlib>
#include <cstring>
#include <memory>

#include "eigh.h"
#include "krylov.h"
#include "noalloc.h"
#include "plan.h"
#include "svd.h"
#include "tridiag.h"
#include "vecops.h"

using namespace qlg;

struct qlg_ctx_s {
  Ctx c;
};
struct qlg_plan_s {
  std::unique_ptr<Plan> p;
};
struct qlg_svd_s {
  std::unique_ptr<SvdResult> r;
};

namespace {
thread_local std::string g_noctx_error;

// Device buffers allocated during a call come from the call context's pool.
struct PoolScope {
  std::shared_ptr<DevPool> saved;
  explicit PoolScope(Ctx *c) : saved(current_pool()) { current_pool() = c ? c->pool : nullptr; }
  ~PoolScope() { current_pool() = saved; }
};

template <class F>
int guarded(Ctx *c, F &&f) {
  try {
    PoolScope scope(c);
    if (c) QLG_HIP(hipSetDevice(c->device));
    f();
    return QLG_OK;
  } catch (const Error &e) {
    (c ? c->last_error : g_noctx_error) = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    (c ? c->last_error : g_noctx_error) = "host allocation failed";
    return QLG_ERR_ALLOC;
  } catch (const std::exception &e) {
    (c ? c->last_error : g_noctx_error) = e.what();
    return QLG_ERR_INVALID;
  }
}
}  // namespace

extern "C" {

int qlg_ctx_create(int device, void *stream, qlg_ctx *out) {
  if (!out) return QLG_ERR_INVALID;
  *out = nullptr;
  auto h = new (std::nothrow) qlg_ctx_s();
  if (!h) return QLG_ERR_ALLOC;
  int rc = guarded(nullptr, [&] {
    int ndev = 0;
    QLG_HIP(hipGetDeviceCount(&ndev));
    QLG_CHECK(device >= 0 && device < ndev, "device index out of range");
    h->c.device = device;
    QLG_HIP(hipSetDevice(device));
    if (stream) {
      h->c.stream = static_cast<hipStream_t>(stream);
    } else {
      QLG_HIP(hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking));
      h->c.own_stream = true;
    }
    h->c.partials.ensure(2 * kMaxPartials * sizeof(double));
    h->c.scalars.ensure(64 * sizeof(double));
    noalloc_init(device);   // address of the fixed staging array on this device
    QLG_HIP(hipHostMalloc((void **)&h->c.host_scalars, 64 * sizeof(double), hipHostMallocDefault));
    const char *pool = std::getenv("QLG_POOL");
    if (!(pool && pool[0] == '0')) h->c.pool = std::make_shared<DevPool>(device);
  });
  if (rc != QLG_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return QLG_OK;
}

int qlg_ctx_destroy(qlg_ctx ctx) {
  if (!ctx) return QLG_OK;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  destroy_eigh_pool(ctx->c.eig);
  ctx->c.eig = nullptr;
  if (ctx->c.host_scalars) (void)hipHostFree(ctx->c.host_scalars);
  if (ctx->c.own_stream && ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return QLG_OK;
}

const char *qlg_last_error(qlg_ctx ctx) { return ctx ? ctx->c.last_error.c_str() : g_noctx_error.c_str(); }

int qlg_ctx_synchronize(qlg_ctx ctx) {
  if (!ctx) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] { QLG_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

int qlg_ctx_release_cached(qlg_ctx ctx) {
  if (!ctx) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    if (!ctx->c.pool) return;
    QLG_HIP(hipStreamSynchronize(ctx->c.stream));
    std::lock_guard<std::mutex> lk(ctx->c.pool->mu);
    ctx->c.pool->trim(0);
  });
}

int qlg_ctx_memory_stats(qlg_ctx ctx, int64_t stats[4]) {
  if (!ctx || !stats) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    for (int i = 0; i < 4; ++i) stats[i] = 0;
    if (!ctx->c.pool) return;
    DevPool &pl = *ctx->c.pool;
    std::lock_guard<std::mutex> lk(pl.mu);
    stats[0] = (int64_t)pl.cached;
    stats[1] = pl.n_malloc;
    stats[2] = pl.n_reuse;
    stats[3] = pl.n_free;
  });
}

int qlg_device_alloc(qlg_ctx ctx, int64_t bytes, void **ptr) {
  if (!ctx || !ptr || bytes < 0) return QLG_ERR_INVALID;
  *ptr = nullptr;
  return guarded(&ctx->c, [&] {
    if (hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1)) != hipSuccess) {
      *ptr = nullptr;
      throw Error(QLG_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
  });
}

int qlg_device_free(qlg_ctx ctx, void *ptr) {
  if (!ctx) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    if (ptr) QLG_HIP(hipFree(ptr));
  });
}

static int copy_sync(qlg_ctx ctx, void *dst, const void *src, int64_t bytes, hipMemcpyKind kind) {
  if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    if (bytes > 0) QLG_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, kind, ctx->c.stream));
    QLG_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int qlg_memcpy_h2d(qlg_ctx ctx, void *dst, const void *src, int64_t bytes) {
  return copy_sync(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}
int qlg_memcpy_d2h(qlg_ctx ctx, void *dst, const void *src, int64_t bytes) {
  return copy_sync(ctx, dst, src, bytes, hipMemcpyDeviceToHost);
}
int qlg_memcpy_d2d(qlg_ctx ctx, void *dst, const void *src, int64_t bytes) {
  return copy_sync(ctx, dst, src, bytes, hipMemcpyDeviceToDevice);
}

// QLG_PREC_* values; the reduced modes exist for real data only
static void check_precision(int precision, int dtype) {
  QLG_CHECK(precision == QLG_PREC_F64 || precision == QLG_PREC_F32 || precision == QLG_PREC_BF16X3 ||
                precision == QLG_PREC_BF16X6,
            "unknown precision");
  if (precision == QLG_PREC_F32 && dtype != QLG_F64)
    throw Error(QLG_ERR_UNSUPPORTED, "complex data runs FP64 or split-BF16 sector GEMMs (no FP32 form)");
}

static int make_plan(qlg_ctx ctx, int nphys, const qlg_bsten *L, const qlg_bsten *const *W, const qlg_bsten *R,
                     const qlg_bsten *psi, int precision, qlg_plan *out) {
  if (!ctx || !out) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    check_precision(precision, psi ? psi->dtype : QLG_F64);
    auto h = std::make_unique<qlg_plan_s>();
    h->p.reset(build_plan(&ctx->c, nphys, L, W, R, psi, precision));
    *out = h.release();
  });
}

int qlg_plan_two_site(qlg_ctx ctx, const qlg_bsten *L, const qlg_bsten *Wl, const qlg_bsten *Wr, const qlg_bsten *R,
                      const qlg_bsten *psi, int precision, qlg_plan *out) {
  const qlg_bsten *W[2] = {Wl, Wr};
  return make_plan(ctx, 2, L, W, R, psi, precision, out);
}

int qlg_plan_one_site(qlg_ctx ctx, const qlg_bsten *L, const qlg_bsten *W, const qlg_bsten *R, const qlg_bsten *psi,
                      int precision, qlg_plan *out) {
  const qlg_bsten *Ws[2] = {W, nullptr};
  return make_plan(ctx, 1, L, Ws, R, psi, precision, out);
}

int qlg_plan_super_block(qlg_ctx ctx, int64_t ngroups, const int64_t *nleft, const qlg_term *left,
                         const int64_t *nright, const qlg_term *right, const qlg_bsten *psi, int precision,
                         qlg_plan *out) {
  if (!ctx || !out) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    check_precision(precision, psi ? psi->dtype : QLG_F64);
    auto h = std::make_unique<qlg_plan_s>();
    h->p.reset(build_super_block_plan(&ctx->c, ngroups, nleft, left, nright, right, psi, precision));
    *out = h.release();
  });
}

int qlg_plan_batch(qlg_ctx ctx, int64_t nplans, const qlg_plan *plans, qlg_plan *out) {
  if (!ctx || !out || !plans || nplans < 1) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    std::vector<Plan *> members(nplans);
    for (int64_t i = 0; i < nplans; ++i) {
      QLG_CHECK(plans[i] != nullptr, "qlg_plan_batch: null member plan");
      members[i] = plans[i]->p.get();
    }
    auto h = std::make_unique<qlg_plan_s>();
    h->p.reset(build_batch_plan(&ctx->c, nplans, members.data()));
    *out = h.release();
  });
}

int qlg_plan_batch_offsets(qlg_plan plan, int64_t *nplans, int64_t *offsets) {
  if (!plan || !nplans) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.batch_off.size() >= 2, "not a batch plan");
    *nplans = (int64_t)p.batch_off.size() - 1;
    if (offsets)
      for (size_t i = 0; i < p.batch_off.size(); ++i) offsets[i] = p.batch_off[i];
  });
}

int qlg_plan_use_graphs(qlg_plan plan, int on) {
  if (!plan) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.kind == PLAN_HEFF, "graph capture applies to effective-Hamiltonian plans");
    p.use_graphs = on != 0;
    if (!p.use_graphs) p.drop_graphs();
  });
}

int qlg_plan_set_reorth(qlg_plan plan, int passes) {
  if (!plan) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.kind == PLAN_HEFF, "reorthogonalisation applies to effective-Hamiltonian plans");
    QLG_CHECK(passes >= 0 && passes <= 2, "reorthogonalisation passes must be 0, 1 or 2");
    if (p.reorth != passes) p.drop_graphs();
    p.reorth = passes;
  });
}

int qlg_lanczos_gs_batch(qlg_plan plan, void *psi_inout, const qlg_lanczos_params *params,
                         qlg_lanczos_result *res) {
  if (!plan || !psi_inout || !params || !res) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] { lanczos_gs_batch(p, static_cast<double *>(psi_inout), *params, res); });
}

int qlg_plan_destroy(qlg_plan plan) {
  if (!plan) return QLG_OK;
  (void)hipSetDevice(plan->p->ctx->device);
  (void)hipStreamSynchronize(plan->p->ctx->stream);
  delete plan;
  return QLG_OK;
}

int qlg_plan_set_precision(qlg_plan plan, int precision) {
  if (!plan) return QLG_ERR_INVALID;
  return guarded(plan->p->ctx, [&] {
    QLG_CHECK(plan->p->kind == PLAN_HEFF, "precision applies to effective-Hamiltonian plans");
    check_precision(precision, plan->p->dtype);
    if (plan->p->precision != precision) plan->p->drop_graphs();   // captured launches name the old variant
    plan->p->precision = precision;
  });
}

int qlg_plan_flops(qlg_plan plan, double *algo_flops, double *exec_flops) {
  if (!plan) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  if (algo_flops) *algo_flops = p.algo_flops;
  if (exec_flops) *exec_flops = p.st[0].flops + p.st[1].flops + p.st[2].flops;
  return QLG_OK;
}

int qlg_plan_vector_layout(qlg_plan plan, int64_t *nblocks, int64_t *size, int32_t *coords, int64_t *offsets) {
  if (!plan) return QLG_ERR_INVALID;
  const Layout &v = plan->p->vec;
  if (nblocks) *nblocks = (int64_t)v.blk.size();
  if (size) *size = v.size;
  for (size_t b = 0; b < v.blk.size(); ++b) {
    if (coords)
      for (int i = 0; i < v.rank; ++i) coords[b * v.rank + i] = v.blk[b][i];
    if (offsets) offsets[b] = v.off[b];
  }
  return QLG_OK;
}

int qlg_plan_workspace(qlg_plan plan, int64_t *bytes) {
  if (!plan || !bytes) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  int64_t b = (int64_t)(p.X.bytes + p.Y.bytes + p.C.bytes + p.basis.bytes);
  for (auto &s : p.st) b += (int64_t)(s.d_tiles.bytes + s.d_tasks.bytes + s.d_segs.bytes);
  *bytes = b;
  return QLG_OK;
}

int qlg_plan_enable_timing(qlg_plan plan, int on) {
  if (!plan) return QLG_ERR_INVALID;
  plan->p->timing = on != 0;
  return QLG_OK;
}

int qlg_plan_stage_ms(qlg_plan plan, double *ms3) {
  if (!plan || !ms3) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.timing && p.ev[3], "timing is not enabled or no matvec has run");
    QLG_HIP(hipEventSynchronize(p.ev[3]));
    for (int i = 0; i < 3; ++i) {
      float ms = 0;
      QLG_HIP(hipEventElapsedTime(&ms, p.ev[i], p.ev[i + 1]));
      ms3[i] = ms;
    }
  });
}

int qlg_plan_stage_flops(qlg_plan plan, double *flops3) {
  if (!plan || !flops3) return QLG_ERR_INVALID;
  for (int i = 0; i < 3; ++i) flops3[i] = plan->p->st[i].flops;
  return QLG_OK;
}

int qlg_plan_stage_work(qlg_plan plan, double *algo_flops3, double *algo_bytes3) {
  if (!plan) return QLG_ERR_INVALID;
  for (int i = 0; i < 3; ++i) {
    if (algo_flops3) algo_flops3[i] = plan->p->st[i].algo_flops;
    if (algo_bytes3) algo_bytes3[i] = plan->p->st[i].algo_bytes;
  }
  return QLG_OK;
}

int qlg_matvec(qlg_plan plan, const void *psi, void *out) {
  if (!plan || !psi || !out) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.kind == PLAN_HEFF, "qlg_matvec needs an effective-Hamiltonian plan");
    p.matvec(static_cast<const double *>(psi), static_cast<double *>(out));
  });
}

int qlg_plan_contract(qlg_ctx ctx, const qlg_bsten *A, const qlg_bsten *B, int k, int conj_a, int conj_b,
                      qlg_plan *out) {
  if (!ctx || !out) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    auto h = std::make_unique<qlg_plan_s>();
    h->p.reset(build_tensordot_plan(&ctx->c, A, B, k, conj_a, conj_b));
    *out = h.release();
  });
}

int qlg_plan_env(qlg_ctx ctx, const qlg_bsten *env, const qlg_bsten *mps, const qlg_bsten *mpo, char side,
                 qlg_plan *out) {
  if (!ctx || !out) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    auto h = std::make_unique<qlg_plan_s>();
    h->p.reset(build_env_plan(&ctx->c, env, mps, mpo, side));
    *out = h.release();
  });
}

int qlg_plan_transpose(qlg_ctx ctx, const qlg_bsten *t, const int32_t *perm, qlg_plan *out) {
  if (!ctx || !out) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    auto h = std::make_unique<qlg_plan_s>();
    h->p.reset(build_transpose_plan(&ctx->c, t, perm));
    *out = h.release();
  });
}

int qlg_execute(qlg_plan plan, const void *in0, const void *in1, void *out) {
  if (!plan || !out) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.kind != PLAN_HEFF, "use qlg_matvec for effective-Hamiltonian plans");
    p.execute(in0, in1, out);
  });
}

int qlg_dot(qlg_ctx ctx, int dtype, int64_t n, const void *x, const void *y, double *out) {
  if (!ctx || !out || n < 0) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    Ctx &c = ctx->c;
    vec_dot(c, dtype, n, static_cast<const double *>(x), static_cast<const double *>(y), c.scalars.as<double>());
    const int nv = dtype == QLG_C128 ? 2 : 1;
    QLG_HIP(hipMemcpyAsync(c.host_scalars, c.scalars.p, nv * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QLG_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < nv; ++i) out[i] = c.host_scalars[i];
  });
}

int qlg_nrm2(qlg_ctx ctx, int dtype, int64_t n, const void *x, double *out) {
  if (!ctx || !out || n < 0) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    Ctx &c = ctx->c;
    vec_nrm2(c, dtype, n, static_cast<const double *>(x), c.scalars.as<double>());
    QLG_HIP(hipMemcpyAsync(c.host_scalars, c.scalars.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QLG_HIP(hipStreamSynchronize(c.stream));
    out[0] = c.host_scalars[0];
  });
}

int qlg_scal(qlg_ctx ctx, int dtype, int64_t n, const double *alpha, void *x) {
  if (!ctx || !alpha || n < 0) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    vec_scal(ctx->c, dtype, n, alpha[0], dtype == QLG_C128 ? alpha[1] : 0.0, static_cast<double *>(x));
  });
}

int qlg_axpy(qlg_ctx ctx, int dtype, int64_t n, const double *alpha, const void *x, void *y) {
  if (!ctx || !alpha || n < 0) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    vec_axpy(ctx->c, dtype, n, alpha[0], dtype == QLG_C128 ? alpha[1] : 0.0, static_cast<const double *>(x),
             static_cast<double *>(y));
  });
}

int qlg_linear_combine(qlg_ctx ctx, int dtype, int64_t n, int64_t k, const double *coefs, const void *const *bases,
                       const double *beta, void *y) {
  if (!ctx || !coefs || !bases || !beta || n < 0 || k < 0) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    Ctx &c = ctx->c;
    const bool cplx = dtype == QLG_C128;
    DevBuf cb, pb;
    cb.ensure(std::max<int64_t>(k, 1) * (cplx ? 2 : 1) * sizeof(double));
    pb.ensure(std::max<int64_t>(k, 1) * sizeof(double *));
    if (k > 0) {
      QLG_HIP(hipMemcpyAsync(cb.p, coefs, k * (cplx ? 2 : 1) * sizeof(double), hipMemcpyHostToDevice, c.stream));
      QLG_HIP(hipMemcpyAsync(pb.p, bases, k * sizeof(void *), hipMemcpyHostToDevice, c.stream));
    }
    const double br = beta[0], bi = cplx ? beta[1] : 0.0;
    vec_lincomb(c, dtype, n, (int)k, cb.as<double>(), pb.as<const double *>(), cplx, br, bi,
                br != 0.0 || bi != 0.0, static_cast<double *>(y));
    QLG_HIP(hipStreamSynchronize(c.stream));
  });
}

int qlg_linear_combine_noalloc(qlg_ctx ctx, int dtype, int64_t n, int64_t k, const double *coefs,
                               const void *const *bases, const double *beta, void *y) {
  if (!ctx || !coefs || !bases || !beta || n < 0 || k < 0) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] {
    Ctx &c = ctx->c;
    QLG_CHECK(dtype == QLG_F64 || dtype == QLG_C128, "dtype must be QLG_F64 or QLG_C128");
    QLG_CHECK(k <= kNoAllocMaxBases, "qlg_linear_combine_noalloc: k = " + std::to_string(k) +
                                          " exceeds the staging limit of " + std::to_string(kNoAllocMaxBases) +
                                          " basis vectors");
    const bool cplx = dtype == QLG_C128;
    // coefficients then basis pointers, in the library's fixed staging array
    double *stage = noalloc_scratch(c.device);
    double *dcoef = stage;
    const double **dptr = reinterpret_cast<const double **>(stage + 2 * kNoAllocMaxBases);
    if (k > 0) {
      QLG_HIP(hipMemcpyAsync(dcoef, coefs, k * (cplx ? 2 : 1) * sizeof(double), hipMemcpyHostToDevice, c.stream));
      QLG_HIP(hipMemcpyAsync(dptr, bases, k * sizeof(void *), hipMemcpyHostToDevice, c.stream));
    }
    const double br = beta[0], bi = cplx ? beta[1] : 0.0;
    vec_lincomb(c, dtype, n, (int)k, dcoef, dptr, cplx, br, bi, br != 0.0 || bi != 0.0, static_cast<double *>(y));
    // the host tables are read by the copies above, and the staging array is
    // shared by every call on this device: return only when the work is done
    QLG_HIP(hipStreamSynchronize(c.stream));
  });
}

int qlg_lanczos_gs(qlg_plan plan, void *psi_inout, const qlg_lanczos_params *params, qlg_lanczos_result *res) {
  if (!plan || !psi_inout || !params || !res) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.kind == PLAN_HEFF, "qlg_lanczos_gs needs an effective-Hamiltonian plan");
    lanczos_gs(p, static_cast<double *>(psi_inout), *params, *res);
  });
}

int qlg_lanczos_expmv(qlg_plan plan, void *psi_inout, double tau, const qlg_lanczos_params *params, int64_t *iters) {
  if (!plan || !psi_inout || !params || !iters) return QLG_ERR_INVALID;
  Plan &p = *plan->p;
  return guarded(p.ctx, [&] {
    QLG_CHECK(p.kind == PLAN_HEFF, "qlg_lanczos_expmv needs an effective-Hamiltonian plan");
    lanczos_expmv(p, static_cast<double *>(psi_inout), tau, *params, *iters);
  });
}

int qlg_block_svd(qlg_ctx ctx, const qlg_bsten *t, int32_t ldims, const int32_t *ldiv, double trunc_err,
                  int64_t Dmin, int64_t Dmax, qlg_svd *out) {
  if (!ctx || !t || !ldiv || !out) return QLG_ERR_INVALID;
  *out = nullptr;
  return guarded(&ctx->c, [&] {
    auto h = std::make_unique<qlg_svd_s>();
    h->r.reset(block_svd(ctx->c, t, ldims, ldiv, trunc_err, Dmin, Dmax));
    *out = h.release();
  });
}

int qlg_svd_result(qlg_svd r, double *trunc_err, int64_t *D, int64_t *bond_nsect) {
  if (!r) return QLG_ERR_INVALID;
  if (trunc_err) *trunc_err = r->r->trunc_err;
  if (D) *D = r->r->D;
  if (bond_nsect) *bond_nsect = (int64_t)r->r->bond_deg.size();
  return QLG_OK;
}

int qlg_svd_bond(qlg_svd r, int32_t *qn, int64_t *deg) {
  if (!r) return QLG_ERR_INVALID;
  const SvdResult &s = *r->r;
  for (size_t i = 0; i < s.bond_deg.size(); ++i) {
    if (qn)
      for (int k = 0; k < s.qn_len; ++k) qn[i * s.qn_len + k] = s.bond_qn[i][k];
    if (deg) deg[i] = s.bond_deg[i];
  }
  return QLG_OK;
}

int qlg_svd_tensor(qlg_svd r, int which, int64_t *nblocks, int64_t *size, int32_t *coords, int64_t *offsets,
                   const void **data) {
  if (!r || which < 0 || which > 2) return QLG_ERR_INVALID;
  const SvdResult &s = *r->r;
  const int64_t nb = (int64_t)s.offsets[which].size();
  if (nblocks) *nblocks = nb;
  if (size) *size = s.size[which];
  if (coords)
    for (size_t i = 0; i < s.coords[which].size(); ++i) coords[i] = s.coords[which][i];
  if (offsets)
    for (int64_t i = 0; i < nb; ++i) offsets[i] = s.offsets[which][i];
  if (data) *data = s.data[which].p;
  return QLG_OK;
}

int qlg_svd_destroy(qlg_svd r) {
  if (!r) return QLG_OK;
  Ctx *c = r->r->ctx;
  if (c) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
  }
  delete r;
  return QLG_OK;
}

int qlg_eigh_batched(qlg_ctx ctx, int dtype, int64_t ngroups, const qlg_eigh_group *groups) {
  if (!ctx) return QLG_ERR_INVALID;
  return guarded(&ctx->c, [&] { eigh_batched(ctx->c, dtype, ngroups, groups); });
}

int qlg_tridiag_gs(const double *a, const double *b, int64_t n, char jobz, double *eng, double *vec) {
  if (!a || (n > 1 && !b) || !eng || n < 1 || (jobz != 'N' && jobz != 'V') || (jobz == 'V' && !vec))
    return QLG_ERR_INVALID;
  return tridiag_gs(n, a, b, jobz == 'V', eng, vec) ? QLG_OK : QLG_ERR_NUMERIC;
}

int qlg_tridiag_expme1(const double *a, const double *b, int64_t n, double tau, double *res) {
  if (!a || (n > 1 && !b) || !res || n < 1) return QLG_ERR_INVALID;
  return tridiag_expme1(n, a, b, tau, res) ? QLG_OK : QLG_ERR_NUMERIC;
}

}  // extern "C"
