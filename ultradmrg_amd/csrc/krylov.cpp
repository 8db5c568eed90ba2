// Krylov drivers on the GPU: Lanczos ground state and Lanczos exp(-i tau H)
// psi.  The Krylov basis lives in one plan-owned device buffer; vector work
// and the matvec run on the context stream; each iteration performs exactly
// one device-to-host copy (the new diagonal element and the residual norm)
// and solves the small FP64 tridiagonal problem on the host.
//
// Recurrences follow include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:114-248
// and include/qlmps/algorithm/tdvp/lanczos_expmv_solver_impl.h:70-184:
//   gamma_m = H v_{m-1} - a_{m-1} v_{m-1} - b_{m-2} v_{m-2},  b_{m-1} = ||gamma_m||,
//   v_m = gamma_m / b_{m-1}.
// The residual of iteration m+1 is formed right after a_m is known (reading
// a_m and b_{m-1} from device memory), so its norm arrives with a_m in the
// same copy.  If the solver stops at m, that speculative residual is simply
// discarded: the ground state only combines v_0..v_m.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "krylov.h"
#include "tridiag.h"
#include "vecops.h"

namespace qlg {

namespace {

struct Workspace {
  Plan &p;
  int64_t n, nd;
  int64_t nslots;
  double *slot0;
  // device scalars: a_m and b_m side by side (one copy fetches both), then
  // scratch for norms and dots
  double *ab, *dev_tmp;
  double *A(int64_t m) { return ab + 2 * m; }
  double *B(int64_t m) { return ab + 2 * m + 1; }
  Workspace(Plan &pl, int64_t maxit) : p(pl) {
    n = p.vec.size;
    nd = n * elem_doubles(p.dtype);
    nslots = maxit + 1;
    if (p.basis_vectors < nslots) {
      p.basis.release();
      p.basis.ensure((size_t)nslots * nd * sizeof(double));
      p.basis_vectors = nslots;
    }
    slot0 = p.basis.as<double>();
    p.scal_buf.ensure((size_t)(2 * (maxit + 2) + 8) * sizeof(double));
    ab = p.scal_buf.as<double>();
    dev_tmp = ab + 2 * (maxit + 2);
    p.coef_buf.ensure((size_t)(2 * (maxit + 2)) * sizeof(double));
    p.ptr_buf.ensure((size_t)(maxit + 2) * sizeof(double *));
  }
  double *slot(int64_t i) { return slot0 + i * nd; }
  // copy nvals doubles from device to host and wait
  void fetch(const double *dev, double *host, int nvals) {
    Ctx &c = *p.ctx;
    QLG_HIP(hipMemcpyAsync(c.host_scalars, dev, nvals * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QLG_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < nvals; ++i) host[i] = c.host_scalars[i];
  }
  // y (canonical layout) <- sum_i coef_i * slot(i) for i < k (real or complex coefficients)
  void combine(int64_t k, const double *coefs, bool cplx_coef, double *y_canonical) {
    Ctx &c = *p.ctx;
    p.vtmp_out.ensure((size_t)nd * sizeof(double));
    double *y = p.vtmp_out.as<double>();
    std::vector<const double *> ptrs(k);
    for (int64_t i = 0; i < k; ++i) ptrs[i] = slot(i);
    QLG_HIP(hipMemcpyAsync(p.coef_buf.p, coefs, k * (cplx_coef ? 2 : 1) * sizeof(double), hipMemcpyHostToDevice,
                           c.stream));
    QLG_HIP(hipMemcpyAsync(p.ptr_buf.p, ptrs.data(), k * sizeof(double *), hipMemcpyHostToDevice, c.stream));
    vec_lincomb(c, p.dtype, n, (int)k, p.coef_buf.as<double>(), p.ptr_buf.as<const double *>(), cplx_coef, 0.0, 0.0,
                false, y);
    p.to_canonical(y, y_canonical);
    // the host arrays above must stay alive until the copies are done
    QLG_HIP(hipStreamSynchronize(c.stream));
  }
  // FP64 Rayleigh quotient <g|H|g> / <g|g> of the fused vector the last
  // combine() formed (the basis slots are free by then).  Used by the
  // reduced-precision modes: the energy error becomes second order in the
  // error of the ground-state vector.
  double rayleigh_f64() {
    Ctx &c = *p.ctx;
    const double *g = p.vtmp_out.as<double>();
    const int prec = p.precision;
    p.precision = QLG_PREC_F64;
    p.matvec_fused(g, slot(0));
    p.precision = prec;
    vec_dot(c, p.dtype, n, slot(0), g, dev_tmp);
    vec_nrm2(c, p.dtype, n, g, dev_tmp + 1);
    double h[2];
    fetch(dev_tmp, h, 2);
    QLG_CHECK(h[1] > 0.0, "ground state has zero norm");
    return h[0] / (h[1] * h[1]);
  }
};

}  // namespace

void lanczos_gs(Plan &p, double *psi, const qlg_lanczos_params &prm, qlg_lanczos_result &res) {
  QLG_CHECK(prm.max_iterations >= 2, "max_iterations must be >= 2");
  Ctx &c = *p.ctx;
  const int64_t maxit = prm.max_iterations;
  Workspace w(p, maxit);
  const int64_t dim = w.n;
  std::vector<double> a(maxit + 1, 0.0), b(maxit + 1, 0.0), N(maxit + 1, 0.0);
  double h[2];

  // v0 = psi / ||psi||   (QuasiNormalize, lanczos_vmps_solver_impl.h:141)
  p.to_fused(psi, w.slot(0));  // the Krylov basis uses the fused layout
  vec_nrm2(c, p.dtype, w.n, w.slot(0), w.dev_tmp);
  w.fetch(w.dev_tmp, h, 1);
  QLG_CHECK(h[0] > 0.0, "initial state has zero norm");
  vec_scal(c, p.dtype, w.n, 1.0 / h[0], 0.0, w.slot(0));
  // a_0 and the first residual
  p.matvec_fused(w.slot(0), w.slot(1));
  vec_dot(c, p.dtype, w.n, w.slot(1), w.slot(0), w.A(0));
  vec_lanczos_residual(c, p.dtype, w.n, w.slot(1), w.slot(0), w.A(0), nullptr, nullptr, w.B(0));
  w.fetch(w.A(0), h, 2);
  a[0] = h[0];
  double norm_gamma = h[1];
  double energy0 = a[0];
  int64_t m = 0;
  while (true) {
    m += 1;
    // slot m holds gamma_m (un-normalised residual), norm_gamma = ||gamma_m||
    if (norm_gamma == 0.0) {
      if (m == 1) {
        p.to_canonical(w.slot(0), psi);
        if (p.precision != QLG_PREC_F64) {
          p.vtmp_out.ensure((size_t)w.nd * sizeof(double));
          QLG_HIP(hipMemcpyAsync(p.vtmp_out.p, w.slot(0), (size_t)w.nd * sizeof(double), hipMemcpyDeviceToDevice,
                                 c.stream));
          energy0 = w.rayleigh_f64();
        }
        QLG_HIP(hipStreamSynchronize(c.stream));
        res.iters = m;
        res.gs_eng = energy0;
        return;
      }
      std::vector<double> vec(m);
      double e;
      if (!tridiag_gs(m, a.data(), b.data(), true, &e, vec.data()))
        throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
      w.combine(m, vec.data(), false, psi);
      res.iters = m;
      res.gs_eng = p.precision == QLG_PREC_F64 ? energy0 : w.rayleigh_f64();
      return;
    }
    N[m] = norm_gamma * norm_gamma;
    b[m - 1] = norm_gamma;
    vec_scal(c, p.dtype, w.n, 1.0 / norm_gamma, 0.0, w.slot(m));
    p.matvec_fused(w.slot(m), w.slot(m + 1));
    vec_dot(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), w.A(m));
    // speculative residual for iteration m+1: H v_m - a_m v_m - b_{m-1} v_{m-1}
    vec_lanczos_residual(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), w.A(m), w.slot(m - 1), w.B(m - 1), w.B(m));
    w.fetch(w.A(m), h, 2);
    a[m] = h[0];
    norm_gamma = h[1];
    double e_new;
    if (!tridiag_gs(m + 1, a.data(), b.data(), false, &e_new, nullptr))
      throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
    if ((energy0 - e_new) < prm.error || m == dim || m == maxit - 1) {
      std::vector<double> vec(m + 1);
      if (!tridiag_gs(m + 1, a.data(), b.data(), true, &e_new, vec.data()))
        throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
      w.combine(m + 1, vec.data(), false, psi);
      res.iters = m + 1;
      res.gs_eng = p.precision == QLG_PREC_F64 ? e_new : w.rayleigh_f64();
      return;
    }
    energy0 = e_new;
  }
}

// Batched Lanczos ground states of independent chains (a batch plan).  Every
// chain follows the single-chain recurrence and stop rule above with its own
// scalars; the chains advance in lock step, so each iteration is one grouped
// matvec plus segmented vector kernels for all of them.  A chain that has
// stopped gets a zero scale factor, so its segment of later basis vectors is
// zero and costs nothing but the launch slots.  With graphs enabled, the
// device work of iteration m is captured once as a hipGraph and replayed by
// later calls on the same plan.
void lanczos_gs_batch(Plan &p, double *psi, const qlg_lanczos_params &prm, qlg_lanczos_result *res) {
  QLG_CHECK(prm.max_iterations >= 2, "max_iterations must be >= 2");
  QLG_CHECK(p.batch_off.size() >= 2, "batched Lanczos needs a batch plan");
  QLG_CHECK(p.dtype == QLG_F64 && p.precision == QLG_PREC_F64, "batched Lanczos runs real FP64 chains");
  Ctx &c = *p.ctx;
  hipStream_t st = c.stream;
  const int ns = (int)p.batch_off.size() - 1;
  const int64_t maxit = prm.max_iterations;
  Workspace w(p, maxit);
  // chunk table
  std::vector<SegChunk> chunks;
  std::vector<int32_t> cbeg(ns + 1, 0);
  for (int s = 0; s < ns; ++s) {
    cbeg[s] = (int32_t)chunks.size();
    for (int64_t b = p.batch_off[s]; b < p.batch_off[s + 1]; b += kSegChunk)
      chunks.push_back({b, std::min(b + kSegChunk, p.batch_off[s + 1]), s, 0});
  }
  cbeg[ns] = (int32_t)chunks.size();
  const int nch = (int)chunks.size();
  QLG_CHECK(nch > 0, "batched Lanczos: empty vectors");
  // device workspace: chunks | chunk_begin | scalars ab[(maxit+1)][2 ns] | factors f[(maxit+1)][ns] |
  // partials[nch] | coefficient table | basis pointers
  const size_t o_chunks = 0, o_cbeg = o_chunks + sizeof(SegChunk) * nch;
  const size_t o_ab = (o_cbeg + sizeof(int32_t) * (ns + 1) + 255) / 256 * 256;
  const size_t o_f = o_ab + sizeof(double) * 2 * ns * (maxit + 1);
  const size_t o_part = o_f + sizeof(double) * ns * (maxit + 1);
  const size_t o_coef = o_part + sizeof(double) * nch;
  const size_t o_ptr = o_coef + sizeof(double) * ns * (maxit + 1);
  const size_t total = o_ptr + sizeof(double *) * (maxit + 1);
  if (p.seg_ws.bytes < total || p.graph_maxit != maxit || p.graph_basis != w.slot0) p.drop_graphs();
  p.seg_ws.ensure(total);
  p.graph_maxit = maxit;
  p.graph_basis = w.slot0;
  char *ws = p.seg_ws.as<char>();
  SegTable tab;
  tab.chunks = reinterpret_cast<const SegChunk *>(ws + o_chunks);
  tab.chunk_begin = reinterpret_cast<const int32_t *>(ws + o_cbeg);
  tab.nchunks = nch;
  tab.nseg = ns;
  double *ab = reinterpret_cast<double *>(ws + o_ab);
  double *fac = reinterpret_cast<double *>(ws + o_f);
  double *part = reinterpret_cast<double *>(ws + o_part);
  double *coefs = reinterpret_cast<double *>(ws + o_coef);
  const double **ptrs = reinterpret_cast<const double **>(ws + o_ptr);
  if (p.host_seg_n < (size_t)std::max<int64_t>(2 * ns, (maxit + 1) * ns)) {
    if (p.host_seg) (void)hipHostFree(p.host_seg);
    p.host_seg = nullptr;
    p.host_seg_n = 0;
    const size_t nh = (size_t)std::max<int64_t>(2 * ns, (maxit + 1) * ns);
    QLG_HIP(hipHostMalloc((void **)&p.host_seg, nh * sizeof(double), hipHostMallocDefault));
    p.host_seg_n = nh;
  }
  double *hs = p.host_seg;
  QLG_HIP(hipMemcpyAsync(ws + o_chunks, chunks.data(), sizeof(SegChunk) * nch, hipMemcpyHostToDevice, st));
  QLG_HIP(hipMemcpyAsync(ws + o_cbeg, cbeg.data(), sizeof(int32_t) * (ns + 1), hipMemcpyHostToDevice, st));
  auto fetch = [&](const double *dev, int nvals, std::vector<double> &out) {
    QLG_HIP(hipMemcpyAsync(hs, dev, nvals * sizeof(double), hipMemcpyDeviceToHost, st));
    QLG_HIP(hipStreamSynchronize(st));
    out.assign(hs, hs + nvals);
  };
  auto upload_row = [&](double *dev, const std::vector<double> &v) {
    for (int s = 0; s < ns; ++s) hs[s] = v[s];
    QLG_HIP(hipMemcpyAsync(dev, hs, ns * sizeof(double), hipMemcpyHostToDevice, st));
    QLG_HIP(hipStreamSynchronize(st));   // hs is reused by the next fetch
  };
  // per-chain state
  std::vector<std::vector<double>> a(ns, std::vector<double>(maxit + 1, 0.0)), b = a;
  std::vector<double> energy0(ns), norm_gamma(ns), f(ns), h;
  std::vector<int> active(ns, 1);
  std::vector<std::vector<double>> fin(ns);
  // v0 = psi / ||psi|| per chain
  p.to_fused(psi, w.slot(0));
  seg_dot(st, tab, w.slot(0), w.slot(0), part, ab, true);
  fetch(ab, ns, h);
  for (int s = 0; s < ns; ++s) {
    QLG_CHECK(h[s] > 0.0, "initial state of chain " + std::to_string(s) + " has zero norm");
    f[s] = 1.0 / h[s];
  }
  upload_row(fac, f);
  seg_scal(st, tab, fac, w.slot(0));
  p.matvec_fused(w.slot(0), w.slot(1));   // also settles the tile shapes before any capture
  seg_dot(st, tab, w.slot(1), w.slot(0), part, ab, false);
  seg_lanczos_residual(st, tab, w.slot(1), w.slot(0), ab, nullptr, nullptr, part, ab + ns);
  fetch(ab, 2 * ns, h);
  for (int s = 0; s < ns; ++s) {
    a[s][0] = h[s];
    norm_gamma[s] = h[ns + s];
    energy0[s] = a[s][0];
  }
  auto step = [&](int64_t m) {
    seg_scal(st, tab, fac + m * ns, w.slot(m));
    p.matvec_fused(w.slot(m), w.slot(m + 1));
    seg_dot(st, tab, w.slot(m + 1), w.slot(m), part, ab + m * 2 * ns, false);
    seg_lanczos_residual(st, tab, w.slot(m + 1), w.slot(m), ab + m * 2 * ns, w.slot(m - 1),
                         ab + (m - 1) * 2 * ns + ns, part, ab + m * 2 * ns + ns);
  };
  int64_t m = 0;
  while (true) {
    m += 1;
    bool any = false;
    for (int s = 0; s < ns; ++s) {
      if (!active[s]) {
        f[s] = 0.0;
        continue;
      }
      if (norm_gamma[s] == 0.0) {   // invariant subspace
        active[s] = 0;
        f[s] = 0.0;
        res[s].iters = m;
        res[s].gs_eng = energy0[s];
        if (m == 1) {
          fin[s] = {1.0};
        } else {
          std::vector<double> vec(m);
          double e;
          if (!tridiag_gs(m, a[s].data(), b[s].data(), true, &e, vec.data()))
            throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
          fin[s] = vec;
        }
        continue;
      }
      b[s][m - 1] = norm_gamma[s];
      f[s] = 1.0 / norm_gamma[s];
      any = true;
    }
    if (!any) break;
    upload_row(fac + m * ns, f);
    if (p.use_graphs) {
      if ((int64_t)p.step_graphs.size() <= m) p.step_graphs.resize(m + 1, nullptr);
      if (!p.step_graphs[m]) {
        hipGraph_t g = nullptr;
        QLG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        p.capturing = true;
        try {
          step(m);
        } catch (...) {
          p.capturing = false;
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        p.capturing = false;
        QLG_HIP(hipStreamEndCapture(st, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        QLG_HIP(e);
        p.step_graphs[m] = ex;
      }
      QLG_HIP(hipGraphLaunch(p.step_graphs[m], st));
    } else {
      step(m);
    }
    fetch(ab + m * 2 * ns, 2 * ns, h);
    for (int s = 0; s < ns; ++s) {
      if (!active[s]) continue;
      a[s][m] = h[s];
      norm_gamma[s] = h[ns + s];
      double e_new;
      if (!tridiag_gs(m + 1, a[s].data(), b[s].data(), false, &e_new, nullptr))
        throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
      const int64_t dim = p.batch_off[s + 1] - p.batch_off[s];
      if ((energy0[s] - e_new) < prm.error || m == dim || m == maxit - 1) {
        std::vector<double> vec(m + 1);
        if (!tridiag_gs(m + 1, a[s].data(), b[s].data(), true, &e_new, vec.data()))
          throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
        fin[s] = vec;
        active[s] = 0;
        res[s].iters = m + 1;
        res[s].gs_eng = e_new;
      } else {
        energy0[s] = e_new;
      }
    }
  }
  // ground states: per-chain combination of its own basis vectors
  int64_t k = 0;
  for (int s = 0; s < ns; ++s) k = std::max<int64_t>(k, (int64_t)fin[s].size());
  for (int64_t j = 0; j < k; ++j)
    for (int s = 0; s < ns; ++s) hs[j * ns + s] = j < (int64_t)fin[s].size() ? fin[s][j] : 0.0;
  QLG_HIP(hipMemcpyAsync(coefs, hs, sizeof(double) * k * ns, hipMemcpyHostToDevice, st));
  std::vector<const double *> bp(k);
  for (int64_t j = 0; j < k; ++j) bp[j] = w.slot(j);
  QLG_HIP(hipMemcpyAsync(ptrs, bp.data(), sizeof(double *) * k, hipMemcpyHostToDevice, st));
  p.vtmp_out.ensure((size_t)w.nd * sizeof(double));
  seg_lincomb(st, tab, (int)k, coefs, ptrs, p.vtmp_out.as<double>());
  p.to_canonical(p.vtmp_out.as<double>(), psi);
  QLG_HIP(hipStreamSynchronize(st));
}

void lanczos_expmv(Plan &p, double *psi, double tau, const qlg_lanczos_params &prm, int64_t &iters) {
  QLG_CHECK(p.dtype == QLG_C128, "Lanczos expmv needs a complex (QLG_C128) plan");
  QLG_CHECK(prm.max_iterations >= 2, "max_iterations must be >= 2");
  Ctx &c = *p.ctx;
  const int64_t maxit = prm.max_iterations;
  Workspace w(p, maxit);
  const int64_t dim = w.n;
  std::vector<double> a(maxit + 1, 0.0), b(maxit + 1, 0.0);
  std::vector<double> coef(2 * (maxit + 1), 0.0), last(2 * (maxit + 1), 0.0);
  double h[2];

  // Normalize keeps the initial norm (lanczos_expmv_solver_impl.h:93)
  p.to_fused(psi, w.slot(0));  // the Krylov basis uses the fused layout
  vec_nrm2(c, p.dtype, w.n, w.slot(0), w.dev_tmp);
  w.fetch(w.dev_tmp, h, 1);
  const double initial_norm = h[0];
  QLG_CHECK(initial_norm > 0.0, "initial state has zero norm");
  vec_scal(c, p.dtype, w.n, 1.0 / initial_norm, 0.0, w.slot(0));
  p.matvec_fused(w.slot(0), w.slot(1));
  // a_m = Re <H v_m | v_m>: the complex dot writes (re, im) over (a_m, b_m);
  // the residual reads the real part and then overwrites the imaginary one
  // with b_m
  vec_dot(c, p.dtype, w.n, w.slot(1), w.slot(0), w.A(0));
  vec_lanczos_residual(c, p.dtype, w.n, w.slot(1), w.slot(0), w.A(0), nullptr, nullptr, w.B(0));
  w.fetch(w.A(0), h, 2);
  a[0] = h[0];
  double norm_gamma = h[1];
  int64_t m = 0;
  while (true) {
    m += 1;
    if (norm_gamma == 0.0) {
      iters = m;
      if (m == 1) {
        // psi <- ||psi0|| exp(-i tau a_0) v_0
        const double ph = -tau * a[0];
        p.to_canonical(w.slot(0), psi);
        vec_scal(c, p.dtype, w.n, initial_norm * std::cos(ph), initial_norm * std::sin(ph), psi);
        QLG_HIP(hipStreamSynchronize(c.stream));
        return;
      }
      if (!tridiag_expme1(m, a.data(), b.data(), tau, coef.data()))
        throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
      for (int64_t i = 0; i < 2 * m; ++i) coef[i] *= initial_norm;
      w.combine(m, coef.data(), true, psi);
      return;
    }
    b[m - 1] = norm_gamma;
    vec_scal(c, p.dtype, w.n, 1.0 / norm_gamma, 0.0, w.slot(m));
    p.matvec_fused(w.slot(m), w.slot(m + 1));
    vec_dot(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), w.A(m));
    vec_lanczos_residual(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), w.A(m), w.slot(m - 1), w.B(m - 1), w.B(m));
    w.fetch(w.A(m), h, 2);
    a[m] = h[0];
    norm_gamma = h[1];
    std::fill(coef.begin(), coef.end(), 0.0);
    if (!tridiag_expme1(m + 1, a.data(), b.data(), tau, coef.data()))
      throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
    double dist = 0.0;
    for (int64_t i = 0; i < 2 * (m + 1); ++i) dist += (coef[i] - last[i]) * (coef[i] - last[i]);
    dist = std::sqrt(dist);
    if (dist < prm.error || m == dim || m == maxit - 1) {
      iters = m + 1;
      for (int64_t i = 0; i < 2 * (m + 1); ++i) coef[i] *= initial_norm;
      w.combine(m + 1, coef.data(), true, psi);
      return;
    }
    last = coef;
  }
}

}  // namespace qlg
