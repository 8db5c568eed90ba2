// Krylov drivers on the GPU: Lanczos ground state and Lanczos exp(-i tau H)
// psi.  The Krylov basis lives in one plan-owned device buffer; vector work
// and the matvec run on the context stream.
//
// Single-chain solves run ahead of the host (lanczos_gs / lanczos_expmv):
// every iteration ends with a device kernel that solves the FP64 tridiagonal
// problem and applies the stop rule (krylov_dev.hip), and the host queues
// iterations in chunks, looking at the device's stop flag once per chunk;
// kernels queued past the stop return at once.  With graphs enabled on the
// plan, each chunk is captured as a hipGraph on first use and replayed by
// later solves.  Optional full reorthogonalisation (classical Gram-Schmidt
// against the whole basis, once or twice per iteration) is off by default,
// as in the reference.  The host-iteration forms (one device-to-host copy and
// a host tridiagonal solve per iteration) remain for Krylov dimensions beyond
// the device kernels' limits and for QLG_KRYLOV_HOST=1.
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
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "krylov.h"
#include "krylov_dev.h"
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

static void lanczos_gs_host(Plan &p, double *psi, const qlg_lanczos_params &prm, qlg_lanczos_result &res) {
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

static void lanczos_expmv_host(Plan &p, double *psi, double tau, const qlg_lanczos_params &prm, int64_t &iters) {
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


// ---------------------------------------------------------------------------
// run-ahead single-chain solves
// ---------------------------------------------------------------------------
namespace {

// iterations queued per look at the stop flag (QLG_LANCZOS_CHUNK, default 8)
int chunk_iterations() {
  const char *v = std::getenv("QLG_LANCZOS_CHUNK");
  const int x = v ? std::atoi(v) : 8;
  return x >= 1 ? x : 8;
}

// QLG_KRYLOV_HOST=1: the iteration-by-iteration host form
bool host_iteration_forced() {
  const char *v = std::getenv("QLG_KRYLOV_HOST");
  return v && v[0] == '1';
}

// Device workspace of a run-ahead solve (in Plan::lz_ws):
//   [LzCtl][ab: (a_k, b_k) pairs, maxit + 2][tmp: 8][coef_cur][coef_prev]
//   (2 (maxit + 1) doubles each)[rc: Gram-Schmidt coefficients, 2 (maxit + 2)]
//   [basis slot pointers, maxit + 1]
struct RunAhead {
  Plan &p;
  Workspace w;
  LzCtl *ctl;
  double *ab, *tmp, *coef_cur, *coef_prev, *rc, *rpart = nullptr;
  const double **ptrs;
  int64_t maxit;
  RunAhead(Plan &pl, int64_t mi) : p(pl), w(pl, mi), maxit(mi) {
    const size_t n_ab = 2 * (maxit + 2), n_coef = 2 * (maxit + 1);
    const size_t o_ab = 256, o_tmp = o_ab + n_ab * 8, o_cur = o_tmp + 64, o_prev = o_cur + n_coef * 8,
                 o_rc = o_prev + n_coef * 8, o_ptr = (o_rc + n_ab * 8 + 255) / 256 * 256,
                 total = o_ptr + (maxit + 1) * sizeof(double *);
    p.lz_ws.ensure(total);
    char *b = p.lz_ws.as<char>();
    ctl = reinterpret_cast<LzCtl *>(b);
    ab = reinterpret_cast<double *>(b + o_ab);
    tmp = reinterpret_cast<double *>(b + o_tmp);
    coef_cur = reinterpret_cast<double *>(b + o_cur);
    coef_prev = reinterpret_cast<double *>(b + o_prev);
    rc = reinterpret_cast<double *>(b + o_rc);
    ptrs = reinterpret_cast<const double **>(b + o_ptr);
    if (p.reorth > 0) {
      p.reorth_ws.ensure((size_t)(maxit + 2) * kMaxPartials * 2 * sizeof(double));
      rpart = p.reorth_ws.as<double>();
    }
    const size_t nh = 8 + n_ab;
    if (p.host_lz_n < nh) {
      if (p.host_lz) (void)hipHostFree(p.host_lz);
      p.host_lz = nullptr;
      p.host_lz_n = 0;
      QLG_HIP(hipHostMalloc((void **)&p.host_lz, nh * sizeof(double), hipHostMallocDefault));
      p.host_lz_n = nh;
    }
    std::vector<const double *> hp(maxit + 1);
    for (int64_t i = 0; i <= maxit; ++i) hp[i] = w.slot(i);
    Ctx &c = *p.ctx;
    QLG_HIP(hipMemcpyAsync(ptrs, hp.data(), hp.size() * sizeof(double *), hipMemcpyHostToDevice, c.stream));
    QLG_HIP(hipStreamSynchronize(c.stream));   // hp is a host temporary
  }
  double *A(int64_t m) { return ab + 2 * m; }
  double *B(int64_t m) { return ab + 2 * m + 1; }
  LzCtl fetch_ctl() {
    Ctx &c = *p.ctx;
    QLG_HIP(hipMemcpyAsync(p.host_lz, ctl, sizeof(LzCtl), hipMemcpyDeviceToHost, c.stream));
    QLG_HIP(hipStreamSynchronize(c.stream));
    LzCtl h;
    std::memcpy(&h, p.host_lz, sizeof(LzCtl));
    return h;
  }
  void fetch_ab(int64_t n, std::vector<double> &a, std::vector<double> &b) {
    Ctx &c = *p.ctx;
    QLG_HIP(hipMemcpyAsync(p.host_lz, ab, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QLG_HIP(hipStreamSynchronize(c.stream));
    a.assign(n, 0.0);
    b.assign(n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      a[i] = p.host_lz[2 * i];
      b[i] = p.host_lz[2 * i + 1];
    }
  }
  // residual of iteration m in slot m+1 is final: reorthogonalise it and
  // replace b_m by its norm
  void reorth(int64_t m, const int *skip) {
    for (int pass = 0; pass < p.reorth; ++pass)
      lz_reorth(p.ctx->stream, p.dtype, w.n, ptrs, (int)(m + 1), w.slot(m + 1), rpart, rc, B(m), skip);
  }
  // queue iterations [m0, m1) with `body`, as a captured graph when the plan
  // uses graphs (one graph per chunk index, replayed by later solves with the
  // same key)
  template <class F>
  void run_chunk(int64_t m0, int64_t m1, int64_t chunk, const std::string &key, F body) {
    Ctx &c = *p.ctx;
    if (!p.use_graphs) {
      for (int64_t m = m0; m < m1; ++m) body(m);
      return;
    }
    if (p.chunk_key != key) {
      for (auto &g : p.chunk_graphs)
        if (g) (void)hipGraphExecDestroy(g);
      p.chunk_graphs.clear();
      p.chunk_key = key;
    }
    if ((int64_t)p.chunk_graphs.size() <= chunk) p.chunk_graphs.resize(chunk + 1, nullptr);
    if (!p.chunk_graphs[chunk]) {
      hipGraph_t g = nullptr;
      QLG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      p.capturing = true;
      try {
        for (int64_t m = m0; m < m1; ++m) body(m);
      } catch (...) {
        p.capturing = false;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      p.capturing = false;
      QLG_HIP(hipStreamEndCapture(c.stream, &g));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      QLG_HIP(e);
      p.chunk_graphs[chunk] = ex;
    }
    QLG_HIP(hipGraphLaunch(p.chunk_graphs[chunk], c.stream));
  }
};

// the plan's matvecs read the solve's stop flag while it is queued
struct SkipGuard {
  Plan &p;
  SkipGuard(Plan &pl, const int *flag) : p(pl) { p.skip_flag = flag; }
  ~SkipGuard() { p.skip_flag = nullptr; }
};

std::string graph_key(const RunAhead &r, const char *what, double err, double tau) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s|%lld|%p|%p|%d|%d|%.17g|%.17g|%lld|%d", what, (long long)r.maxit,
                (const void *)r.w.slot0, (const void *)r.ctl, r.p.precision, r.p.reorth, err, tau,
                (long long)r.w.n, chunk_iterations());
  return buf;
}

}  // namespace

void lanczos_gs(Plan &p, double *psi, const qlg_lanczos_params &prm, qlg_lanczos_result &res) {
  QLG_CHECK(prm.max_iterations >= 2, "max_iterations must be >= 2");
  if (host_iteration_forced() || prm.max_iterations > 2047) {
    QLG_CHECK(p.reorth == 0, "reorthogonalisation needs the device-resident solver (max_iterations <= 2047)");
    lanczos_gs_host(p, psi, prm, res);
    return;
  }
  Ctx &c = *p.ctx;
  hipStream_t st = c.stream;
  const int64_t maxit = prm.max_iterations;
  RunAhead r(p, maxit);
  Workspace &w = r.w;
  const int64_t dim = w.n;
  double h[2];
  // v0 = psi / ||psi||   (QuasiNormalize, lanczos_vmps_solver_impl.h:141)
  p.to_fused(psi, w.slot(0));
  vec_nrm2(c, p.dtype, w.n, w.slot(0), w.dev_tmp);
  w.fetch(w.dev_tmp, h, 1);
  QLG_CHECK(h[0] > 0.0, "initial state has zero norm");
  vec_scal(c, p.dtype, w.n, 1.0 / h[0], 0.0, w.slot(0));
  lz_reset(st, r.ctl);
  // iteration 0: a_0, the first residual and its norm
  p.matvec_fused(w.slot(0), w.slot(1));   // also settles the tile shapes before any capture
  {
    SkipGuard sg(p, &r.ctl->stop);
    vec_dot(c, p.dtype, w.n, w.slot(1), w.slot(0), r.A(0), &r.ctl->stop);
    vec_lanczos_residual(c, p.dtype, w.n, w.slot(1), w.slot(0), r.A(0), nullptr, nullptr, r.B(0), &r.ctl->stop);
    r.reorth(0, &r.ctl->stop);
    lz_gs_check(st, r.ab, r.ctl, 0, dim, maxit, prm.error);
    auto body = [&](int64_t m) {
      const int *skip = &r.ctl->stop;
      vec_scal_inv_dev(c, p.dtype, w.n, r.B(m - 1), w.slot(m), skip);
      p.matvec_fused(w.slot(m), w.slot(m + 1));
      vec_dot(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), r.A(m), skip);
      vec_lanczos_residual(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), r.A(m), w.slot(m - 1), r.B(m - 1), r.B(m),
                           skip);
      r.reorth(m, skip);
      lz_gs_check(st, r.ab, r.ctl, (int)m, dim, maxit, prm.error);
    };
    const int64_t K = chunk_iterations();
    const std::string key = graph_key(r, "gs", prm.error, 0.0);
    for (int64_t m0 = 1, chunk = 0;; m0 += K, ++chunk) {
      const int64_t m1 = std::min(m0 + K, maxit);
      if (m0 < m1) r.run_chunk(m0, m1, chunk, key, body);
      if (r.fetch_ctl().stop) break;
      QLG_CHECK(m1 < maxit, "device Lanczos: no stop by max_iterations");
    }
  }
  const LzCtl cs = r.fetch_ctl();
  const int64_t iters = cs.iters;
  QLG_CHECK(iters >= 1 && iters <= maxit, "device Lanczos: invalid stop");
  std::vector<double> a, b, vec(iters);
  r.fetch_ab(iters, a, b);
  double e;
  if (!tridiag_gs(iters, a.data(), b.data(), true, &e, vec.data()))
    throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
  w.combine(iters, vec.data(), false, psi);
  res.iters = iters;
  res.gs_eng = p.precision == QLG_PREC_F64 ? e : w.rayleigh_f64();
}

void lanczos_expmv(Plan &p, double *psi, double tau, const qlg_lanczos_params &prm, int64_t &iters) {
  QLG_CHECK(p.dtype == QLG_C128, "Lanczos expmv needs a complex (QLG_C128) plan");
  QLG_CHECK(prm.max_iterations >= 2, "max_iterations must be >= 2");
  if (host_iteration_forced() || prm.max_iterations > 64) {
    QLG_CHECK(p.reorth == 0, "reorthogonalisation needs the device-resident solver (max_iterations <= 64)");
    lanczos_expmv_host(p, psi, tau, prm, iters);
    return;
  }
  Ctx &c = *p.ctx;
  hipStream_t st = c.stream;
  const int64_t maxit = prm.max_iterations;
  RunAhead r(p, maxit);
  Workspace &w = r.w;
  const int64_t dim = w.n;
  double h[2];
  // Normalize keeps the initial norm (lanczos_expmv_solver_impl.h:93)
  p.to_fused(psi, w.slot(0));
  vec_nrm2(c, p.dtype, w.n, w.slot(0), w.dev_tmp);
  w.fetch(w.dev_tmp, h, 1);
  const double initial_norm = h[0];
  QLG_CHECK(initial_norm > 0.0, "initial state has zero norm");
  vec_scal(c, p.dtype, w.n, 1.0 / initial_norm, 0.0, w.slot(0));
  lz_reset(st, r.ctl);
  p.matvec_fused(w.slot(0), w.slot(1));
  {
    SkipGuard sg(p, &r.ctl->stop);
    // a_m = Re <H v_m | v_m>: the complex dot writes (re, im) over (a_m, b_m);
    // the residual reads the real part and then overwrites b_m with its norm
    vec_dot(c, p.dtype, w.n, w.slot(1), w.slot(0), r.A(0), &r.ctl->stop);
    vec_lanczos_residual(c, p.dtype, w.n, w.slot(1), w.slot(0), r.A(0), nullptr, nullptr, r.B(0), &r.ctl->stop);
    r.reorth(0, &r.ctl->stop);
    lz_expmv_check(st, r.ab, r.ctl, r.coef_cur, r.coef_prev, nullptr, 0, dim, maxit, prm.error, tau);
    auto body = [&](int64_t m) {
      const int *skip = &r.ctl->stop;
      vec_scal_inv_dev(c, p.dtype, w.n, r.B(m - 1), w.slot(m), skip);
      p.matvec_fused(w.slot(m), w.slot(m + 1));
      vec_dot(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), r.A(m), skip);
      vec_lanczos_residual(c, p.dtype, w.n, w.slot(m + 1), w.slot(m), r.A(m), w.slot(m - 1), r.B(m - 1), r.B(m),
                           skip);
      r.reorth(m, skip);
      lz_expmv_check(st, r.ab, r.ctl, r.coef_cur, r.coef_prev, nullptr, (int)m, dim, maxit, prm.error, tau);
    };
    const int64_t K = chunk_iterations();
    const std::string key = graph_key(r, "expmv", prm.error, tau);
    for (int64_t m0 = 1, chunk = 0;; m0 += K, ++chunk) {
      const int64_t m1 = std::min(m0 + K, maxit);
      if (m0 < m1) r.run_chunk(m0, m1, chunk, key, body);
      if (r.fetch_ctl().stop) break;
      QLG_CHECK(m1 < maxit, "device expmv: no stop by max_iterations");
    }
  }
  const LzCtl cs = r.fetch_ctl();
  if (cs.iters < 0) throw Error(QLG_ERR_NUMERIC, "tridiagonal eigen-solver did not converge");
  QLG_CHECK(cs.iters >= 1 && cs.iters <= maxit, "device expmv: invalid stop");
  iters = cs.iters;
  // psi <- ||psi_0|| sum_i c_i v_i with the device's coefficients
  p.vtmp_out.ensure((size_t)w.nd * sizeof(double));
  double *y = p.vtmp_out.as<double>();
  vec_lincomb(c, p.dtype, w.n, (int)iters, r.coef_cur, r.ptrs, true, 0.0, 0.0, false, y);
  vec_scal(c, p.dtype, w.n, initial_norm, 0.0, y);
  p.to_canonical(y, psi);
  QLG_HIP(hipStreamSynchronize(st));
}

}  // namespace qlg
