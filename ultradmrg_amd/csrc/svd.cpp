// Block SVD with truncation on the GPU: the qlten SVD call of the reference's
// two-site updates, SVD(t, ldims, div, trunc_err, Dmin, Dmax, u, s, vt,
// actual_trunc_err, D) (two_site_update_finite_vmps_impl.h:302-307,
// dmrg_impl.h:249-254, two_site_update_finite_tdvp_impl.h:339-344).
//
// The tensor is a direct sum of sector matrices M_q (rows: the left
// coordinates whose charge sets the new bond charge q = div - flow(left
// axes), columns: the right coordinates).  Each is gathered, shorter side
// first (transposed when it has more rows than columns), into padded batches
// of a few size buckets, scaled to unit Frobenius norm, and decomposed
// through its Gram matrix X X^H (grouped FP64 GEMM, rocSOLVER divide and
// conquer on four lanes, eigh.cpp):
//   X = E S P,  rows of R = E^H X are s_j p_j,  s_j = |R_j|
// (the back-projected row norm has absolute error ~eps s_max, where
// sqrt(lambda_j) would have eps s_max^2 / s_j).  Eigenvectors of tiny
// eigenvalues are mixtures within their cluster, and the rows of small values
// carry components of their neighbours' directions, which only matters for
// singular values below ~1e-6 of the sector's largest: when such values are
// kept, the sector's rows are finished by one-sided Jacobi rotations applied
// to R and E^H alike (values and vectors then resolve to ~eps of the sector's
// largest value, as a direct SVD does).  The kept rows P (unit by construction, but
// mutually orthogonal only to ~eps s_max^2 / (s_i s_j)) are orthonormalised
// by Newton-Schulz steps toward the polar factor, P <- 1.5 P - 0.5 (P P^H) P,
// until |P P^H - I| < 1e-13; rows of exactly zero singular values are
// completed to an orthonormal set.
//
// Truncation (the reference's qlten rule is not part of the reference
// sources; this library's): pool the singular values of all sectors, keep the
// smallest D whose discarded weight sum(s^2) / total <= trunc_err, clamped to
// [Dmin, Dmax].  Complex batches are rotated by two fixed Householder
// reflectors before the eigensolver (the complex divide-and-conquer solver
// returns NaN vectors on block-diagonal input) and rotated back after.
#include "svd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <numeric>

#include "eigh.h"
#include "plan.h"

namespace qlg {

namespace {

using Key = std::vector<int32_t>;

struct Sector {
  QN q{0, 0};
  std::map<Key, std::pair<int64_t, int64_t>> rows, cols;   // key -> (extent, offset)
  int64_t R = 0, C = 0, n = 0, m = 0;
  bool tr = false;
  int bucket = -1, slot = 0;
  int64_t sv0 = 0;   // first of its n values in the singular-value arrays
};

struct Bucket {
  int64_t b = 0, mp = 0;   // padded short side (Gram order) and long side
  std::vector<int> secs;
  int64_t xbase = 0, gbase = 0, rbase = 0;
};

int64_t bucket_bound(int64_t n) {
  static const int64_t bounds[] = {1, 256, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192};
  for (int64_t b : bounds)
    if (n <= b) return b;
  return (n + 4095) / 4096 * 4096;
}

int64_t up16(int64_t x) { return (x + 15) / 16 * 16; }

// device copy of a host vector on the context stream (the host data is
// copied before the call returns)
template <class T>
void upload(Ctx &c, DevBuf &d, const std::vector<T> &v) {
  d.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
  if (!v.empty()) QLG_HIP(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c.stream));
}

template <class T>
void download(Ctx &c, std::vector<T> &v, const void *src, size_t n) {
  v.resize(n);
  if (n) QLG_HIP(hipMemcpyAsync(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
  QLG_HIP(hipStreamSynchronize(c.stream));
}

// A batch of independent GEMMs C = op(A) op(B) through the grouped FP64 GEMM,
// operands addressed absolutely (slot 0)
struct GemmBatch {
  Ctx &c;
  bool cplx;
  Stage st;
  GemmBatch(Ctx &cx, bool cp) : c(cx), cplx(cp) {}
  int64_t el(const void *p) const { return (int64_t)(reinterpret_cast<uintptr_t>(p) / (cplx ? 16 : 8)); }
  // C[M x N] (ldc) = A[M x K] (lda; stored [K x M] if trans_a) * B[K x N] (ldb; stored [N x K] if trans_b)
  void add(const double *A, int64_t lda, bool ta, bool conj_a, const double *B, int64_t ldb, bool tb, bool conj_b,
           double *Cm, int64_t ldc, int64_t M, int64_t N, int64_t K) {
    if (M <= 0 || N <= 0) return;
    QLG_CHECK(M < (1 << 30) && N < (1 << 30) && K < (1 << 30), "block SVD: matrix too large");
    GemmSeg g{};
    g.a_slot = SLOT_ABS;
    g.a_off = el(A);
    g.lda = lda;
    g.b_slot = SLOT_ABS;
    g.b_off = el(B);
    g.ldb = ldb;
    g.K = (int32_t)K;
    g.flags = (ta ? SEG_TRANS_A : 0) | (tb ? SEG_TRANS_B : 0) | (conj_a ? SEG_CONJ_A : 0) | (conj_b ? SEG_CONJ_B : 0);
    GemmTask t{};
    t.c_slot = SLOT_ABS;
    t.c_off = el(Cm);
    t.ldc = ldc;
    t.M = (int32_t)M;
    t.N = (int32_t)N;
    t.seg_begin = (int32_t)st.segs.size();
    t.seg_count = K > 0 ? 1 : 0;
    if (K > 0) st.segs.push_back(g);
    st.tasks.push_back(t);
  }
  void run() {
    if (st.tasks.empty()) return;
    int64_t big = 0;
    for (const GemmTask &t : st.tasks) big = std::max<int64_t>(big, std::min(t.M, t.N));
    const int tile = big >= 128 ? 64 : 32;
    st.make_tiles(tile, tile);
    st.upload(c.stream);
    Slots s{};
    launch_grouped_gemm_f64(st.bm, st.bn, cplx, st.d_tiles.as<GemmTile>(), st.d_tasks.as<GemmTask>(),
                            st.d_segs.as<GemmSeg>(), s, (int)st.tiles.size(), c.stream);
  }
};

// deterministic unit vectors for the complex batches' reflectors (SplitMix64 +
// Box-Muller), two per order b, interleaved (re, im)
std::vector<double> reflector(int64_t b, int which, int attempt) {
  uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)(b * 4 + which + 1) + 0x632BE59BD9B4E019ull * (uint64_t)(attempt + 1);
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  std::vector<double> u(2 * b);
  double nrm = 0.0;
  for (int64_t i = 0; i < b; ++i) {
    const double r = std::sqrt(-2.0 * std::log(next())), th = 2.0 * M_PI * next();
    u[2 * i] = r * std::cos(th);
    u[2 * i + 1] = r * std::sin(th);
    nrm += u[2 * i] * u[2 * i] + u[2 * i + 1] * u[2 * i + 1];
  }
  nrm = std::sqrt(nrm);
  for (double &x : u) x /= nrm;
  return u;
}

}  // namespace

SvdResult *block_svd(Ctx &c, const qlg_bsten *desc, int ldims, const int32_t *ldiv, double trunc_err, int64_t Dmin,
                     int64_t Dmax) {
  QLG_CHECK(desc != nullptr, "block SVD: null tensor");
  const Ten t = parse_tensor(desc, desc->rank, "t");
  QLG_CHECK(ldims >= 1 && ldims < t.rank, "block SVD: ldims must split the axes");
  QLG_CHECK(ldiv != nullptr, "block SVD: null divergence");
  QLG_CHECK(trunc_err >= 0.0 && Dmin >= 1 && Dmax >= Dmin, "block SVD: need trunc_err >= 0 and 1 <= Dmin <= Dmax");
  const bool cplx = t.dtype == QLG_C128;
  const int ed = cplx ? 2 : 1;
  hipStream_t stream = c.stream;
  const char *dbg = std::getenv("QLG_SVD_DEBUG");
  const bool debug = dbg && dbg[0] == '1';
  // QLG_SVD_PROFILE=1: one line per call on stderr with the sector shapes and
  // the wall time of each phase (the stream is drained at every phase end)
  const char *prf = std::getenv("QLG_SVD_PROFILE");
  const bool profile = prf && prf[0] == '1';
  std::vector<std::pair<const char *, double>> phase;
  auto t_last = std::chrono::steady_clock::now();
  auto mark = [&](const char *name) {
    if (!profile) return;
    QLG_HIP(hipStreamSynchronize(stream));
    const auto now = std::chrono::steady_clock::now();
    phase.push_back({name, std::chrono::duration<double, std::milli>(now - t_last).count()});
    t_last = now;
  };
  QN div{0, 0};
  for (int k = 0; k < t.qn_len; ++k) div[k] = ldiv[k];

  // ---- sectors: the new bond charge of every block, row / column keys -------------
  std::map<QN, int> sid;
  std::vector<Sector> sec;
  std::vector<int> bsec(t.blk.size(), -1);
  for (size_t b = 0; b < t.blk.size(); ++b) {
    const Coords &cd = t.blk[b];
    QN q = div;
    int64_t re = 1, ce = 1;
    Key lk(cd.begin(), cd.begin() + ldims), rk(cd.begin() + ldims, cd.begin() + t.rank);
    for (int k = 0; k < t.rank; ++k) {
      const int64_t dg = t.ax[k].deg[cd[k]];
      if (k < ldims) {
        const QN &sq = t.ax[k].qn[cd[k]];
        q = t.ax[k].dir > 0 ? qn_sub(q, sq) : qn_add(q, sq);
        re *= dg;
      } else {
        ce *= dg;
      }
    }
    auto it = sid.find(q);
    int s;
    if (it == sid.end()) {
      s = (int)sec.size();
      sid[q] = s;
      sec.push_back(Sector{});
      sec.back().q = q;
    } else {
      s = it->second;
    }
    sec[s].rows[lk] = {re, 0};
    sec[s].cols[rk] = {ce, 0};
    bsec[b] = s;
  }
  // ascending charge order for the output bond
  std::vector<int> order;
  for (auto &kv : sid) order.push_back(kv.second);
  for (Sector &s : sec) {
    for (auto &kv : s.rows) {
      kv.second.second = s.R;
      s.R += kv.second.first;
    }
    for (auto &kv : s.cols) {
      kv.second.second = s.C;
      s.C += kv.second.first;
    }
    s.n = std::min(s.R, s.C);
    s.m = std::max(s.R, s.C);
    s.tr = s.R > s.C;
  }
  // ---- buckets ------------------------------------------------------------------------
  std::map<int64_t, std::vector<int>> raw;
  for (int s : order) raw[bucket_bound(sec[s].n)].push_back(s);
  std::vector<Bucket> bk;
  int64_t xtot = 0, gtot = 0, rtot = 0, nsv = 0;
  for (auto &kv : raw) {
    Bucket B;
    int64_t nmax = 0, mmax = 0;
    for (int s : kv.second) {
      nmax = std::max(nmax, sec[s].n);
      mmax = std::max(mmax, sec[s].m);
    }
    B.b = kv.first == 1 ? 1 : up16(nmax);
    B.mp = std::max<int64_t>(16, up16(mmax));
    B.secs = kv.second;
    B.xbase = xtot;
    B.gbase = gtot;
    B.rbase = rtot;
    const int64_t nb = (int64_t)B.secs.size();
    xtot += nb * B.b * B.mp;
    gtot += nb * B.b * B.b;
    rtot += nb * B.b * B.mp;
    for (size_t j = 0; j < B.secs.size(); ++j) {
      Sector &s = sec[B.secs[j]];
      s.bucket = (int)bk.size();
      s.slot = (int)j;
      s.sv0 = nsv;
      nsv += s.n;
    }
    bk.push_back(B);
  }
  QLG_CHECK((xtot + 1) * ed * 8 < (int64_t(1) << 40), "block SVD: tensor too large");

  // ---- gather, normalise, Gram ------------------------------------------------------------
  DevBuf X, G, R, scale, sv, dtab, dmats, dmats2, du;
  X.ensure((size_t)std::max<int64_t>(xtot, 1) * ed * 8);
  G.ensure((size_t)std::max<int64_t>(gtot, 1) * ed * 8);
  R.ensure((size_t)std::max<int64_t>(rtot, 1) * ed * 8);
  QLG_HIP(hipMemsetAsync(X.p, 0, (size_t)std::max<int64_t>(xtot, 1) * ed * 8, stream));
  QLG_HIP(hipMemsetAsync(G.p, 0, (size_t)std::max<int64_t>(gtot, 1) * ed * 8, stream));
  {
    std::vector<SvdBlk> tab;
    for (size_t b = 0; b < t.blk.size(); ++b) {
      const Sector &s = sec[bsec[b]];
      const Bucket &B = bk[s.bucket];
      const Coords &cd = t.blk[b];
      Key lk(cd.begin(), cd.begin() + ldims), rk(cd.begin() + ldims, cd.begin() + t.rank);
      const auto &rr = s.rows.at(lk);
      const auto &cc = s.cols.at(rk);
      if (rr.first == 0 || cc.first == 0) continue;
      tab.push_back(SvdBlk{t.off[b], rr.first, cc.first, B.xbase + s.slot * B.b * B.mp, rr.second, cc.second, B.mp,
                           s.tr ? 1 : 0, 0});
    }
    upload(c, dtab, tab);
    int64_t mx = 0;
    for (const SvdBlk &b : tab) mx = std::max(mx, b.rows * b.cols);
    svd_gather(stream, dtab.as<SvdBlk>(), (int64_t)tab.size(), mx, static_cast<const double *>(t.data),
               X.as<double>(), ed);
  }
  std::vector<SvdMat> xm;   // per sector (bucket order): its X [n x m]
  for (const Bucket &B : bk)
    for (size_t j = 0; j < B.secs.size(); ++j) {
      const Sector &s = sec[B.secs[j]];
      xm.push_back(SvdMat{B.xbase + (int64_t)j * B.b * B.mp, s.n, s.m, B.mp, 0});
    }
  int64_t xmax = 0, rowmax = 0;   // largest sector matrix, most rows
  for (const SvdMat &m : xm) {
    xmax = std::max(xmax, m.rows * m.cols);
    rowmax = std::max(rowmax, m.rows);
  }
  scale.ensure(std::max<size_t>(1, xm.size()) * 8);
  upload(c, dmats, xm);
  {
    DevBuf part;
    part.ensure(std::max<size_t>(1, xm.size() * (size_t)svd_grid_y(xmax)) * 8);
    svd_normalise(stream, dmats.as<SvdMat>(), (int64_t)xm.size(), xmax, X.as<double>(), part.as<double>(),
                  scale.as<double>(), ed);
  }
  {
    GemmBatch gb(c, cplx);
    for (const Bucket &B : bk)
      for (size_t j = 0; j < B.secs.size(); ++j) {
        const Sector &s = sec[B.secs[j]];
        const double *x = X.as<double>() + (B.xbase + (int64_t)j * B.b * B.mp) * ed;
        double *g = G.as<double>() + (B.gbase + (int64_t)j * B.b * B.b) * ed;
        gb.add(x, B.mp, false, false, x, B.mp, true, cplx, g, B.b, s.n, s.n, s.m);   // X X^H
      }
    gb.run();
    std::vector<SvdMat> pm;
    for (const Bucket &B : bk)
      for (size_t j = 0; j < B.secs.size(); ++j)
        if (B.b > 1) pm.push_back(SvdMat{B.gbase + (int64_t)j * B.b * B.b, B.b, sec[B.secs[j]].n, B.b, 0});
    upload(c, dmats2, pm);
    svd_pad_diag(stream, dmats2.as<SvdMat>(), (int64_t)pm.size(), G.as<double>(), ed);
  }
  mark("gram");

  // ---- eigen-decompositions (rows of E^H = eigenvectors^H, ascending) ------------------
  std::vector<char> forced(sec.size(), 0);   // sectors left to the Jacobi refinement (E = I)
  DevBuf W;
  W.ensure((size_t)std::max<int64_t>(gtot, 1) * 8);   // eigenvalues, b per matrix (an upper bound)
  {
    std::vector<int64_t> wbase(bk.size());
    int64_t wt = 0;
    for (size_t i = 0; i < bk.size(); ++i) {
      wbase[i] = wt;
      wt += (int64_t)bk[i].secs.size() * bk[i].b;
    }
    // complex: rotate by H2 H1, solve, rotate the vectors back; a batch with NaN
    // vectors is solved again with other reflectors; a matrix that still has
    // NaN vectors (a Gram matrix the rotation leaves diagonal: c I, or zero)
    // gets E = I and its sector is finished by the Jacobi refinement below
    std::vector<int> todo;
    for (size_t i = 0; i < bk.size(); ++i)
      if (bk[i].b > 1) todo.push_back((int)i);
    std::vector<int32_t> nan_flag;   // the last check: one flag per matrix of nan_bucket, in order
    std::vector<int> nan_bucket;
    DevBuf G0;
    if (cplx && !todo.empty()) {
      G0.ensure((size_t)std::max<int64_t>(gtot, 1) * ed * 8);
      QLG_HIP(hipMemcpyAsync(G0.p, G.p, (size_t)gtot * ed * 8, hipMemcpyDeviceToDevice, stream));
    }
    for (int attempt = 0; attempt < (cplx ? 2 : 1) && !todo.empty(); ++attempt) {
      std::vector<SvdMat> gm;
      std::vector<double> uh;
      int64_t bmax = 0;
      if (cplx) {
        for (int i : todo) {
          const Bucket &B = bk[i];
          bmax = std::max(bmax, B.b);
          if (attempt > 0)
            QLG_HIP(hipMemcpyAsync(G.as<double>() + B.gbase * ed, G0.as<double>() + B.gbase * ed,
                                   (size_t)B.secs.size() * B.b * B.b * ed * 8, hipMemcpyDeviceToDevice, stream));
          for (int which = 0; which < 2; ++which) {
            const auto u = reflector(B.b, which, attempt);
            const int64_t off = (int64_t)uh.size() / 2;
            uh.insert(uh.end(), u.begin(), u.end());
            for (size_t j = 0; j < B.secs.size(); ++j)
              gm.push_back(SvdMat{B.gbase + (int64_t)j * B.b * B.b, B.b, B.b, B.b, off});
          }
        }
        upload(c, du, uh);
        // gm holds, per bucket, all matrices with reflector 1, then with reflector 2
        upload(c, dmats2, gm);
        size_t pos = 0;
        for (int i : todo) {
          const int64_t nb = (int64_t)bk[i].secs.size();
          svd_reflect(stream, dmats2.as<SvdMat>() + pos, nb, bmax, G.as<double>(), du.as<double>());
          svd_reflect(stream, dmats2.as<SvdMat>() + pos + nb, nb, bmax, G.as<double>(), du.as<double>());
          pos += 2 * nb;
        }
      }
      std::vector<qlg_eigh_group> groups;
      for (int i : todo) {
        const Bucket &B = bk[i];
        groups.push_back(qlg_eigh_group{B.b, (int64_t)B.secs.size(), G.as<double>() + B.gbase * ed,
                                        W.as<double>() + wbase[i]});
      }
      eigh_batched(c, t.dtype, (int64_t)groups.size(), groups.data());
      if (!cplx) break;
      // undo the rotation on the rows of E^H: r <- r H2, then r <- r H1
      size_t pos = 0;
      for (int i : todo) {
        const int64_t nb = (int64_t)bk[i].secs.size();
        svd_reflect_rows(stream, dmats2.as<SvdMat>() + pos + nb, nb, G.as<double>(), du.as<double>());
        svd_reflect_rows(stream, dmats2.as<SvdMat>() + pos, nb, G.as<double>(), du.as<double>());
        pos += 2 * nb;
      }
      // NaN check on the device, one flag per matrix
      std::vector<SvdMat> nm;
      int64_t nmax = 0;
      for (int i : todo) {
        const Bucket &B = bk[i];
        for (size_t j = 0; j < B.secs.size(); ++j)
          nm.push_back(SvdMat{B.gbase + (int64_t)j * B.b * B.b, B.b, B.b, B.b, (int64_t)nm.size()});
        nmax = std::max(nmax, B.b * B.b * ed);
      }
      DevBuf dnm, dnf;
      upload(c, dnm, nm);
      dnf.ensure(std::max<size_t>(1, nm.size()) * sizeof(int32_t));
      QLG_HIP(hipMemsetAsync(dnf.p, 0, std::max<size_t>(1, nm.size()) * sizeof(int32_t), stream));
      svd_nan_flags(stream, dnm.as<SvdMat>(), (int64_t)nm.size(), nmax, G.as<double>(), dnf.as<int32_t>(), ed);
      download<int32_t>(c, nan_flag, dnf.p, nm.size());
      nan_bucket.clear();
      std::vector<int> left;
      size_t fp = 0;
      for (int i : todo) {
        bool bad = false;
        for (size_t j = 0; j < bk[i].secs.size(); ++j, ++fp) bad |= nan_flag[fp] != 0;
        if (bad) left.push_back(i);
        nan_bucket.push_back(i);
      }
      todo.swap(left);
    }
    // matrices still with NaN vectors: E^H row (b - n + i) = e_i, so R = X
    {
      size_t pos = 0;
      for (int i : nan_bucket) {
        const Bucket &B = bk[i];
        for (size_t j = 0; j < B.secs.size(); ++j, ++pos) {
          if (std::find(todo.begin(), todo.end(), i) == todo.end() || nan_flag[pos] == 0) continue;
          const Sector &sj = sec[B.secs[j]];
          std::vector<double> m((size_t)B.b * B.b * ed, 0.0);
          for (int64_t r = 0; r < sj.n; ++r) m[((B.b - sj.n + r) * B.b + r) * ed] = 1.0;
          QLG_HIP(hipMemcpyAsync(G.as<double>() + (B.gbase + (int64_t)j * B.b * B.b) * ed, m.data(), m.size() * 8,
                                 hipMemcpyHostToDevice, stream));
          QLG_HIP(hipStreamSynchronize(stream));
          forced[B.secs[j]] = 1;
        }
      }
    }
    for (Bucket &B : bk)   // the b = 1 bucket: E^H = [1]
      if (B.b == 1) {
        std::vector<double> one((size_t)B.secs.size() * ed, 0.0);
        for (size_t j = 0; j < B.secs.size(); ++j) one[j * ed] = 1.0;
        QLG_HIP(hipMemcpyAsync(G.as<double>() + B.gbase * ed, one.data(), one.size() * 8, hipMemcpyHostToDevice,
                               stream));
        QLG_HIP(hipStreamSynchronize(stream));
      }
  }

  mark("eigh");
  // ---- back-projection R = E^H X (valid eigenvectors: the last n rows) ------------------
  {
    GemmBatch gb(c, cplx);
    for (const Bucket &B : bk)
      for (size_t j = 0; j < B.secs.size(); ++j) {
        const Sector &s = sec[B.secs[j]];
        const double *e = G.as<double>() + (B.gbase + (int64_t)j * B.b * B.b + (B.b - s.n) * B.b) * ed;
        const double *x = X.as<double>() + (B.xbase + (int64_t)j * B.b * B.mp) * ed;
        double *r = R.as<double>() + (B.rbase + (int64_t)j * B.b * B.mp) * ed;
        gb.add(e, B.b, false, false, x, B.mp, false, false, r, B.mp, s.n, s.m, s.n);
      }
    gb.run();
  }
  std::vector<SvdMat> rm;   // per sector: its R rows [n x m], aux = first singular value
  for (const Bucket &B : bk)
    for (size_t j = 0; j < B.secs.size(); ++j) {
      const Sector &s = sec[B.secs[j]];
      rm.push_back(SvdMat{B.rbase + (int64_t)j * B.b * B.mp, s.n, s.m, B.mp, s.sv0});
    }
  sv.ensure((size_t)std::max<int64_t>(nsv, 1) * 8);
  upload(c, dmats, rm);
  svd_row_norms(stream, dmats.as<SvdMat>(), (int64_t)rm.size(), rowmax, R.as<double>(), sv.as<double>(), ed);
  std::vector<double> hsv, hscale;
  download<double>(c, hscale, scale.p, xm.size());
  download<double>(c, hsv, sv.p, (size_t)nsv);

  // ---- cluster refinement: sectors whose kept values reach below 1e-6 of their largest --
  // (decided on the first values; the refined values then decide the truncation)
  auto truncate = [&](const std::vector<double> &vals, std::vector<int64_t> &keep, double &err, int64_t &D) {
    // vals: per sector scaled values (tensor scale), sector order of `sec`
    std::vector<std::pair<double, int>> all;
    for (size_t si = 0; si < sec.size(); ++si)
      for (int64_t i = 0; i < sec[si].n; ++i) all.push_back({vals[sec[si].sv0 + i], (int)si});
    std::stable_sort(all.begin(), all.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    const int64_t N = (int64_t)all.size();
    std::vector<double> tail(N + 1, 0.0);
    for (int64_t i = N - 1; i >= 0; --i) tail[i] = tail[i + 1] + all[i].first * all[i].first;
    const double total = tail[0];
    int64_t d = N;
    for (int64_t k = 0; k <= N; ++k)
      if (total <= 0.0 || tail[k] / total <= trunc_err) {
        d = k;
        break;
      }
    d = std::max<int64_t>(std::min<int64_t>(std::max<int64_t>(d, std::min<int64_t>(Dmin, N)), std::min(Dmax, N)), 1);
    D = d;
    err = total > 0.0 ? tail[d] / total : 0.0;
    keep.assign(sec.size(), 0);
    for (int64_t i = 0; i < d; ++i) ++keep[all[i].second];
  };
  auto scaled_values = [&]() {
    std::vector<double> v(nsv, 0.0);
    int mi = 0;
    for (const Bucket &B : bk)
      for (size_t j = 0; j < B.secs.size(); ++j, ++mi) {
        const Sector &s = sec[B.secs[j]];
        for (int64_t i = 0; i < s.n; ++i) v[s.sv0 + i] = hsv[s.sv0 + i] * hscale[mi];
      }
    return v;
  };
  std::vector<int64_t> keep;
  double err = 0.0;
  int64_t D = 0;
  truncate(scaled_values(), keep, err, D);
  mark("project");
  int jac_sweeps = 0;
  {
    // Sectors whose kept values reach below 1e-6 of their largest: the rows of
    // such small values carry components of their neighbours' directions
    // (eps s_max^2 / (s_i s_j) relative) that Newton-Schulz cannot untangle.
    // One-sided Jacobi on the sector's rows of R, with the same rotations on
    // its rows of E^H (X = E R is kept exactly), resolves them; the refined
    // values then decide the truncation.  Refined values move the pooled cut,
    // which can make another sector keep values below the threshold: the check
    // repeats until every such sector is refined (each at most once).
    constexpr double kTau = 1e-6;
    std::vector<char> refined(sec.size(), 0);
    for (int pass = 0; pass < 8; ++pass) {
      // sectors to refine in this pass
      std::vector<int> active;
      for (size_t si = 0; si < sec.size(); ++si) {
        const Sector &s = sec[si];
        if (refined[si] || s.n < 2) continue;
        if (keep[si] == 0 && !forced[si]) continue;
        double vmax = 0.0, vkeep = 1e300;
        std::vector<double> v(hsv.begin() + s.sv0, hsv.begin() + s.sv0 + s.n);
        std::vector<double> sorted = v;
        std::sort(sorted.begin(), sorted.end(), std::greater<double>());
        vmax = sorted[0];
        if (keep[si] > 0) vkeep = sorted[keep[si] - 1];
        if (!forced[si] && vkeep >= kTau * vmax) continue;
        refined[si] = 1;
        active.push_back((int)si);
      }
      if (active.empty()) break;
      // Multi-level Gram (real data).  The Gram eigenvectors resolve values
      // down to ~1e-8 of a sector's largest; the rows below that carry their
      // neighbours' directions.  In each refined sector, the rows below
      // 1e-4 (then 1e-8) of its largest value and above its rounding floor
      // are scaled to unit largest norm, and rotated by the eigenvectors of
      // their own Gram matrix (the same rotation on their rows of E^H keeps
      // X = E R exact).  Each level resolves another ~8 decades, so the
      // Jacobi sweeps below start from nearly orthogonal rows.
      if (!cplx) {
        constexpr double kLevelTau = 1e-4;
        for (int level = 0; level < 2; ++level) {
          struct Sel {
            int si;
            int64_t k, kb, perm, rs, es, gs, ws;
          };
          std::vector<Sel> sels;
          std::vector<int32_t> perm;
          std::vector<double> fdown, fup;
          int64_t rs_tot = 0, es_tot = 0;
          const double thr_rel = level == 0 ? kLevelTau : kLevelTau * kLevelTau;
          for (int si : active) {
            const Sector &s = sec[si];
            const Bucket &B = bk[s.bucket];
            double vmax = 0.0;
            for (int64_t r = 0; r < s.n; ++r) vmax = std::max(vmax, hsv[s.sv0 + r]);
            const double floor = 64.0 * 2.220446049250313e-16 * vmax, thr = thr_rel * vmax;
            std::vector<int32_t> rows;
            double smax = 0.0;
            for (int64_t r = 0; r < s.n; ++r) {
              const double v = hsv[s.sv0 + r];
              if (v > floor && v < thr) {
                rows.push_back((int32_t)r);
                smax = std::max(smax, v);
              }
            }
            if (rows.size() < 2 || !(smax > 0.0)) continue;
            Sel q{};
            q.si = si;
            q.k = (int64_t)rows.size();
            q.kb = up16(q.k);
            q.perm = (int64_t)perm.size();
            q.rs = rs_tot;
            rs_tot += q.k * s.m;
            q.es = es_tot;
            es_tot += q.k * B.b;
            perm.insert(perm.end(), rows.begin(), rows.end());
            fdown.insert(fdown.end(), rows.size(), 1.0 / smax);
            fup.insert(fup.end(), rows.size(), smax);
            sels.push_back(q);
          }
          if (sels.empty()) break;
          // Gram matrices grouped by padded order (eigh batches are contiguous)
          std::map<int64_t, std::vector<size_t>> byk;
          for (size_t i = 0; i < sels.size(); ++i) byk[sels[i].kb].push_back(i);
          int64_t gs_tot = 0, ws_tot = 0;
          for (auto &kv : byk)
            for (size_t i : kv.second) {
              sels[i].gs = gs_tot;
              gs_tot += kv.first * kv.first;
              sels[i].ws = ws_tot;
              ws_tot += kv.first;
            }
          DevBuf Rs, Rs2, Es, Es2, Gs, Ws, dperm, dfd, dfu, dgr, dge, dpm;
          Rs.ensure((size_t)rs_tot * 8);
          Rs2.ensure((size_t)rs_tot * 8);
          Es.ensure((size_t)es_tot * 8);
          Es2.ensure((size_t)es_tot * 8);
          Gs.ensure((size_t)gs_tot * 8);
          Ws.ensure((size_t)ws_tot * 8);
          upload(c, dperm, perm);
          upload(c, dfd, fdown);
          upload(c, dfu, fup);
          std::vector<SvdRows> gr, ge, sr, se;
          std::vector<SvdMat> pm;
          int64_t rmax = 0, emax = 0;
          for (const Sel &q : sels) {
            const Sector &s = sec[q.si];
            const Bucket &B = bk[s.bucket];
            const int64_t r0 = B.rbase + (int64_t)s.slot * B.b * B.mp;
            const int64_t e0 = B.gbase + (int64_t)s.slot * B.b * B.b + (B.b - s.n) * B.b;
            gr.push_back(SvdRows{r0, B.mp, q.rs, s.m, q.k, s.m, q.perm});
            ge.push_back(SvdRows{e0, B.b, q.es, B.b, q.k, B.b, q.perm});
            sr.push_back(SvdRows{q.rs, s.m, r0, B.mp, q.k, s.m, q.perm});
            se.push_back(SvdRows{q.es, B.b, e0, B.b, q.k, B.b, q.perm});
            pm.push_back(SvdMat{q.gs, q.kb, q.k, q.kb, 0});
            rmax = std::max(rmax, q.k * s.m);
            emax = std::max(emax, q.k * B.b);
          }
          upload(c, dgr, gr);
          upload(c, dge, ge);
          svd_gather_rows(stream, dgr.as<SvdRows>(), (int64_t)gr.size(), rmax, dperm.as<int32_t>(), dfd.as<double>(),
                          R.as<double>(), Rs.as<double>(), 1);
          svd_gather_rows(stream, dge.as<SvdRows>(), (int64_t)ge.size(), emax, dperm.as<int32_t>(), nullptr,
                          G.as<double>(), Es.as<double>(), 1);
          QLG_HIP(hipMemsetAsync(Gs.p, 0, (size_t)gs_tot * 8, stream));
          {
            GemmBatch gb(c, false);
            for (const Sel &q : sels) {
              const Sector &s = sec[q.si];
              gb.add(Rs.as<double>() + q.rs, s.m, false, false, Rs.as<double>() + q.rs, s.m, true, false,
                     Gs.as<double>() + q.gs, q.kb, q.k, q.k, s.m);
            }
            gb.run();
          }
          upload(c, dpm, pm);
          svd_pad_diag(stream, dpm.as<SvdMat>(), (int64_t)pm.size(), Gs.as<double>(), 1);
          std::vector<qlg_eigh_group> groups;
          for (auto &kv : byk) {
            const Sel &q0 = sels[kv.second.front()];
            groups.push_back(qlg_eigh_group{kv.first, (int64_t)kv.second.size(), Gs.as<double>() + q0.gs,
                                            Ws.as<double>() + q0.ws});
          }
          eigh_batched(c, QLG_F64, (int64_t)groups.size(), groups.data());
          {
            // rows of the eigenvector buffer kb - k .. kb - 1: the valid (ascending) eigenvectors
            GemmBatch gb(c, false);
            for (const Sel &q : sels) {
              const Sector &s = sec[q.si];
              const Bucket &B = bk[s.bucket];
              const double *V = Gs.as<double>() + q.gs + (q.kb - q.k) * q.kb;
              gb.add(V, q.kb, false, false, Rs.as<double>() + q.rs, s.m, false, false, Rs2.as<double>() + q.rs, s.m,
                     q.k, s.m, q.k);
              gb.add(V, q.kb, false, false, Es.as<double>() + q.es, B.b, false, false, Es2.as<double>() + q.es, B.b,
                     q.k, B.b, q.k);
            }
            gb.run();
          }
          upload(c, dgr, sr);
          upload(c, dge, se);
          svd_scatter_rows(stream, dgr.as<SvdRows>(), (int64_t)sr.size(), rmax, dperm.as<int32_t>(), dfu.as<double>(),
                           Rs2.as<double>(), R.as<double>(), 1);
          svd_scatter_rows(stream, dge.as<SvdRows>(), (int64_t)se.size(), emax, dperm.as<int32_t>(), nullptr,
                           Es2.as<double>(), G.as<double>(), 1);
          svd_row_norms(stream, dmats.as<SvdMat>(), (int64_t)rm.size(), rowmax, R.as<double>(), sv.as<double>(), ed);
          download<double>(c, hsv, sv.p, (size_t)nsv);   // waits for the rotations before the buffers go
        }
      }
      // Sweeps.  Before each one a sector's rows are re-sorted by their
      // current norms (de Rijk's pivoting: rows of like size meet in the same
      // blocks, which cuts the sweep count on graded spectra), rows at the
      // sector's rounding floor (64 eps of its largest) are left out, and a
      // sector whose last sweep rotated nothing is finished.
      DevBuf dj, dpairs, didx, dflag;
      std::vector<int32_t> hflag;
      for (int sweep = 0; sweep < 30 && !active.empty(); ++sweep) {
        std::vector<SvdJacB> jobs;
        std::vector<int32_t> idx, pairs;
        std::vector<int> job_sec;
        int maxnbp = 0;
        for (int si : active) {
          const Sector &s = sec[si];
          std::vector<int32_t> ord(s.n);
          std::iota(ord.begin(), ord.end(), 0);
          std::stable_sort(ord.begin(), ord.end(),
                           [&](int32_t a, int32_t b) { return hsv[s.sv0 + a] > hsv[s.sv0 + b]; });
          const double floor = 64.0 * 2.220446049250313e-16 * hsv[s.sv0 + ord[0]];
          int64_t k = 0;
          while (k < s.n && hsv[s.sv0 + ord[k]] > floor) ++k;
          if (k < 2) continue;
          const Bucket &B = bk[s.bucket];
          const int nblk = (int)((k + 15) / 16), nbp = nblk + (nblk & 1);
          const int job = (int)jobs.size();
          jobs.push_back(SvdJacB{B.rbase + (int64_t)s.slot * B.b * B.mp, B.mp, s.m,
                                 B.gbase + (int64_t)s.slot * B.b * B.b + (B.b - s.n) * B.b, B.b, k,
                                 (int64_t)idx.size(), nblk, 0});
          job_sec.push_back(si);
          idx.insert(idx.end(), ord.begin(), ord.begin() + k);
          for (int pi = 0; pi < nbp / 2; ++pi) {
            pairs.push_back(job);
            pairs.push_back(pi);
          }
          maxnbp = std::max(maxnbp, nbp);
        }
        if (jobs.empty()) break;
        upload(c, dj, jobs);
        upload(c, dpairs, pairs);
        upload(c, didx, idx);
        dflag.ensure(jobs.size() * sizeof(int32_t));
        QLG_HIP(hipMemsetAsync(dflag.p, 0, jobs.size() * sizeof(int32_t), stream));
        for (int round = 0; round < maxnbp - 1; ++round)
          svd_jacobi_round(stream, dj.as<SvdJacB>(), dpairs.as<int32_t>(), (int64_t)pairs.size() / 2,
                           didx.as<int32_t>(), round, R.as<double>(), G.as<double>(), dflag.as<int32_t>(), ed);
        download<int32_t>(c, hflag, dflag.p, jobs.size());
        std::vector<int> still;
        for (size_t q = 0; q < jobs.size(); ++q)
          if (hflag[q] != 0) still.push_back(job_sec[q]);
        if (debug) std::fprintf(stderr, "[qlg svd] Jacobi pass %d sweep %d: %zu jobs, %zu still rotating\n", pass,
                                sweep, jobs.size(), still.size());
        ++jac_sweeps;
        active.swap(still);
        if (!active.empty()) {   // the next sweep's ordering
          svd_row_norms(stream, dmats.as<SvdMat>(), (int64_t)rm.size(), rowmax, R.as<double>(), sv.as<double>(), ed);
          download<double>(c, hsv, sv.p, (size_t)nsv);
        }
      }
      svd_row_norms(stream, dmats.as<SvdMat>(), (int64_t)rm.size(), rowmax, R.as<double>(), sv.as<double>(), ed);
      download<double>(c, hsv, sv.p, (size_t)nsv);
      truncate(scaled_values(), keep, err, D);
    }
  }
  if (debug)
    for (size_t si = 0; si < sec.size(); ++si) {
      const Sector &s = sec[si];
      std::fprintf(stderr, "[qlg svd] sector q=(%d,%d) R=%lld C=%lld tr=%d bucket b=%lld mp=%lld keep=%lld values:", s.q[0],
                   s.q[1], (long long)s.R, (long long)s.C, (int)s.tr, (long long)bk[s.bucket].b,
                   (long long)bk[s.bucket].mp, (long long)keep[si]);
      for (int64_t i = 0; i < s.n; ++i) std::fprintf(stderr, " %.3e", hsv[s.sv0 + i]);
      std::fprintf(stderr, "\n");
    }
  mark("refine");
  // ---- kept rows: P = R_j / s_j in descending order, U side from E^H ----------------------
  // per kept sector: perm (descending order of its rows), factors 1 / s
  std::vector<int> kept;
  for (int s : order)
    if (keep[s] > 0) kept.push_back(s);
  std::vector<int32_t> perm;
  std::vector<double> inv, mask;
  std::vector<int64_t> pbase(sec.size(), 0), kbase(sec.size(), 0);
  std::vector<SvdRows> pjobs, ejobs;
  int64_t ptot = 0, etot = 0;
  const std::vector<double> svals = scaled_values();
  std::vector<double> kept_sv;
  for (int si : kept) {
    const Sector &s = sec[si];
    const Bucket &B = bk[s.bucket];
    std::vector<int32_t> ord(s.n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int32_t a, int32_t b) { return hsv[s.sv0 + a] > hsv[s.sv0 + b]; });
    const int64_t k = keep[si];
    kbase[si] = (int64_t)perm.size();
    // values at the rounding floor of the sector (64 eps of its largest) are
    // zero within FP64: their rows of R carry rounding errors of the large
    // values, not a direction, so they are reported as 0 and their vectors
    // completed to an orthonormal set
    const double floor = 64.0 * 2.220446049250313e-16 * hsv[s.sv0 + ord[0]];
    for (int64_t i = 0; i < k; ++i) {
      const int32_t r = ord[i];
      perm.push_back(r);
      const double v = hsv[s.sv0 + r];
      const bool live = v > floor && v > 0.0;
      inv.push_back(live ? 1.0 / v : 0.0);
      mask.push_back(live ? 1.0 : 0.0);
      kept_sv.push_back(live ? svals[s.sv0 + r] : 0.0);
    }
    pbase[si] = ptot;
    pjobs.push_back(SvdRows{B.rbase + (int64_t)s.slot * B.b * B.mp, B.mp, ptot, s.m, k, s.m, kbase[si]});
    ptot += k * s.m;
    // rows of E^H for the kept values (row index of E^H: b - n + r)
    ejobs.push_back(SvdRows{B.gbase + (int64_t)s.slot * B.b * B.b + (B.b - s.n) * B.b, B.b, etot, s.n, k, s.n,
                            kbase[si]});
    etot += k * s.n;
  }
  DevBuf P, EK, dperm, dinv, dmask, dpj, dej;
  P.ensure((size_t)std::max<int64_t>(ptot, 1) * ed * 8);
  EK.ensure((size_t)std::max<int64_t>(etot, 1) * ed * 8);
  upload(c, dperm, perm);
  upload(c, dinv, inv);
  upload(c, dmask, mask);
  upload(c, dpj, pjobs);
  upload(c, dej, ejobs);
  int64_t pmax = 0, emax = 0;
  for (const SvdRows &j : pjobs) pmax = std::max(pmax, j.k * j.cols);
  for (const SvdRows &j : ejobs) emax = std::max(emax, j.k * j.cols);
  svd_gather_rows(stream, dpj.as<SvdRows>(), (int64_t)pjobs.size(), pmax, dperm.as<int32_t>(), dinv.as<double>(),
                  R.as<double>(), P.as<double>(), ed);
  svd_gather_rows(stream, dej.as<SvdRows>(), (int64_t)ejobs.size(), emax, dperm.as<int32_t>(), nullptr,
                  G.as<double>(), EK.as<double>(), ed);

  // ---- Newton-Schulz on the kept rows of P --------------------------------------------------
  int ns_rounds = 0;
  {
    std::vector<int64_t> abase(kept.size());
    int64_t atot = 0;
    for (size_t i = 0; i < kept.size(); ++i) {
      abase[i] = atot;
      atot += keep[kept[i]] * keep[kept[i]];
    }
    DevBuf A, T, dev, dam;
    A.ensure((size_t)std::max<int64_t>(atot, 1) * ed * 8);
    T.ensure((size_t)std::max<int64_t>(ptot, 1) * ed * 8);
    std::vector<int> active(kept.size());
    std::iota(active.begin(), active.end(), 0);
    for (int round = 0; round < 8 && !active.empty(); ++round) {
      ns_rounds = round + 1;
      GemmBatch g1(c, cplx);
      std::vector<SvdMat> am;
      for (int i : active) {
        const Sector &s = sec[kept[i]];
        const int64_t k = keep[kept[i]];
        const double *p = P.as<double>() + pbase[kept[i]] * ed;
        g1.add(p, s.m, false, false, p, s.m, true, cplx, A.as<double>() + abase[i] * ed, k, k, k, s.m);
        am.push_back(SvdMat{abase[i], k, k, k, kbase[kept[i]]});
      }
      g1.run();
      upload(c, dam, am);
      int64_t amax = 0;
      for (const SvdMat &m : am) amax = std::max(amax, m.rows * m.rows);
      const int gy = svd_grid_y(amax);
      dev.ensure(std::max<size_t>(1, am.size() * (size_t)gy) * 8);
      svd_ns_dev(stream, dam.as<SvdMat>(), (int64_t)am.size(), amax, A.as<double>(), dmask.as<double>(),
                 dev.as<double>(), ed);
      std::vector<double> hdp, hd(am.size(), 0.0);
      download<double>(c, hdp, dev.p, am.size() * (size_t)gy);
      for (size_t q = 0; q < am.size(); ++q)
        for (int y = 0; y < gy; ++y) hd[q] = std::max(hd[q], hdp[q * gy + y]);
      if (debug)
        for (size_t q = 0; q < active.size(); ++q)
          std::fprintf(stderr, "[qlg svd] NS round %d sector q=(%d,%d) k=%lld dev %.3e\n", round,
                       sec[kept[active[q]]].q[0], sec[kept[active[q]]].q[1], (long long)keep[kept[active[q]]], hd[q]);
      std::vector<int> still;
      GemmBatch g2(c, cplx);
      std::vector<SvdMat> um;
      for (size_t q = 0; q < active.size(); ++q) {
        if (hd[q] < 1e-13) continue;
        if (!(hd[q] <= 0.5)) {
          const Sector &bad = sec[kept[active[q]]];
          double vmin = 1e300, vmax = 0.0;
          for (int64_t i = 0; i < bad.n; ++i) vmax = std::max(vmax, hsv[bad.sv0 + i]);
          for (int64_t i = 0; i < keep[kept[active[q]]]; ++i)
            vmin = std::min(vmin, hsv[bad.sv0 + perm[kbase[kept[active[q]]] + i]]);
          char msg[320];
          std::snprintf(msg, sizeof msg,
                        "block SVD: singular vectors too far from orthonormal (sector (%d,%d), %lld x %lld, "
                        "%lld kept, smallest kept / largest value %.3e, deviation %.3e after %d steps)",
                        bad.q[0], bad.q[1], (long long)bad.R, (long long)bad.C,
                        (long long)keep[kept[active[q]]], vmin / vmax, hd[q], round);
          throw Error(QLG_ERR_NUMERIC, msg);
        }
        const int i = active[q];
        const Sector &s = sec[kept[i]];
        const int64_t k = keep[kept[i]];
        const double *p = P.as<double>() + pbase[kept[i]] * ed;
        g2.add(A.as<double>() + abase[i] * ed, k, false, false, p, s.m, false, false, T.as<double>() + pbase[kept[i]] * ed,
               s.m, k, s.m, k);
        um.push_back(SvdMat{pbase[kept[i]], k, s.m, s.m, pbase[kept[i]]});
        still.push_back(i);
      }
      g2.run();
      upload(c, dam, um);
      int64_t umax = 0;
      for (const SvdMat &m : um) umax = std::max(umax, m.rows * m.cols * ed);
      svd_ns_update(stream, dam.as<SvdMat>(), (int64_t)um.size(), umax, P.as<double>(), T.as<double>(), ed);
      active.swap(still);
    }
    // rows of exactly zero singular values: completed to an orthonormal set
    std::vector<SvdMat> zm;
    for (int si : kept) {
      bool any = false;
      for (int64_t i = 0; i < keep[si]; ++i) any |= mask[kbase[si] + i] == 0.0;
      if (any) zm.push_back(SvdMat{pbase[si], keep[si], sec[si].m, sec[si].m, kbase[si]});
    }
    if (!zm.empty()) {
      upload(c, dam, zm);
      svd_complete_rows(stream, dam.as<SvdMat>(), (int64_t)zm.size(), P.as<double>(), dmask.as<double>(), ed);
    }
  }

  mark("orthonormalise");
  // ---- outputs -----------------------------------------------------------------------------------
  auto *res = new SvdResult();
  std::unique_ptr<SvdResult> guard(res);
  res->ctx = &c;
  res->qn_len = t.qn_len;
  res->dtype = t.dtype;
  res->trunc_err = err;
  res->D = D;
  std::vector<int> mid_of(sec.size(), -1);
  for (size_t mi = 0; mi < kept.size(); ++mi) {
    mid_of[kept[mi]] = (int)mi;
    res->bond_qn.push_back(sec[kept[mi]].q);
    res->bond_deg.push_back(keep[kept[mi]]);
  }
  // Sector matrix M (R x C), kept rows P (k x m) and kept rows of E^H (EK, k x n):
  //   R <= C (X = M):   U[r][c] = conj(EK[c][r]),  V^H[c][x] = P[c][x]
  //   R >  C (X = M^T): U[r][c] = P[c][r],         V^H[c][x] = conj(EK[c][x])
  // u blocks (lk..., mi) are row ranges of U, vt blocks (mi, rk...) column ranges of V^H.
  struct OutBlk {
    std::vector<int32_t> coords;
    SvdOut job;
    bool from_p;
  };
  std::vector<OutBlk> ub, vb;
  {
    int64_t ebase = 0;
    for (int si : kept) {
      const Sector &s = sec[si];
      const int64_t k = keep[si];
      const int32_t mi = mid_of[si];
      for (auto &kv : s.rows) {
        const int64_t ext = kv.second.first, ro = kv.second.second;
        OutBlk o;
        o.coords = kv.first;
        o.coords.push_back(mi);
        o.from_p = s.tr;
        o.job = s.tr ? SvdOut{0, ext, k, pbase[si] + ro, 1, s.m, 0, 0}
                     : SvdOut{0, ext, k, ebase + ro, 1, s.n, cplx ? 1 : 0, 0};
        ub.push_back(o);
      }
      for (auto &kv : s.cols) {
        const int64_t ext = kv.second.first, co = kv.second.second;
        OutBlk o;
        o.coords.push_back(mi);
        o.coords.insert(o.coords.end(), kv.first.begin(), kv.first.end());
        o.from_p = !s.tr;
        o.job = s.tr ? SvdOut{0, k, ext, ebase + co, s.n, 1, cplx ? 1 : 0, 0}
                     : SvdOut{0, k, ext, pbase[si] + co, s.m, 1, 0, 0};
        vb.push_back(o);
      }
      ebase += k * s.n;
    }
  }
  // canonical layout: blocks in lexicographic order of their coordinates, offsets = prefix sums
  auto finish = [&](std::vector<OutBlk> &blks, int which, std::vector<SvdOut> &fp, std::vector<SvdOut> &fe) {
    std::sort(blks.begin(), blks.end(), [](const OutBlk &a, const OutBlk &b) { return a.coords < b.coords; });
    int64_t off = 0;
    for (OutBlk &o : blks) {
      o.job.dst = off;
      for (int32_t v : o.coords) res->coords[which].push_back(v);
      res->offsets[which].push_back(off);
      off += o.job.rows * o.job.cols;
      (o.from_p ? fp : fe).push_back(o.job);
    }
    res->size[which] = off;
    res->data[which].ensure((size_t)std::max<int64_t>(off, 1) * ed * 8);
  };
  {
    std::vector<SvdOut> up, ue, vp, ve;
    finish(ub, 0, up, ue);
    finish(vb, 2, vp, ve);
    DevBuf d0, d1, d2, d3;
    upload(c, d0, up);
    upload(c, d1, ue);
    upload(c, d2, vp);
    upload(c, d3, ve);
    auto most = [](const std::vector<SvdOut> &v) {
      int64_t m = 0;
      for (const SvdOut &o : v) m = std::max(m, o.rows * o.cols);
      return m;
    };
    svd_assemble(stream, d0.as<SvdOut>(), (int64_t)up.size(), most(up), P.as<double>(), res->data[0].as<double>(), ed);
    svd_assemble(stream, d1.as<SvdOut>(), (int64_t)ue.size(), most(ue), EK.as<double>(), res->data[0].as<double>(),
                 ed);
    svd_assemble(stream, d2.as<SvdOut>(), (int64_t)vp.size(), most(vp), P.as<double>(), res->data[2].as<double>(), ed);
    svd_assemble(stream, d3.as<SvdOut>(), (int64_t)ve.size(), most(ve), EK.as<double>(), res->data[2].as<double>(),
                 ed);
    QLG_HIP(hipStreamSynchronize(stream));
  }
  // s: diagonal blocks (mi, mi), real
  {
    std::vector<SvdRows> dj;
    int64_t off = 0, kb = 0;
    for (size_t mi = 0; mi < kept.size(); ++mi) {
      const int64_t k = keep[kept[mi]];
      res->coords[1].push_back((int32_t)mi);
      res->coords[1].push_back((int32_t)mi);
      res->offsets[1].push_back(off);
      dj.push_back(SvdRows{kb, 0, off, 0, k, k, 0});
      off += k * k;
      kb += k;
    }
    res->size[1] = off;
    res->data[1].ensure((size_t)std::max<int64_t>(off, 1) * 8);
    DevBuf dsv, dd;
    upload(c, dsv, kept_sv);
    upload(c, dd, dj);
    int64_t kmax = 0;
    for (const SvdRows &j : dj) kmax = std::max(kmax, j.k * j.k);
    svd_diag(stream, dd.as<SvdRows>(), (int64_t)dj.size(), kmax, dsv.as<double>(), res->data[1].as<double>());
    QLG_HIP(hipStreamSynchronize(stream));
  }
  mark("assemble");
  if (profile) {
    int64_t nmax = 0, mmax = 0, nsum = 0;
    for (const Sector &s : sec) {
      nmax = std::max(nmax, s.n);
      mmax = std::max(mmax, s.m);
      nsum += s.n;
    }
    double tot = 0.0;
    std::string line;
    char buf[96];
    for (auto &ph : phase) {
      std::snprintf(buf, sizeof buf, " %s %.2f", ph.first, ph.second);
      line += buf;
      tot += ph.second;
    }
    std::string bl;
    for (const Bucket &B : bk) {
      std::snprintf(buf, sizeof buf, " %lldx%zu", (long long)B.b, B.secs.size());
      bl += buf;
    }
    std::fprintf(stderr,
                 "[qlg svd prof] sectors %zu n_max %lld m_max %lld n_sum %lld D %lld buckets%s | ms%s total %.2f "
                 "| jacobi sweeps %d ns rounds %d\n",
                 sec.size(), (long long)nmax, (long long)mmax, (long long)nsum, (long long)D, bl.c_str(), line.c_str(),
                 tot, jac_sweeps, ns_rounds);
  }
  return guard.release();
}

}  // namespace qlg
