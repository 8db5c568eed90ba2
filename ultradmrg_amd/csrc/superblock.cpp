// DMRG-form (super-block) effective Hamiltonian.
//
// The reference's DMRG executor applies
//     H psi = sum_j [ sum_i L_i (x) W_l(i, j) ] psi [ sum_k W_r(j, k) (x) R_k ]
// (include/qlmps/algorithm/dmrg/lanczos_dmrg_solver_impl.h:199-283, terms from
// include/qlmps/algorithm/dmrg/dmrg_impl.h:319-341) with rank-2 block operators
// L_i[a (OUT), a' (IN)], R_k[b (IN), b' (OUT)] and rank-2 site operators
// W[s_in (IN), s_out (OUT)].  The same operator in MPO form is
//     L[a', w, a] = L_w[a, a'],  R[b', z, b] = R_z[b, b'],
//     W_l[w, s, s', y] = W_l(w, y)[s, s'],  W_r[y, s, s', z] = W_r(y, z)[s, s'],
// with one bond element per distinct block operator (w, z) or term group (y).
// The bond elements' charges follow from the operators' divergences so that
// every MPO-form tensor has divergence zero:
//     q(w_i) = -div(L_i),  q(z_k) = div(R_k),
//     q(y_j) = q(w_i) - div(W_l(i, j)) = q(z_k) + div(W_r(j, k)).
// Bond elements of equal charge share one sector.  The stacked tensors are
// built on the device once and handed to the ordinary two-site planner.
#include <algorithm>
#include <map>
#include <memory>
#include <set>

#include "plan.h"

namespace qlg {

void launch_trans_copy_multi(bool cplx, const TransDesc *d, const double *const *srcs, int n, double *dst,
                             hipStream_t st);

namespace {

struct BondIndex {
  // element e -> (sector, position); sectors sorted by charge
  std::vector<QN> sect_qn;
  std::vector<int64_t> sect_deg;
  std::vector<int32_t> el_sect;
  std::vector<int64_t> el_pos;
  void build(const std::vector<QN> &el_qn) {
    std::map<QN, int32_t> ids;
    for (const QN &q : el_qn) ids.emplace(q, 0);
    int32_t s = 0;
    for (auto &kv : ids) {
      kv.second = s++;
      sect_qn.push_back(kv.first);
      sect_deg.push_back(0);
    }
    for (const QN &q : el_qn) {
      const int32_t sid = ids[q];
      el_sect.push_back(sid);
      el_pos.push_back(sect_deg[sid]++);
    }
  }
};

// a host-built tensor with arrays that a qlg_bsten descriptor can point at
struct Built {
  int rank = 0;
  std::vector<int32_t> nsect, qn, coords;
  std::vector<int64_t> deg, offs;
  std::vector<int8_t> dir;
  std::vector<Coords> blk;
  std::map<Coords, int64_t> index;   // block -> position in blk
  int64_t size = 0;
  qlg_bsten d{};
  void add_axis(const std::vector<QN> &q, const std::vector<int64_t> &g, int direction, int qn_len) {
    nsect.push_back((int32_t)q.size());
    for (size_t s = 0; s < q.size(); ++s) {
      for (int c = 0; c < qn_len; ++c) qn.push_back(q[s][c]);
      deg.push_back(g[s]);
    }
    dir.push_back((int8_t)direction);
    ++rank;
  }
  int64_t axis_deg(int axis, int s) const {
    int64_t base = 0;
    for (int a = 0; a < axis; ++a) base += nsect[a];
    return deg[base + s];
  }
  void finish(const std::set<Coords> &blocks, int qn_len, int dtype, const void *data) {
    blk.assign(blocks.begin(), blocks.end());
    int64_t o = 0;
    for (size_t b = 0; b < blk.size(); ++b) {
      index[blk[b]] = (int64_t)b;
      offs.push_back(o);
      int64_t n = 1;
      for (int a = 0; a < rank; ++a) n *= axis_deg(a, blk[b][a]);
      o += n;
      for (int a = 0; a < rank; ++a) coords.push_back(blk[b][a]);
    }
    size = o;
    d = qlg_bsten{rank, qn_len, nsect.data(), qn.data(), deg.data(), dir.data(), (int64_t)blk.size(),
                  coords.data(), offs.data(), dtype, data};
  }
};

QN neg(const QN &q) { return {-q[0], -q[1]}; }

}  // namespace

Plan *build_super_block_plan(Ctx *ctx, int64_t ngroups, const int64_t *nleft, const qlg_term *left,
                             const int64_t *nright, const qlg_term *right, const qlg_bsten *psid, int precision) {
  QLG_CHECK(ngroups >= 1 && nleft && left && nright && right && psid, "super block: empty term list");
  hipStream_t st = ctx->stream;
  // ---- distinct block operators (shared by descriptor pointer) ----------------------
  std::map<const qlg_bsten *, int> lid, rid;
  std::vector<Ten> Ls, Rs;
  struct T2 {
    int op;      // block operator id
    Ten site;    // site operator
  };
  std::vector<std::vector<T2>> gl(ngroups), gr(ngroups);
  int64_t pl = 0, pr = 0;
  for (int64_t j = 0; j < ngroups; ++j) {
    QLG_CHECK(nleft[j] >= 0 && nright[j] >= 0, "super block: negative term count");
    for (int64_t t = 0; t < nleft[j]; ++t, ++pl) {
      const qlg_term &tm = left[pl];
      auto it = lid.find(tm.op0);
      if (it == lid.end()) {
        it = lid.emplace(tm.op0, (int)Ls.size()).first;
        Ls.push_back(parse_tensor(tm.op0, 2, "L_i"));
      }
      gl[j].push_back(T2{it->second, parse_tensor(tm.op1, 2, "W_l(i,j)")});
    }
    for (int64_t t = 0; t < nright[j]; ++t, ++pr) {
      const qlg_term &tm = right[pr];
      auto it = rid.find(tm.op1);
      if (it == rid.end()) {
        it = rid.emplace(tm.op1, (int)Rs.size()).first;
        Rs.push_back(parse_tensor(tm.op1, 2, "R_k"));
      }
      gr[j].push_back(T2{it->second, parse_tensor(tm.op0, 2, "W_r(j,k)")});
    }
  }
  QLG_CHECK(!Ls.empty() && !Rs.empty(), "super block: no block operators");
  const int dtype = Ls[0].dtype, qn_len = Ls[0].qn_len;
  auto same_axes = [](const Ten &a, const Ten &b) {
    return a.ax[0].same_sectors(b.ax[0]) && a.ax[1].same_sectors(b.ax[1]) && a.ax[0].dir == b.ax[0].dir &&
           a.ax[1].dir == b.ax[1].dir;
  };
  for (const Ten &t : Ls) QLG_CHECK(same_axes(t, Ls[0]) && t.dtype == dtype, "super block: L_i axes differ");
  for (const Ten &t : Rs) QLG_CHECK(same_axes(t, Rs[0]) && t.dtype == dtype, "super block: R_k axes differ");
  QLG_CHECK(Ls[0].ax[0].dir == 1 && Ls[0].ax[1].dir == -1, "super block: L_i must be [a (OUT), a' (IN)]");
  QLG_CHECK(Rs[0].ax[0].dir == -1 && Rs[0].ax[1].dir == 1, "super block: R_k must be [b (IN), b' (OUT)]");
  const Ten *site0 = nullptr;
  for (auto &g : gl)
    for (auto &t : g) {
      if (!site0) site0 = &t.site;
      QLG_CHECK(same_axes(t.site, *site0) && t.site.dtype == dtype, "super block: site operator axes differ");
    }
  for (auto &g : gr)
    for (auto &t : g) QLG_CHECK(same_axes(t.site, *site0) && t.site.dtype == dtype, "super block: site axes differ");
  QLG_CHECK(site0 && site0->ax[0].dir == -1 && site0->ax[1].dir == 1,
            "super block: site operators must be [s_in (IN), s_out (OUT)]");
  auto div_of = [](const Ten &t, bool *empty) {
    *empty = t.blk.empty();
    return t.blk.empty() ? QN{0, 0} : t.flow(t.blk[0]);
  };
  // ---- bond charges ----------------------------------------------------------------------
  std::vector<QN> qw(Ls.size()), qz(Rs.size());
  std::vector<bool> lempty(Ls.size()), rempty(Rs.size());
  for (size_t i = 0; i < Ls.size(); ++i) {
    bool e;
    qw[i] = neg(div_of(Ls[i], &e));
    lempty[i] = e;
  }
  for (size_t k = 0; k < Rs.size(); ++k) {
    bool e;
    qz[k] = div_of(Rs[k], &e);
    rempty[k] = e;
  }
  std::vector<QN> qy(ngroups);
  std::vector<bool> ylive(ngroups, false);
  for (int64_t j = 0; j < ngroups; ++j) {
    bool have = false;
    QN q{0, 0};
    auto setq = [&](const QN &v) {
      if (!have) {
        q = v;
        have = true;
      } else {
        QLG_CHECK(q == v, "super block: inconsistent charges in term group " + std::to_string(j));
      }
    };
    for (auto &t : gl[j]) {
      bool e;
      const QN d = div_of(t.site, &e);
      if (!e && !lempty[t.op]) setq(qn_sub(qw[t.op], d));
    }
    for (auto &t : gr[j]) {
      bool e;
      const QN d = div_of(t.site, &e);
      if (!e && !rempty[t.op]) setq(qn_add(qz[t.op], d));
    }
    qy[j] = q;
    ylive[j] = have;
  }
  BondIndex W, Y, Z;
  W.build(qw);
  Y.build(qy);
  Z.build(qz);

  // ---- L[a', w, a] --------------------------------------------------------------------------
  const Axis &ax_a = Ls[0].ax[0], &ax_ap = Ls[0].ax[1];
  const Axis &ax_b = Rs[0].ax[0], &ax_bp = Rs[0].ax[1];
  const Axis &ax_s = site0->ax[0], &ax_sp = site0->ax[1];
  Built L, R, WL, WR;
  L.add_axis(ax_ap.qn, ax_ap.deg, -1, qn_len);
  L.add_axis(W.sect_qn, W.sect_deg, +1, qn_len);
  L.add_axis(ax_a.qn, ax_a.deg, +1, qn_len);
  std::set<Coords> lblocks;
  for (size_t i = 0; i < Ls.size(); ++i)
    for (const Coords &c : Ls[i].blk) lblocks.insert(Coords{c[1], W.el_sect[i], c[0]});
  R.add_axis(ax_bp.qn, ax_bp.deg, +1, qn_len);
  R.add_axis(Z.sect_qn, Z.sect_deg, -1, qn_len);
  R.add_axis(ax_b.qn, ax_b.deg, -1, qn_len);
  std::set<Coords> rblocks;
  for (size_t k = 0; k < Rs.size(); ++k)
    for (const Coords &c : Rs[k].blk) rblocks.insert(Coords{c[1], Z.el_sect[k], c[0]});
  const int ed = elem_doubles(dtype);
  auto bufL = std::make_unique<DevBuf>(), bufR = std::make_unique<DevBuf>();
  L.finish(lblocks, qn_len, dtype, nullptr);
  R.finish(rblocks, qn_len, dtype, nullptr);
  bufL->ensure(std::max<int64_t>(L.size, 1) * ed * sizeof(double));
  bufR->ensure(std::max<int64_t>(R.size, 1) * ed * sizeof(double));
  QLG_HIP(hipMemsetAsync(bufL->p, 0, std::max<int64_t>(L.size, 1) * ed * sizeof(double), st));
  QLG_HIP(hipMemsetAsync(bufR->p, 0, std::max<int64_t>(R.size, 1) * ed * sizeof(double), st));
  L.d.data = bufL->p;
  R.d.data = bufR->p;
  // transposed copies: op[z][x] -> stacked[x][pos][z], every block of every
  // operator in one launch (a source pointer per block); the descriptor
  // buffers live until the stream is drained below
  DevBuf dd[2], dsrc[2];
  std::vector<TransDesc> hv[2];
  std::vector<const double *> hs[2];
  auto stack = [&](int side, const std::vector<Ten> &ops, const BondIndex &bi, const Built &B, const Axis &ax_row,
                   const Axis &ax_col, double *dst) {
    std::vector<TransDesc> &v = hv[side];
    std::vector<const double *> &srcs = hs[side];
    for (size_t o = 0; o < ops.size(); ++o) {
      const Ten &t = ops[o];
      const int64_t degw = bi.sect_deg[bi.el_sect[o]];
      for (size_t b = 0; b < t.blk.size(); ++b) {
        const Coords &c = t.blk[b];       // (row sector z, col sector x) of the operator
        const int64_t nz = ax_row.deg[c[0]], nx = ax_col.deg[c[1]];
        const int64_t sb = B.index.at(Coords{c[1], bi.el_sect[o], c[0]});
        TransDesc d{};
        d.src = t.off[b];
        d.sld = nx;
        d.dst = B.offs[sb] + bi.el_pos[o] * nz;
        d.dsx = degw * nz;
        d.nz = (int32_t)nz;
        d.nx = (int32_t)nx;
        v.push_back(d);
        srcs.push_back(static_cast<const double *>(t.data));
      }
    }
    if (v.empty()) return;
    QLG_CHECK(v.size() < (size_t(1) << 31), "super block: too many operator blocks");
    dd[side].ensure(v.size() * sizeof(TransDesc));
    dsrc[side].ensure(srcs.size() * sizeof(const double *));
    QLG_HIP(hipMemcpyAsync(dd[side].p, v.data(), v.size() * sizeof(TransDesc), hipMemcpyHostToDevice, st));
    QLG_HIP(hipMemcpyAsync(dsrc[side].p, srcs.data(), srcs.size() * sizeof(const double *), hipMemcpyHostToDevice,
                           st));
    launch_trans_copy_multi(ed == 2, dd[side].as<TransDesc>(), dsrc[side].as<const double *>(), (int)v.size(), dst,
                            st);
  };
  stack(0, Ls, W, L, ax_a, ax_ap, bufL->as<double>());
  stack(1, Rs, Z, R, ax_b, ax_bp, bufR->as<double>());
  // ---- host copies of the site operators: one device-to-device gather of their
  // block data into a staging buffer, one copy to the host, one wait ----------------------
  std::map<const Ten *, int64_t> site_base;   // staging offset (elements) of each term's site operator
  int64_t stage_n = 0;
  for (auto *grp : {&gl, &gr})
    for (auto &g : *grp)
      for (auto &t : g) {
        site_base[&t.site] = stage_n;
        stage_n += raw_elems(t.site);
      }
  std::vector<double> stage_h((size_t)std::max<int64_t>(stage_n, 1) * ed);
  {
    DevBuf stage_d;
    stage_d.ensure((size_t)std::max<int64_t>(stage_n, 1) * ed * sizeof(double));
    for (auto &kv : site_base) {
      const int64_t n = raw_elems(*kv.first);
      if (n > 0)
        QLG_HIP(hipMemcpyAsync(stage_d.as<double>() + kv.second * ed, kv.first->data, n * ed * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
    }
    if (stage_n > 0)
      QLG_HIP(hipMemcpyAsync(stage_h.data(), stage_d.p, stage_n * ed * sizeof(double), hipMemcpyDeviceToHost, st));
    QLG_HIP(hipStreamSynchronize(st));   // also completes the stacking launches
  }
  // ---- W_l[w, s, s', y] and W_r[y, s, s', z] (small: assembled on the host) ------------------
  auto site_tensor = [&](Built &B, const BondIndex &bin, const BondIndex &bout,
                         const std::vector<std::vector<T2>> &groups, bool left_side, std::vector<double> &host) {
    B.add_axis(bin.sect_qn, bin.sect_deg, -1, qn_len);
    B.add_axis(ax_s.qn, ax_s.deg, -1, qn_len);
    B.add_axis(ax_sp.qn, ax_sp.deg, +1, qn_len);
    B.add_axis(bout.sect_qn, bout.sect_deg, +1, qn_len);
    // dense host copies of the site operators, then blocks where any term lands
    struct Entry {
      int64_t ein, eout;
      std::vector<double> dense;
    };
    std::vector<Entry> entries;
    std::set<Coords> blocks;
    const auto so_s = sector_offsets(ax_s), so_sp = sector_offsets(ax_sp);
    const int64_t ds = ax_s.dim(), dsp = ax_sp.dim();
    for (int64_t j = 0; j < (int64_t)groups.size(); ++j) {
      if (!ylive[j]) continue;
      for (const T2 &t : groups[j]) {
        const int64_t ein = left_side ? t.op : j, eout = left_side ? j : t.op;
        if ((left_side ? lempty[t.op] : rempty[t.op]) || t.site.blk.empty()) continue;
        Entry e{ein, eout, densify(t.site, stage_h.data() + site_base.at(&t.site) * ed, ed)};
        for (const Coords &c : t.site.blk)
          blocks.insert(Coords{bin.el_sect[ein], c[0], c[1], bout.el_sect[eout]});
        entries.push_back(std::move(e));
      }
    }
    B.finish(blocks, qn_len, dtype, nullptr);
    host.assign(std::max<int64_t>(B.size, 1) * ed, 0.0);
    for (const Entry &e : entries) {
      const int32_t wi = bin.el_sect[e.ein], wo = bout.el_sect[e.eout];
      for (size_t si = 0; si < ax_s.deg.size(); ++si)
        for (size_t so = 0; so < ax_sp.deg.size(); ++so) {
          auto it = B.index.find(Coords{wi, (int32_t)si, (int32_t)so, wo});
          if (it == B.index.end()) continue;
          const int64_t base = B.offs[it->second];
          const int64_t gi = bin.sect_deg[wi], gs = ax_s.deg[si], gsp = ax_sp.deg[so], go = bout.sect_deg[wo];
          for (int64_t a = 0; a < gs; ++a)
            for (int64_t b = 0; b < gsp; ++b) {
              const int64_t src = ((so_s[si] + a) * dsp + so_sp[so] + b) * ed;
              const int64_t dst = (((bin.el_pos[e.ein] * gs + a) * gsp + b) * go + bout.el_pos[e.eout]) * ed;
              for (int q = 0; q < ed; ++q) host[base * ed + dst + q] += e.dense[src + q];
            }
          (void)gi;
        }
    }
  };
  std::vector<double> hwl, hwr;
  site_tensor(WL, W, Y, gl, true, hwl);
  site_tensor(WR, Y, Z, gr, false, hwr);
  auto bufWL = std::make_unique<DevBuf>(), bufWR = std::make_unique<DevBuf>();
  bufWL->ensure(hwl.size() * sizeof(double));
  bufWR->ensure(hwr.size() * sizeof(double));
  QLG_HIP(hipMemcpyAsync(bufWL->p, hwl.data(), hwl.size() * sizeof(double), hipMemcpyHostToDevice, st));
  QLG_HIP(hipMemcpyAsync(bufWR->p, hwr.data(), hwr.size() * sizeof(double), hipMemcpyHostToDevice, st));
  QLG_HIP(hipStreamSynchronize(st));
  WL.d.data = bufWL->p;
  WR.d.data = bufWR->p;
  const qlg_bsten *Wd[2] = {&WL.d, &WR.d};
  Plan *p = build_plan(ctx, 2, &L.d, Wd, &R.d, psid, precision);
  // the plan keeps repacked copies; the stacked operators stay with it for its lifetime
  p->owned.push_back(std::move(bufL));
  p->owned.push_back(std::move(bufR));
  p->owned.push_back(std::move(bufWL));
  p->owned.push_back(std::move(bufWR));
  return p;
}

}  // namespace qlg
