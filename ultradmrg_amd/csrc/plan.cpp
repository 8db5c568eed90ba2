// Host-side planner: parses block-sparse descriptors, computes the closure of
// the Krylov vector layout under H_eff, counts the reference schedule's
// FLOPs, and emits the three grouped-GEMM task lists (see plan.h).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <set>

namespace qlg {

// ---------------------------------------------------------------------------
Ten parse_tensor(const qlg_bsten *d, int expect_rank, const char *name) {
  QLG_CHECK(d != nullptr, std::string(name) + ": null descriptor");
  QLG_CHECK(d->rank == expect_rank, std::string(name) + ": rank " + std::to_string(d->rank) + " != " +
                                        std::to_string(expect_rank));
  QLG_CHECK(d->qn_len >= 1 && d->qn_len <= QLG_MAX_QN_LEN, std::string(name) + ": qn_len must be 1 or 2");
  QLG_CHECK(d->dtype == QLG_F64 || d->dtype == QLG_C128, std::string(name) + ": unknown dtype");
  Ten t;
  t.rank = d->rank;
  t.qn_len = d->qn_len;
  t.dtype = d->dtype;
  t.data = d->data;
  t.ax.resize(t.rank);
  int64_t s0 = 0;
  for (int i = 0; i < t.rank; ++i) {
    const int ns = d->nsect[i];
    QLG_CHECK(ns >= 1, std::string(name) + ": axis without sectors");
    Axis &a = t.ax[i];
    a.dir = d->dir[i];
    QLG_CHECK(a.dir == 1 || a.dir == -1, std::string(name) + ": direction must be +1 or -1");
    for (int s = 0; s < ns; ++s) {
      QN q{0, 0};
      for (int c = 0; c < t.qn_len; ++c) q[c] = d->sect_qn[(s0 + s) * t.qn_len + c];
      a.qn.push_back(q);
      const int64_t dg = d->sect_deg[s0 + s];
      QLG_CHECK(dg >= 1, std::string(name) + ": sector degeneracy must be >= 1");
      a.deg.push_back(dg);
    }
    s0 += ns;
  }
  QLG_CHECK(d->nblocks >= 0, std::string(name) + ": negative block count");
  t.blk.resize(d->nblocks);
  t.off.resize(d->nblocks);
  QN div{0, 0};
  for (int64_t b = 0; b < d->nblocks; ++b) {
    Coords c{};
    for (int i = 0; i < t.rank; ++i) {
      const int32_t v = d->blk_coords[b * t.rank + i];
      QLG_CHECK(v >= 0 && v < (int32_t)t.ax[i].qn.size(), std::string(name) + ": block sector out of range");
      c[i] = v;
    }
    t.blk[b] = c;
    t.off[b] = d->blk_offset[b];
    QLG_CHECK(t.off[b] >= 0, std::string(name) + ": negative block offset");
    if (b == 0) div = t.flow(c);
    QLG_CHECK(t.flow(c) == div, std::string(name) + ": blocks with different divergence");
    QLG_CHECK(t.find.emplace(c, b).second, std::string(name) + ": duplicate block");
  }
  return t;
}

void check_pair(const Axis &a, const Axis &b, const char *what) {
  QLG_CHECK(a.same_sectors(b) && a.dir == -b.dir, std::string("contracted axes do not match: ") + what);
}

// ---------------------------------------------------------------------------
// Launch order of the tiles.  Workgroup b of a launch runs on XCD b % 8, and
// each XCD has its own L2: all tiles of one task (which share A row panels
// and B column panels) are placed on one XCD, tasks are spread over the XCDs
// by longest-processing-time-first, and each XCD walks its tasks from the
// most to the least expensive.  Slots past an XCD's share are empty tiles.
void Stage::make_tiles(int bm_, int bn_) {
  constexpr int kXcd = 8;
  bm = bm_;
  bn = bn_;
  tiles.clear();
  std::vector<int> order(tasks.size());
  std::vector<double> cost(tasks.size(), 0.0);
  for (size_t t = 0; t < tasks.size(); ++t) {
    order[t] = (int)t;
    int64_t kp = 0;
    for (int s = 0; s < tasks[t].seg_count; ++s) kp += (segs[tasks[t].seg_begin + s].K + 15) / 16 * 16;
    const double nt = double((tasks[t].M + bm - 1) / bm) * double((tasks[t].N + bn - 1) / bn);
    cost[t] = nt * (double(bm) * bn * std::max<int64_t>(kp, 16) + 4096.0);
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::vector<std::vector<GemmTile>> per(kXcd);
  std::vector<double> load(kXcd, 0.0);
  for (int t : order) {
    const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += cost[t];
    const GemmTask &k = tasks[t];
    for (int m0 = 0; m0 < k.M; m0 += bm)
      for (int n0 = 0; n0 < k.N; n0 += bn) per[x].push_back(GemmTile{t, m0, n0, 0});
  }
  size_t slots = 0;
  for (auto &v : per) slots = std::max(slots, v.size());
  tiles.reserve(slots * kXcd);
  for (size_t s = 0; s < slots; ++s)
    for (int x = 0; x < kXcd; ++x) tiles.push_back(s < per[x].size() ? per[x][s] : GemmTile{-1, 0, 0, 0});
  while (!tiles.empty() && tiles.back().task < 0) tiles.pop_back();
}

void Stage::upload(hipStream_t st) {
  d_tiles.ensure(std::max<size_t>(1, tiles.size()) * sizeof(GemmTile));
  d_tasks.ensure(std::max<size_t>(1, tasks.size()) * sizeof(GemmTask));
  d_segs.ensure(std::max<size_t>(1, segs.size()) * sizeof(GemmSeg));
  if (!tiles.empty())
    QLG_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), tiles.size() * sizeof(GemmTile), hipMemcpyHostToDevice, st));
  if (!tasks.empty())
    QLG_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(GemmTask), hipMemcpyHostToDevice, st));
  if (!segs.empty())
    QLG_HIP(hipMemcpyAsync(d_segs.p, segs.data(), segs.size() * sizeof(GemmSeg), hipMemcpyHostToDevice, st));
}

// Bounds of every access a task list can make: rows < M, cols < N, k < K of
// each segment, relative to the buffer behind its slot (in elements).
void validate_stage(const Stage &s, const int64_t *slot_elems) {
  auto inside = [&](int slot, int64_t first, int64_t last, const char *what) {
    QLG_CHECK(slot > SLOT_ABS && slot < NSLOTS, std::string("task list: bad slot for ") + what);
    QLG_CHECK(first >= 0 && last < slot_elems[slot],
              std::string("task list: ") + what + " extent [" + std::to_string(first) + ", " + std::to_string(last) +
                  "] outside buffer of " + std::to_string(slot_elems[slot]) + " elements");
  };
  for (const GemmTask &t : s.tasks) {
    QLG_CHECK(t.M > 0 && t.N > 0 && t.seg_count >= 0, "task list: empty task shape");
    QLG_CHECK(t.ldc >= t.N, "task list: ldc < N");
    inside(t.c_slot, t.c_off, t.c_off + (int64_t)(t.M - 1) * t.ldc + (t.N - 1), "C");
    for (int q = 0; q < t.seg_count; ++q) {
      const GemmSeg &g = s.segs[t.seg_begin + q];
      if (g.K <= 0) continue;
      if (g.flags & SEG_TRANS_A) {
        QLG_CHECK(g.lda >= t.M, "task list: lda < M");
        inside(g.a_slot, g.a_off, g.a_off + (int64_t)(g.K - 1) * g.lda + (t.M - 1), "A^T");
      } else {
        QLG_CHECK(g.lda >= g.K, "task list: lda < K");
        inside(g.a_slot, g.a_off, g.a_off + (int64_t)(t.M - 1) * g.lda + (g.K - 1), "A");
      }
      if (g.flags & SEG_TRANS_B) {
        QLG_CHECK(g.ldb >= g.K, "task list: ldb < K");
        inside(g.b_slot, g.b_off, g.b_off + (int64_t)(t.N - 1) * g.ldb + (g.K - 1), "B^T");
      } else {
        QLG_CHECK(g.ldb >= t.N, "task list: ldb < N");
        inside(g.b_slot, g.b_off, g.b_off + (int64_t)(g.K - 1) * g.ldb + (t.N - 1), "B");
      }
    }
  }
}

// Pick the FP64 tile shape minimising a simple cost model: MFMA time of the
// padded tiles spread over 256 CUs (2 workgroups each), bounded below by the
// longest single tile, plus a fixed per-tile cost.
void autotune_tiles(Stage &s, bool cplx) {
  static const int shapes[][2] = {{32, 32}, {64, 32}, {32, 64}, {64, 64}, {128, 64}, {64, 128}, {128, 128}};
  const int nshapes = cplx ? 4 : 7;   // complex tiles keep two accumulator planes: cap at 64x64
  double best = 1e300;
  int bm = 64, bn = 64;
  std::vector<int64_t> kpad(s.tasks.size(), 0);
  for (size_t t = 0; t < s.tasks.size(); ++t)
    for (int q = 0; q < s.tasks[t].seg_count; ++q) kpad[t] += (s.segs[s.tasks[t].seg_begin + q].K + 15) / 16 * 16;
  const double ebytes = cplx ? 16.0 : 8.0, fmul = cplx ? 4.0 : 1.0;
  for (int si = 0; si < nshapes; ++si) {
    const int *sh = shapes[si];
    double flops = 0, bytes = 0, longest = 0, ntiles = 0;
    for (size_t t = 0; t < s.tasks.size(); ++t) {
      const GemmTask &k = s.tasks[t];
      const double nt = double((k.M + sh[0] - 1) / sh[0]) * double((k.N + sh[1] - 1) / sh[1]);
      const double kp = (double)std::max<int64_t>(kpad[t], 16);
      flops += nt * 2.0 * sh[0] * sh[1] * kp * fmul;
      // operand panels streamed per tile + the output tile
      bytes += nt * ((sh[0] + sh[1]) * kp + double(sh[0]) * sh[1]) * ebytes;
      // one workgroup alone: 4 waves, one 64-cycle FP64 MFMA per 16x16x4 block each, 2.4 GHz
      longest = std::max(longest, sh[0] * sh[1] * kp * fmul / 64.0 / 2.4e9);
      ntiles += nt;
    }
    // MFMA time at the FP64 peak, operand streaming at an L2-assisted 12 TB/s,
    // and ~1.5 us of exposed latency per tile spread over 1024 resident tiles
    const double est = std::max({flops / 78.6e12, bytes / 12e12, longest}) + ntiles * 1.5e-6 / 1024.0;
    if (est < best * 0.97) {
      best = est;
      bm = sh[0];
      bn = sh[1];
    }
  }
  s.make_tiles(bm, bn);
}

// ---------------------------------------------------------------------------
namespace {

struct PhysComp {
  int sect;     // sector id
  int local;    // component inside the sector
};

// all components of an axis, flattened
std::vector<PhysComp> comps(const Axis &a) {
  std::vector<PhysComp> v;
  for (size_t s = 0; s < a.deg.size(); ++s)
    for (int64_t i = 0; i < a.deg[s]; ++i) v.push_back({(int)s, (int)i});
  return v;
}

std::vector<int64_t> sector_offsets(const Axis &a) {
  std::vector<int64_t> o(a.deg.size() + 1, 0);
  for (size_t s = 0; s < a.deg.size(); ++s) o[s + 1] = o[s] + a.deg[s];
  return o;
}

// dense host copy of a small tensor (MPO), index order kept
std::vector<double> dense_copy(const Ten &t, hipStream_t st, int ed) {
  std::vector<int64_t> dims(t.rank);
  std::vector<std::vector<int64_t>> so(t.rank);
  int64_t tot = 1;
  for (int i = 0; i < t.rank; ++i) {
    dims[i] = t.ax[i].dim();
    so[i] = sector_offsets(t.ax[i]);
    tot *= dims[i];
  }
  int64_t raw = 0;
  for (size_t b = 0; b < t.blk.size(); ++b) raw = std::max(raw, t.off[b] + t.blk_elems(b));
  std::vector<double> h(raw * ed);
  if (raw > 0) {
    QLG_HIP(hipMemcpyAsync(h.data(), t.data, raw * ed * sizeof(double), hipMemcpyDeviceToHost, st));
    QLG_HIP(hipStreamSynchronize(st));
  }
  std::vector<double> d(tot * ed, 0.0);
  for (size_t b = 0; b < t.blk.size(); ++b) {
    std::vector<int64_t> bd(t.rank);
    int64_t n = 1;
    for (int i = 0; i < t.rank; ++i) { bd[i] = t.ax[i].deg[t.blk[b][i]]; n *= bd[i]; }
    for (int64_t e = 0; e < n; ++e) {
      int64_t r = e, gi = 0;
      std::vector<int64_t> loc(t.rank);
      for (int i = t.rank - 1; i >= 0; --i) { loc[i] = r % bd[i]; r /= bd[i]; }
      for (int i = 0; i < t.rank; ++i) gi = gi * dims[i] + so[i][t.blk[b][i]] + loc[i];
      for (int c = 0; c < ed; ++c) d[gi * ed + c] = h[(t.off[b] + e) * ed + c];
    }
  }
  return d;
}

using Key6 = std::array<int32_t, 6>;

}  // namespace

// ---------------------------------------------------------------------------
Plan *build_plan(Ctx *ctx, int nphys, const qlg_bsten *Ld, const qlg_bsten *const *Wd, const qlg_bsten *Rd,
                 const qlg_bsten *psid, int precision) {
  QLG_CHECK(nphys == 1 || nphys == 2, "nphys must be 1 or 2");
  std::unique_ptr<Plan> P(new Plan());
  Plan &p = *P;
  p.ctx = ctx;
  p.nphys = nphys;
  p.precision = precision;
  p.L = parse_tensor(Ld, 3, "L");
  for (int k = 0; k < nphys; ++k) p.W[k] = parse_tensor(Wd[k], 4, k == 0 ? "Wl" : "Wr");
  p.R = parse_tensor(Rd, 3, "R");
  Ten psi = parse_tensor(psid, 2 + nphys, "psi");
  p.dtype = psi.dtype;
  const int qn_len = psi.qn_len;
  for (const Ten *t : {&p.L, &p.W[0], &p.W[nphys - 1], &p.R}) {
    QLG_CHECK(t->dtype == p.dtype, "all operands must share one dtype");
    QLG_CHECK(t->qn_len == qn_len, "all operands must share one quantum-number length");
  }
  const Ten &L = p.L, &R = p.R;
  const Ten &W0 = p.W[0], &W1 = p.W[nphys - 1];
  // contracted-axis compatibility (reference index orders)
  check_pair(L.ax[2], psi.ax[0], "L ket / psi left");
  check_pair(L.ax[1], W0.ax[0], "L mpo / W left");
  for (int k = 0; k < nphys; ++k) check_pair(p.W[k].ax[1], psi.ax[1 + k], "W s_in / psi physical");
  if (nphys == 2) check_pair(W0.ax[3], W1.ax[0], "Wl right / Wr left");
  check_pair(W1.ax[3], R.ax[1], "W right / R mpo");
  check_pair(R.ax[2], psi.ax[1 + nphys], "R ket / psi right");
  // result axes must be the state's axes so that H_eff maps the space into itself
  {
    std::vector<const Axis *> out_ax = {&L.ax[0]};
    for (int k = 0; k < nphys; ++k) out_ax.push_back(&p.W[k].ax[2]);
    out_ax.push_back(&R.ax[0]);
    for (int i = 0; i < psi.rank; ++i)
      QLG_CHECK(out_ax[i]->same_sectors(psi.ax[i]) && out_ax[i]->dir == psi.ax[i].dir,
                "H_eff output axis " + std::to_string(i) + " differs from the state's axis");
  }
  const int ed = elem_doubles(p.dtype);
  const double fl = p.dtype == QLG_C128 ? 8.0 : 2.0;
  const int rk = psi.rank;
  const int bax = rk - 1;  // right bond axis of psi

  // ---- indexes of operator blocks ------------------------------------------
  std::map<int, std::vector<int64_t>> L_by_ket;             // a -> L blocks
  for (size_t b = 0; b < L.blk.size(); ++b) L_by_ket[L.blk[b][2]].push_back(b);
  std::map<std::pair<int, int>, std::vector<int64_t>> W_by_in[2];   // (wl, s_in) -> W blocks
  for (int k = 0; k < nphys; ++k)
    for (size_t b = 0; b < p.W[k].blk.size(); ++b)
      W_by_in[k][{p.W[k].blk[b][0], p.W[k].blk[b][1]}].push_back(b);
  std::map<std::pair<int, int>, std::vector<int64_t>> R_by_zb;    // (z, b) -> R blocks
  std::map<int, std::vector<int64_t>> R_by_bra;                  // b' -> R blocks
  for (size_t b = 0; b < R.blk.size(); ++b) {
    R_by_zb[{R.blk[b][1], R.blk[b][2]}].push_back(b);
    R_by_bra[R.blk[b][0]].push_back(b);
  }

  // ---- symbolic reference schedule: closure of the block set + FLOPs --------
  auto symbolic = [&](const std::set<Coords> &in, double *flops) {
    std::set<Coords> out;
    double f = 0;
    // stage T1: (a', w, s..., b)
    std::set<Coords> t1;
    for (const Coords &c : in) {
      auto it = L_by_ket.find(c[0]);
      if (it == L_by_ket.end()) continue;
      int64_t nfree = 1;
      for (int i = 1; i < rk; ++i) nfree *= psi.ax[i].deg[c[i]];
      for (int64_t lb : it->second) {
        const Coords &l = L.blk[lb];
        Coords t{};
        t[0] = l[0];
        t[1] = l[1];
        for (int i = 1; i < rk; ++i) t[1 + i] = c[i];
        t1.insert(t);
        f += fl * double(L.ax[0].deg[l[0]] * L.ax[1].deg[l[1]]) * double(L.ax[2].deg[l[2]]) * double(nfree);
      }
    }
    // MPO stages: carry (a', [s' done...], w_cur, [s remaining...], b)
    // represent as Coords: [a', t_0..t_{k-1}, w, s_k..s_{n-1}, b]
    std::set<Coords> cur = t1;  // k = 0: [a', w, s0, (s1), b]
    for (int k = 0; k < nphys; ++k) {
      std::set<Coords> nxt;
      for (const Coords &c : cur) {
        const int w = c[1 + k], s = c[2 + k];
        auto it = W_by_in[k].find({w, s});
        if (it == W_by_in[k].end()) continue;
        // free dims of the intermediate other than (w, s)
        int64_t mfree = L.ax[0].deg[c[0]] * R.ax[2].deg[c[rk]];   // a' and b
        for (int j = 0; j < k; ++j) mfree *= p.W[j].ax[2].deg[c[1 + j]];
        for (int j = k + 1; j < nphys; ++j) mfree *= psi.ax[1 + j].deg[c[2 + j]];
        const int64_t kdim = p.W[k].ax[0].deg[w] * p.W[k].ax[1].deg[s];
        for (int64_t wb : it->second) {
          const Coords &wc = p.W[k].blk[wb];
          Coords t = c;
          t[1 + k] = wc[2];   // s_out
          t[2 + k] = wc[3];   // w right
          nxt.insert(t);
          f += fl * double(mfree) * double(kdim) * double(p.W[k].ax[2].deg[wc[2]] * p.W[k].ax[3].deg[wc[3]]);
        }
      }
      cur.swap(nxt);
    }
    // cur: [a', t..., z, b] ; contract (z, b) with R
    for (const Coords &c : cur) {
      const int z = c[1 + nphys], b = c[2 + nphys];
      auto it = R_by_zb.find({z, b});
      if (it == R_by_zb.end()) continue;
      int64_t m = L.ax[0].deg[c[0]];
      for (int j = 0; j < nphys; ++j) m *= p.W[j].ax[2].deg[c[1 + j]];
      const int64_t kd = R.ax[1].deg[z] * R.ax[2].deg[b];
      for (int64_t rb : it->second) {
        Coords o{};
        o[0] = c[0];
        for (int j = 0; j < nphys; ++j) o[1 + j] = c[1 + j];
        o[bax] = R.blk[rb][0];
        out.insert(o);
        f += fl * double(m) * double(kd) * double(R.ax[0].deg[R.blk[rb][0]]);
      }
    }
    if (flops) *flops = f;
    return out;
  };

  std::set<Coords> vset(psi.blk.begin(), psi.blk.end());
  for (int iter = 0; iter < 16; ++iter) {
    std::set<Coords> out = symbolic(vset, nullptr);
    size_t before = vset.size();
    vset.insert(out.begin(), out.end());
    if (vset.size() == before) break;
  }
  symbolic(vset, &p.algo_flops);

  // ---- vector layout ----------------------------------------------------------
  p.vec.rank = rk;
  p.vec.ax = psi.ax;
  p.vec.blk.assign(vset.begin(), vset.end());   // std::set order = lexicographic
  p.vec.off.resize(p.vec.blk.size());
  {
    int64_t o = 0;
    for (size_t b = 0; b < p.vec.blk.size(); ++b) {
      p.vec.off[b] = o;
      p.vec.find[p.vec.blk[b]] = b;
      o += p.vec.blk_elems(b);
    }
    p.vec.size = o;
  }
  std::map<int, std::vector<int64_t>> V_by_left;
  for (size_t b = 0; b < p.vec.blk.size(); ++b) V_by_left[p.vec.blk[b][0]].push_back(b);

  // ---- stage 1: chunks of channels (a', b) -----------------------------------------
  struct Chunk {
    int64_t lblk, vblk;
    int w, sl[2];        // w component, physical local components
    int32_t wg, sg[2];   // global component indexes
  };
  struct Channel {
    int ap, b;
    int64_t P;
    std::vector<Chunk> in;
    std::map<Key6, int> out_index;
    std::vector<Key6> out;
    int64_t x_off = 0, y_off = 0, c_off = 0;
  };
  std::map<std::pair<int, int>, int> chan_id;
  std::vector<Channel> chans;
  auto so_w0 = sector_offsets(W0.ax[0]);
  std::vector<int64_t> so_s[2];
  for (int k = 0; k < nphys; ++k) so_s[k] = sector_offsets(psi.ax[1 + k]);
  for (size_t lb = 0; lb < L.blk.size(); ++lb) {
    const Coords &l = L.blk[lb];
    auto it = V_by_left.find(l[2]);
    if (it == V_by_left.end()) continue;
    for (int64_t vb : it->second) {
      const Coords &v = p.vec.blk[vb];
      auto key = std::make_pair(l[0], v[bax]);
      auto ci = chan_id.find(key);
      int cid;
      if (ci == chan_id.end()) {
        cid = (int)chans.size();
        chan_id[key] = cid;
        Channel c;
        c.ap = l[0];
        c.b = v[bax];
        c.P = L.ax[0].deg[l[0]] * psi.ax[bax].deg[v[bax]];
        chans.push_back(c);
      } else {
        cid = ci->second;
      }
      const int64_t Wm = L.ax[1].deg[l[1]];
      const int64_t d0 = psi.ax[1].deg[v[1]], d1 = nphys == 2 ? psi.ax[2].deg[v[2]] : 1;
      for (int w = 0; w < Wm; ++w)
        for (int a = 0; a < d0; ++a)
          for (int c2 = 0; c2 < d1; ++c2) {
            Chunk ch;
            ch.lblk = lb;
            ch.vblk = vb;
            ch.w = w;
            ch.sl[0] = a;
            ch.sl[1] = c2;
            ch.wg = (int32_t)(so_w0[l[1]] + w);
            ch.sg[0] = (int32_t)(so_s[0][v[1]] + a);
            ch.sg[1] = nphys == 2 ? (int32_t)(so_s[1][v[2]] + c2) : 0;
            chans[cid].in.push_back(ch);
          }
    }
  }

  // ---- stage 2: output rows of each channel (structural reachability) ----------
  // W dense copies for the combined operator
  std::vector<double> Wdn[2];
  std::vector<int64_t> wdims[2];
  for (int k = 0; k < nphys; ++k) {
    Wdn[k] = dense_copy(p.W[k], ctx->stream, ed);
    for (int i = 0; i < 4; ++i) wdims[k].push_back(p.W[k].ax[i].dim());
  }
  std::vector<int64_t> so_t[2];
  for (int k = 0; k < nphys; ++k) so_t[k] = sector_offsets(p.W[k].ax[2]);
  auto so_z = sector_offsets(W1.ax[3]);
  // R blocks reachable per (z sector, b sector): need at least one
  for (Channel &c : chans) {
    std::set<Key6> outs;
    for (const Chunk &ch : c.in) {
      const Coords &l = L.blk[ch.lblk];
      const Coords &v = p.vec.blk[ch.vblk];
      // walk the MPO chain structurally: (w sector, s sectors) -> (t sectors, z sector)
      std::vector<std::pair<std::array<int, 2>, int>> frontier;  // (t sectors, current w sector)
      frontier.push_back({{-1, -1}, l[1]});
      for (int k = 0; k < nphys; ++k) {
        std::vector<std::pair<std::array<int, 2>, int>> nf;
        for (auto &fr : frontier) {
          auto it = W_by_in[k].find({fr.second, v[1 + k]});
          if (it == W_by_in[k].end()) continue;
          for (int64_t wb : it->second) {
            auto t = fr.first;
            t[k] = p.W[k].blk[wb][2];
            nf.push_back({t, p.W[k].blk[wb][3]});
          }
        }
        frontier.swap(nf);
      }
      for (auto &fr : frontier) {
        const int z = fr.second;
        if (R_by_zb.find({z, c.b}) == R_by_zb.end()) continue;
        const int64_t dz = W1.ax[3].deg[z];
        const int64_t dt0 = p.W[0].ax[2].deg[fr.first[0]];
        const int64_t dt1 = nphys == 2 ? p.W[1].ax[2].deg[fr.first[1]] : 1;
        for (int a = 0; a < dt0; ++a)
          for (int b2 = 0; b2 < dt1; ++b2)
            for (int zz = 0; zz < dz; ++zz)
              outs.insert(Key6{fr.first[0], a, nphys == 2 ? fr.first[1] : 0, b2, z, zz});
      }
    }
    c.out.assign(outs.begin(), outs.end());
    for (size_t i = 0; i < c.out.size(); ++i) c.out_index[c.out[i]] = (int)i;
  }

  // ---- workspace offsets --------------------------------------------------------------
  int64_t xo = 0, yo = 0, co = 0;
  for (Channel &c : chans) {
    c.x_off = xo;
    xo += (int64_t)c.in.size() * c.P;
    c.y_off = yo;
    yo += (int64_t)c.out.size() * c.P;
    c.c_off = co;
    co += (int64_t)c.out.size() * c.in.size();
  }
  p.x_elems = xo;
  p.y_elems = yo;
  p.c_elems = co;

  // ---- combined operator matrices ---------------------------------------------------
  std::vector<double> Ch(std::max<int64_t>(co, 1) * ed, 0.0);
  auto Widx = [&](int k, int64_t a, int64_t b, int64_t c, int64_t d) {
    return ((a * wdims[k][1] + b) * wdims[k][2] + c) * wdims[k][3] + d;
  };
  for (Channel &c : chans) {
    const int64_t K = (int64_t)c.in.size();
    for (size_t r = 0; r < c.out.size(); ++r) {
      const Key6 &o = c.out[r];
      const int64_t tg0 = so_t[0][o[0]] + o[1];
      const int64_t tg1 = nphys == 2 ? so_t[1][o[2]] + o[3] : 0;
      const int64_t zg = so_z[o[4]] + o[5];
      for (int64_t q = 0; q < K; ++q) {
        const Chunk &ch = c.in[q];
        double vr = 0, vi = 0;
        if (nphys == 1) {
          const int64_t i = Widx(0, ch.wg, ch.sg[0], tg0, zg);
          vr = Wdn[0][i * ed];
          if (ed == 2) vi = Wdn[0][i * ed + 1];
        } else {
          const int64_t ny = wdims[0][3];
          for (int64_t y = 0; y < ny; ++y) {
            const int64_t i0 = Widx(0, ch.wg, ch.sg[0], tg0, y);
            const int64_t i1 = Widx(1, y, ch.sg[1], tg1, zg);
            if (ed == 1) {
              vr += Wdn[0][i0] * Wdn[1][i1];
            } else {
              const double ar = Wdn[0][2 * i0], ai = Wdn[0][2 * i0 + 1];
              const double br = Wdn[1][2 * i1], bi = Wdn[1][2 * i1 + 1];
              vr += ar * br - ai * bi;
              vi += ar * bi + ai * br;
            }
          }
        }
        const int64_t e = c.c_off + (int64_t)r * K + q;
        Ch[e * ed] = vr;
        if (ed == 2) Ch[e * ed + 1] = vi;
      }
    }
  }

  // ---- emit tasks ------------------------------------------------------------------
  Stage &s1 = p.st[0], &s2 = p.st[1], &s3 = p.st[2];
  for (Channel &c : chans) {
    const int64_t Dap = L.ax[0].deg[c.ap], Db = psi.ax[bax].deg[c.b];
    for (size_t q = 0; q < c.in.size(); ++q) {
      const Chunk &ch = c.in[q];
      const Coords &l = L.blk[ch.lblk];
      const Coords &v = p.vec.blk[ch.vblk];
      const int64_t Da = L.ax[2].deg[l[2]], Wm = L.ax[1].deg[l[1]];
      const int64_t d0 = psi.ax[1].deg[v[1]], d1 = nphys == 2 ? psi.ax[2].deg[v[2]] : 1;
      GemmSeg sg{};
      sg.a_slot = SLOT_L;
      sg.a_off = L.off[ch.lblk] + ch.w * Da;
      sg.lda = Wm * Da;
      sg.b_slot = SLOT_IN;
      sg.b_off = p.vec.off[ch.vblk] + (ch.sl[0] * d1 + ch.sl[1]) * Db;
      sg.ldb = d0 * d1 * Db;
      sg.K = (int32_t)Da;
      sg.flags = 0;
      GemmTask tk{};
      tk.c_slot = SLOT_X;
      tk.c_off = c.x_off + (int64_t)q * c.P;
      tk.ldc = Db;
      tk.M = (int32_t)Dap;
      tk.N = (int32_t)Db;
      tk.seg_begin = (int32_t)s1.segs.size();
      tk.seg_count = 1;
      s1.segs.push_back(sg);
      s1.tasks.push_back(tk);
      s1.flops += fl * double(Dap) * double(Db) * double(Da);
    }
    if (!c.out.empty() && !c.in.empty()) {
      GemmSeg sg{};
      sg.a_slot = SLOT_C;
      sg.a_off = c.c_off;
      sg.lda = (int64_t)c.in.size();
      sg.b_slot = SLOT_X;
      sg.b_off = c.x_off;
      sg.ldb = c.P;
      sg.K = (int32_t)c.in.size();
      GemmTask tk{};
      tk.c_slot = SLOT_Y;
      tk.c_off = c.y_off;
      tk.ldc = c.P;
      tk.M = (int32_t)c.out.size();
      tk.N = (int32_t)c.P;
      tk.seg_begin = (int32_t)s2.segs.size();
      tk.seg_count = 1;
      s2.segs.push_back(sg);
      s2.tasks.push_back(tk);
      s2.flops += fl * double(c.out.size()) * double(c.in.size()) * double(c.P);
    }
  }
  // stage 3: one task per output component block
  for (size_t vb = 0; vb < p.vec.blk.size(); ++vb) {
    const Coords &o = p.vec.blk[vb];
    const int ap = o[0], bp = o[bax];
    const int64_t Dap = L.ax[0].deg[ap], Dbp = R.ax[0].deg[bp];
    const int64_t d0 = p.W[0].ax[2].deg[o[1]], d1 = nphys == 2 ? p.W[1].ax[2].deg[o[2]] : 1;
    auto rit = R_by_bra.find(bp);
    for (int a = 0; a < d0; ++a)
      for (int b2 = 0; b2 < d1; ++b2) {
        GemmTask tk{};
        tk.c_slot = SLOT_OUT;
        tk.c_off = p.vec.off[vb] + (a * d1 + b2) * Dbp;
        tk.ldc = d0 * d1 * Dbp;
        tk.M = (int32_t)Dap;
        tk.N = (int32_t)Dbp;
        tk.seg_begin = (int32_t)s3.segs.size();
        tk.seg_count = 0;
        if (rit != R_by_bra.end()) {
          for (int64_t rb : rit->second) {
            const int z = R.blk[rb][1], b = R.blk[rb][2];
            auto ci = chan_id.find({ap, b});
            if (ci == chan_id.end()) continue;
            const Channel &c = chans[ci->second];
            const int64_t Db = R.ax[2].deg[b], Wz = R.ax[1].deg[z];
            for (int zz = 0; zz < Wz; ++zz) {
              auto oi = c.out_index.find(Key6{o[1], a, nphys == 2 ? o[2] : 0, b2, z, zz});
              if (oi == c.out_index.end()) continue;
              GemmSeg sg{};
              sg.a_slot = SLOT_Y;
              sg.a_off = c.y_off + (int64_t)oi->second * c.P;
              sg.lda = Db;
              sg.b_slot = SLOT_R;
              sg.b_off = R.off[rb] + zz * Db;
              sg.ldb = Wz * Db;
              sg.flags = SEG_TRANS_B;
              sg.K = (int32_t)Db;
              s3.segs.push_back(sg);
              tk.seg_count++;
              s3.flops += fl * double(Dap) * double(Dbp) * double(Db);
            }
          }
        }
        s3.tasks.push_back(tk);
      }
  }
  // every operand extent of every task must lie inside its buffer
  {
    auto extent = [](const Ten &t) {
      int64_t e = 0;
      for (size_t b = 0; b < t.blk.size(); ++b) e = std::max(e, t.off[b] + t.blk_elems(b));
      return e;
    };
    int64_t slot_elems[NSLOTS] = {0};
    slot_elems[SLOT_IN] = slot_elems[SLOT_OUT] = p.vec.size;
    slot_elems[SLOT_L] = extent(L);
    slot_elems[SLOT_R] = extent(R);
    slot_elems[SLOT_X] = p.x_elems;
    slot_elems[SLOT_Y] = p.y_elems;
    slot_elems[SLOT_C] = p.c_elems;
    for (int i = 0; i < 3; ++i) validate_stage(p.st[i], slot_elems);
  }
  // QLG_TILES="BMxBN,BMxBN,BMxBN" pins the tile shape of each stage (tuning runs)
  const char *pin = std::getenv("QLG_TILES");
  int pinned[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  if (pin) std::sscanf(pin, "%dx%d,%dx%d,%dx%d", &pinned[0][0], &pinned[0][1], &pinned[1][0], &pinned[1][1],
                       &pinned[2][0], &pinned[2][1]);
  for (int i = 0; i < 3; ++i) {
    if (pinned[i][0] > 0)
      p.st[i].make_tiles(pinned[i][0], pinned[i][1]);
    else
      autotune_tiles(p.st[i], p.dtype == QLG_C128);
    p.st[i].upload(ctx->stream);
  }
  if (const char *v = std::getenv("QLG_VERBOSE"); v && v[0] == '1') {
    for (int i = 0; i < 3; ++i) {
      const Stage &s = p.st[i];
      double padded = 0;
      for (const GemmTile &t : s.tiles) {
        if (t.task < 0) continue;
        const GemmTask &k = s.tasks[t.task];
        int64_t kp = 0;
        for (int q = 0; q < k.seg_count; ++q) kp += (s.segs[k.seg_begin + q].K + 15) / 16 * 16;
        padded += 2.0 * s.bm * s.bn * kp * (p.dtype == QLG_C128 ? 4 : 1);
      }
      std::fprintf(stderr, "[qlg plan] stage %d: tasks %zu segs %zu tiles %zu tile %dx%d gflop %.3f padded-eff %.3f\n",
                   i + 1, s.tasks.size(), s.segs.size(), s.tiles.size(), s.bm, s.bn, s.flops / 1e9,
                   padded > 0 ? s.flops / padded : 0.0);
    }
    std::fprintf(stderr, "[qlg plan] vector %lld elems, X %lld, Y %lld, C %lld, algo gflop %.3f\n",
                 (long long)p.vec.size, (long long)p.x_elems, (long long)p.y_elems, (long long)p.c_elems,
                 p.algo_flops / 1e9);
  }
  p.X.ensure(std::max<int64_t>(p.x_elems, 1) * ed * sizeof(double));
  p.Y.ensure(std::max<int64_t>(p.y_elems, 1) * ed * sizeof(double));
  p.C.ensure(std::max<int64_t>(p.c_elems, 1) * ed * sizeof(double));
  QLG_HIP(hipMemcpyAsync(p.C.p, Ch.data(), std::max<int64_t>(co, 1) * ed * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  QLG_HIP(hipStreamSynchronize(ctx->stream));
  return P.release();
}

// ---------------------------------------------------------------------------
Plan::~Plan() {
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
}

void Plan::matvec(const double *in, double *out) {
  Slots s{};
  s.p[SLOT_ABS] = nullptr;
  s.p[SLOT_IN] = const_cast<double *>(in);
  s.p[SLOT_OUT] = out;
  s.p[SLOT_L] = const_cast<double *>(static_cast<const double *>(L.data));
  s.p[SLOT_R] = const_cast<double *>(static_cast<const double *>(R.data));
  s.p[SLOT_X] = X.as<double>();
  s.p[SLOT_Y] = Y.as<double>();
  s.p[SLOT_C] = C.as<double>();
  hipStream_t stream = ctx->stream;
  if (timing) {
    for (auto &e : ev)
      if (!e) QLG_HIP(hipEventCreate(&e));
    QLG_HIP(hipEventRecord(ev[0], stream));
  }
  for (int i = 0; i < 3; ++i) {
    Stage &g = st[i];
    launch_grouped_gemm_f64(g.bm, g.bn, dtype == QLG_C128, g.d_tiles.as<GemmTile>(), g.d_tasks.as<GemmTask>(),
                            g.d_segs.as<GemmSeg>(), s, (int)g.tiles.size(), stream);
    if (timing) QLG_HIP(hipEventRecord(ev[i + 1], stream));
  }
}

}  // namespace qlg
