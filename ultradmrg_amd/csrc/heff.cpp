// Effective-Hamiltonian planner (two-site and one-site).
//
// The sector pairing is done once per plan and turned into three launches:
//  stage 1  per ket sector a of the state:  T1_a = L_a . Psi_a, where L_a stacks
//           every L block (a', w, a) (rows (a', w)) and Psi_a is the state's
//           fused matrix [D_a x N_a] (all blocks (a, s1, s2, b) side by side);
//  stage 2  per channel (a'-sector, b-sector): Y = C . X on rows gathered from
//           T1 (one per (w, s1, s2) component) and scattered into Y, C being the
//           combined MPO operator sum_y Wl[w,s1,t1,y] Wr[y,s2,t2,z];
//  stage 3  per bra sector b': out = Y_b' . R_b'^T, where R_b' holds every R
//           block (b', z, b) side by side (one long K) and Y_b' the rows of all
//           output blocks with bra b'; output rows land in the fused layout.
// The Krylov vectors inside the library use the fused layout; the boundary
// keeps the canonical one (pack / unpack copies at the solver's entry / exit).
#include <algorithm>
#include <iterator>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <set>

#include "host_pool.h"
#include "plan.h"

namespace qlg {

namespace {
struct PairHash {
  size_t operator()(const std::pair<int, int> &p) const {
    return std::hash<uint64_t>()((uint64_t(uint32_t(p.first)) << 32) | uint32_t(p.second));
  }
};
// QLG_VERBOSE=2: host-side phase times of plan construction (stderr)
struct PhaseTimer {
  bool on;
  std::chrono::steady_clock::time_point t0, last;
  std::string line;
  PhaseTimer() {
    const char *v = std::getenv("QLG_VERBOSE");
    on = v && v[0] == '2';
    t0 = last = std::chrono::steady_clock::now();
  }
  void mark(const char *name) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    line += std::string(" ") + name + " " +
            std::to_string(std::chrono::duration<double, std::milli>(now - last).count()).substr(0, 6);
    last = now;
  }
  ~PhaseTimer() {
    if (on)
      std::fprintf(stderr, "[qlg plan ms]%s total %.3f\n", line.c_str(),
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
};
}  // namespace

Plan *build_plan(Ctx *ctx, int nphys, const qlg_bsten *Ld, const qlg_bsten *const *Wd, const qlg_bsten *Rd,
                 const qlg_bsten *psid, int precision) {
  PhaseTimer pt;
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

  pt.mark("parse");
  // ---- indexes of operator blocks ------------------------------------------
  std::map<int, std::vector<int64_t>> L_by_ket;             // a -> L blocks
  for (size_t b = 0; b < L.blk.size(); ++b) L_by_ket[L.blk[b][2]].push_back(b);
  std::unordered_map<std::pair<int, int>, std::vector<int64_t>, PairHash> W_by_in[2];   // (wl, s_in) -> W blocks
  for (int k = 0; k < nphys; ++k)
    for (size_t b = 0; b < p.W[k].blk.size(); ++b)
      W_by_in[k][{p.W[k].blk[b][0], p.W[k].blk[b][1]}].push_back(b);
  std::unordered_map<std::pair<int, int>, std::vector<int64_t>, PairHash> R_by_zb;    // (z, b) -> R blocks
  std::map<int, std::vector<int64_t>> R_by_bra;                  // b' -> R blocks
  for (size_t b = 0; b < R.blk.size(); ++b) {
    R_by_zb[{R.blk[b][1], R.blk[b][2]}].push_back(b);
    R_by_bra[R.blk[b][0]].push_back(b);
  }

  pt.mark("index");
  // ---- symbolic reference schedule: closure of the block set + FLOPs --------
  // Intermediate block sets are built as vectors, then sorted and
  // de-duplicated once per stage (the lexicographic order of std::array).
  auto sort_unique = [](auto &v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  // L blocks by bra sector a' (in block order), and the bra sectors in order
  std::map<int, std::vector<int64_t>> L_by_bra;
  for (size_t lb = 0; lb < L.blk.size(); ++lb) L_by_bra[L.blk[lb][0]].push_back((int64_t)lb);
  std::vector<int> bras;
  for (auto &kv : L_by_bra) bras.push_back(kv.first);
  // The pass runs per bra sector a' on host threads: a' is the first
  // coordinate of every intermediate and output block and no stage changes
  // it, so the per-a' sorted sets, joined in a' order, are the sorted global
  // set.  The FLOPs are summed per a' and then in a' order.
  auto symbolic = [&](const std::vector<Coords> &in, double *flops) {
    std::unordered_map<int, std::vector<int64_t>> in_by_ket;   // a -> entries of `in`
    for (size_t i = 0; i < in.size(); ++i) in_by_ket[in[i][0]].push_back((int64_t)i);
    struct Part {
      std::vector<Coords> out;
      double f1 = 0, f = 0, f4 = 0;
    };
    std::vector<Part> part(bras.size());
    parallel_for((int64_t)bras.size(), [&](int64_t bi) {
      Part &pr = part[(size_t)bi];
      // stage T1: (a', w, s..., b)
      std::vector<Coords> t1;
      for (int64_t lb : L_by_bra.at(bras[(size_t)bi])) {
        const Coords &l = L.blk[lb];
        auto it = in_by_ket.find(l[2]);
        if (it == in_by_ket.end()) continue;
        for (int64_t ii : it->second) {
          const Coords &c = in[(size_t)ii];
          int64_t nfree = 1;
          for (int i = 1; i < rk; ++i) nfree *= psi.ax[i].deg[c[i]];
          Coords t{};
          t[0] = l[0];
          t[1] = l[1];
          for (int i = 1; i < rk; ++i) t[1 + i] = c[i];
          t1.push_back(t);
          pr.f1 += fl * double(L.ax[0].deg[l[0]] * L.ax[1].deg[l[1]]) * double(L.ax[2].deg[l[2]]) * double(nfree);
        }
      }
      // MPO stages: carry (a', [s' done...], w_cur, [s remaining...], b)
      // represent as Coords: [a', t_0..t_{k-1}, w, s_k..s_{n-1}, b]
      sort_unique(t1);
      std::vector<Coords> cur;  // k = 0: [a', w, s0, (s1), b]
      cur.swap(t1);
      for (int k = 0; k < nphys; ++k) {
        std::vector<Coords> nxt;
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
            nxt.push_back(t);
            pr.f += fl * double(mfree) * double(kdim) * double(p.W[k].ax[2].deg[wc[2]] * p.W[k].ax[3].deg[wc[3]]);
          }
        }
        sort_unique(nxt);
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
          pr.out.push_back(o);
          pr.f4 += fl * double(m) * double(kd) * double(R.ax[0].deg[R.blk[rb][0]]);
        }
      }
      sort_unique(pr.out);
    });
    std::vector<Coords> out;
    double f = 0, f1 = 0, f4 = 0;
    for (Part &pr : part) {
      out.insert(out.end(), pr.out.begin(), pr.out.end());
      f1 += pr.f1;
      f += pr.f;
      f4 += pr.f4;
    }
    // [S1, S2 (+S3), S4] of the reference schedule
    flops[0] = f1;
    flops[1] = f;
    flops[2] = f4;
    return out;
  };

  // closure of psi's blocks under H_eff; the FLOPs are those of the pass
  // that finds the set closed
  std::vector<Coords> vset(psi.blk.begin(), psi.blk.end());
  sort_unique(vset);
  double f3[3] = {0, 0, 0};
  bool closed = false;
  for (int iter = 0; iter < 16 && !closed; ++iter) {
    const std::vector<Coords> out = symbolic(vset, f3);
    std::vector<Coords> merged;
    merged.reserve(vset.size() + out.size());
    std::set_union(vset.begin(), vset.end(), out.begin(), out.end(), std::back_inserter(merged));
    closed = merged.size() == vset.size();
    vset.swap(merged);
  }
  if (!closed) symbolic(vset, f3);
  p.algo_flops = f3[0] + f3[1] + f3[2];
  for (int i = 0; i < 3; ++i) p.st[i].algo_flops = f3[i];

  pt.mark("symbolic");
  // ---- vector layout ----------------------------------------------------------
  p.vec.rank = rk;
  p.vec.ax = psi.ax;
  p.vec.blk.assign(vset.begin(), vset.end());   // sorted: lexicographic order
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


  // ---- fused Krylov layout -----------------------------------------------------
  struct Fused {
    int64_t base = 0, N = 0;
  };
  std::map<int, Fused> fz;
  std::vector<int64_t> fcol(p.vec.blk.size(), 0);
  {
    int64_t base = 0;
    for (auto &kv : V_by_left) {
      const int a = kv.first;
      const int64_t Da = psi.ax[0].deg[a];
      int64_t N = 0;
      for (int64_t vb : kv.second) {
        fcol[vb] = N;
        N += p.vec.blk_elems(vb) / Da;
      }
      fz[a] = Fused{base, N};
      base += Da * N;
    }
    QLG_CHECK(base == p.vec.size, "fused layout size mismatch");
    for (size_t vb = 0; vb < p.vec.blk.size(); ++vb) {
      const int a = p.vec.blk[vb][0];
      const int64_t Da = psi.ax[0].deg[a], nb = p.vec.blk_elems(vb) / Da;
      p.pack_in.push_back(CopyDesc{p.vec.off[vb], fz[a].base + fcol[vb], nb, fz[a].N, (int32_t)Da, (int32_t)nb});
      p.pack_out.push_back(CopyDesc{fz[a].base + fcol[vb], p.vec.off[vb], fz[a].N, nb, (int32_t)Da, (int32_t)nb});
    }
  }

  pt.mark("layout");
  // ---- L stacks (ket-major) and stage 1 --------------------------------------------
  struct LStack {
    int64_t base = 0, rows = 0;
    std::map<int64_t, int64_t> roff;   // L block -> first row in the stack
  };
  std::map<int, LStack> lst;
  std::vector<CopyDesc> lcopy;
  int64_t lint_elems = 0;
  // Stacks, R panels and Y panels keep a leading dimension that is a multiple
  // of 8 (zero padding past K in every row): the split-BF16 kernels stage
  // their bf16 piece planes in aligned 16-byte chunks of 8 k values.
  auto ld8 = [](int64_t k) { return (k + 7) / 8 * 8; };
  int64_t lint_alg = 0, rint_alg = 0, y_alg = 0;   // elements without the padding
  for (auto &kv : L_by_ket) {
    const int a = kv.first;
    if (!fz.count(a)) continue;
    const int64_t Da = L.ax[2].deg[a], lda = ld8(Da);
    LStack s;
    s.base = lint_elems;
    for (int64_t lb : kv.second) {
      const int64_t r = L.ax[0].deg[L.blk[lb][0]] * L.ax[1].deg[L.blk[lb][1]];
      s.roff[lb] = s.rows;
      lcopy.push_back(CopyDesc{L.off[lb], s.base + s.rows * lda, Da, lda, (int32_t)r, (int32_t)Da});
      s.rows += r;
    }
    lint_elems += s.rows * lda;
    lint_alg += s.rows * Da;
    lst[a] = s;
  }
  Stage &s1 = p.st[0], &s2 = p.st[1], &s3 = p.st[2];
  s3.splitk = true;   // one long-K task per bra sector: split-K parts may be tuned in
  s3.split_on = std::getenv("QLG_SPLITK") != nullptr;
  s3.ed = ed;
  std::map<int, int64_t> tbase;
  {
    int64_t t1 = 0;
    for (auto &kv : lst) {
      const int a = kv.first;
      const LStack &s = kv.second;
      const int64_t Da = L.ax[2].deg[a], N = fz[a].N;
      tbase[a] = t1;
      GemmSeg g{};
      g.a_slot = SLOT_L;
      g.a_off = s.base;
      g.lda = ld8(Da);
      g.b_slot = SLOT_IN;
      g.b_off = fz[a].base;
      g.ldb = N;
      g.K = (int32_t)Da;
      GemmTask t{};
      t.c_slot = SLOT_X;
      t.c_off = t1;
      t.ldc = N;
      t.M = (int32_t)s.rows;
      t.N = (int32_t)N;
      t.seg_begin = (int32_t)s1.segs.size();
      t.seg_count = 1;
      s1.segs.push_back(g);
      s1.tasks.push_back(t);
      s1.flops += fl * double(s.rows) * double(N) * double(Da);
      t1 += s.rows * N;
    }
    p.x_elems = t1;
  }

  pt.mark("stage1");
  // ---- R panels (bra-major) ------------------------------------------------------------
  struct RPanel {
    int64_t base = 0, K = 0, ld = 0;
    std::map<std::pair<int, int>, int64_t> col;   // (z sector, b sector) -> column offset
  };
  std::map<int, RPanel> rpn;
  std::vector<CopyDesc> rcopy;
  int64_t rint_elems = 0;
  for (auto &kv : R_by_bra) {
    const int bp = kv.first;
    RPanel P;
    P.base = rint_elems;
    for (int64_t rb : kv.second) {
      P.col[{R.blk[rb][1], R.blk[rb][2]}] = P.K;
      P.K += R.ax[1].deg[R.blk[rb][1]] * R.ax[2].deg[R.blk[rb][2]];
    }
    const int64_t Dbp = R.ax[0].deg[bp];
    P.ld = ld8(P.K);
    for (int64_t rb : kv.second) {
      const int64_t w = R.ax[1].deg[R.blk[rb][1]] * R.ax[2].deg[R.blk[rb][2]];
      rcopy.push_back(CopyDesc{R.off[rb], P.base + P.col[{R.blk[rb][1], R.blk[rb][2]}], w, P.ld, (int32_t)Dbp,
                               (int32_t)w});
    }
    rint_elems += Dbp * P.ld;
    rint_alg += Dbp * P.K;
    rpn[bp] = P;
  }

  pt.mark("rpanels");
  // ---- Y panels (per bra sector) and stage 3 ---------------------------------------------
  struct YPanel {
    int64_t base = 0, rows = 0, K = 0, ld = 0;
  };
  std::map<int, YPanel> ypn;
  // first Y row of (vec block vb, components c0, c1):
  // yrow0[vb] + c0 * ystep0[vb] + c1 * ystep1[vb]
  std::vector<int64_t> yrow0(p.vec.blk.size(), -1), ystep0(p.vec.blk.size(), 0), ystep1(p.vec.blk.size(), 0);
  {
    std::map<int, std::vector<int64_t>> V_by_right;
    for (size_t vb = 0; vb < p.vec.blk.size(); ++vb) V_by_right[p.vec.blk[vb][bax]].push_back(vb);
    int64_t ybase = 0;
    for (auto &kv : V_by_right) {
      const int bp = kv.first;
      auto pit = rpn.find(bp);
      YPanel Y;
      Y.base = ybase;
      Y.K = pit == rpn.end() ? 0 : pit->second.K;
      Y.ld = ld8(Y.K);
      const int64_t Dbp = R.ax[0].deg[bp];
      const int32_t c_row0 = (int32_t)s3.crows.size();
      for (int64_t vb : kv.second) {
        const Coords &o = p.vec.blk[vb];
        const int ap = o[0];
        const int64_t Dap = L.ax[0].deg[ap];
        const int64_t d0 = p.W[0].ax[2].deg[o[1]], d1 = nphys == 2 ? p.W[1].ax[2].deg[o[2]] : 1;
        yrow0[vb] = Y.rows;
        ystep0[vb] = d1 * Dap;
        ystep1[vb] = Dap;
        for (int c0 = 0; c0 < d0; ++c0)
          for (int c1 = 0; c1 < d1; ++c1) {
            const int64_t col0 = fz[ap].base + fcol[vb] + (c0 * d1 + c1) * Dbp;
            for (int64_t r = 0; r < Dap; ++r) s3.crows.push_back(col0 + r * fz[ap].N);
            Y.rows += Dap;
          }
      }
      GemmTask t{};
      t.c_slot = SLOT_OUT;
      t.c_off = 0;
      t.ldc = Dbp;
      t.M = (int32_t)Y.rows;
      t.N = (int32_t)Dbp;
      t.c_row = c_row0;
      t.c_rowtab = 1;
      t.seg_begin = (int32_t)s3.segs.size();
      t.seg_count = 0;
      if (Y.K > 0) {
        GemmSeg g{};
        g.a_slot = SLOT_Y;
        g.a_off = Y.base;
        g.lda = Y.ld;
        g.b_slot = SLOT_R;
        g.b_off = pit->second.base;
        g.ldb = pit->second.ld;
        g.K = (int32_t)Y.K;
        g.flags = SEG_TRANS_B;
        s3.segs.push_back(g);
        t.seg_count = 1;
        s3.flops += fl * double(Y.rows) * double(Dbp) * double(Y.K);
      }
      s3.tasks.push_back(t);
      ybase += Y.rows * Y.ld;
      y_alg += Y.rows * Y.K;
      ypn[bp] = Y;
    }
    p.y_elems = ybase;
  }

  pt.mark("stage3");
  // ---- stage 2: channels (a'-sector, b-sector) -------------------------------------------
  std::vector<double> Wdn[2];
  std::vector<int64_t> wdims[2];
  for (int k = 0; k < nphys; ++k) {
    Wdn[k] = dense_copy(p.W[k], ctx->stream, ed);
    for (int i = 0; i < 4; ++i) wdims[k].push_back(p.W[k].ax[i].dim());
  }
  auto Widx = [&](int k, int64_t a, int64_t b, int64_t c, int64_t d) {
    return ((a * wdims[k][1] + b) * wdims[k][2] + c) * wdims[k][3] + d;
  };
  const auto so_w0 = sector_offsets(W0.ax[0]);
  std::vector<int64_t> so_s[2], so_t[2];
  for (int k = 0; k < nphys; ++k) {
    so_s[k] = sector_offsets(psi.ax[1 + k]);
    so_t[k] = sector_offsets(p.W[k].ax[2]);
  }
  const auto so_z = sector_offsets(W1.ax[3]);
  using Key7 = std::array<int32_t, 7>;   // (t0 sect, t0 comp, t1 sect, t1 comp, z sect, z comp, b' sect)
  struct InRow {
    RowDesc d;
    int64_t wg, sg0, sg1;
    int wsec, ssec0, ssec1;
  };
  struct Chan {
    int ap, b;
    std::vector<InRow> in;
    std::vector<Key7> out;
  };
  pt.mark("s2-mpo");
  std::map<std::pair<int, int>, int> cid;
  std::vector<Chan> ch;
  for (auto &kv : lst) {
    const int a = kv.first;
    const LStack &stk = kv.second;
    const int64_t Na = fz[a].N;
    for (auto &lr : stk.roff) {
      const int64_t lb = lr.first;
      const Coords &l = L.blk[lb];
      const int64_t Wm = L.ax[1].deg[l[1]];
      for (int64_t vb : V_by_left[a]) {
        const Coords &v = p.vec.blk[vb];
        const int b = v[bax];
        const int64_t Db = psi.ax[bax].deg[b];
        const int64_t d0 = psi.ax[1].deg[v[1]], d1 = nphys == 2 ? psi.ax[2].deg[v[2]] : 1;
        auto key = std::make_pair(l[0], b);
        auto it = cid.find(key);
        int c;
        if (it == cid.end()) {
          c = (int)ch.size();
          cid[key] = c;
          ch.push_back(Chan{l[0], b, {}, {}});
        } else {
          c = it->second;
        }
        for (int64_t w = 0; w < Wm; ++w)
          for (int c0 = 0; c0 < d0; ++c0)
            for (int c1 = 0; c1 < d1; ++c1) {
              InRow r;
              r.d.off = tbase[a] + (lr.second + w) * Na + fcol[vb] + (c0 * d1 + c1) * Db;
              r.d.rs = Wm * Na;
              r.wg = so_w0[l[1]] + w;
              r.sg0 = so_s[0][v[1]] + c0;
              r.sg1 = nphys == 2 ? so_s[1][v[2]] + c1 : 0;
              r.wsec = l[1];
              r.ssec0 = v[1];
              r.ssec1 = nphys == 2 ? v[2] : 0;
              ch[c].in.push_back(r);
            }
      }
    }
  }
  // Combined two-site operator, independent of the channel:
  // Wc[(w, s1, t0), (s2, t1, z)] = sum_y Wl[w, s1, t0, y] Wr[y, s2, t1, z],
  // formed once from the nonzero entries of the (sparse) MPO tensors.
  pt.mark("s2-rows");
  std::vector<double> Wc;
  int64_t wc_cols = 0;
  if (nphys == 2) {
    const int64_t nw = wdims[0][0], ns0 = wdims[0][1], nt0 = wdims[0][2], ny = wdims[0][3];
    const int64_t ns1 = wdims[1][1], nt1 = wdims[1][2], nz = wdims[1][3];
    wc_cols = ns1 * nt1 * nz;
    Wc.assign((size_t)nw * ns0 * nt0 * wc_cols * ed, 0.0);
    std::vector<std::vector<int64_t>> wr_nz(ny);   // nonzero (s2, t1, z) columns of Wr per y
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t col = 0; col < wc_cols; ++col) {
        const int64_t i1 = y * wc_cols + col;   // Wr is [y][s2][t1][z] row-major
        if (Wdn[1][i1 * ed] != 0.0 || (ed == 2 && Wdn[1][i1 * ed + 1] != 0.0)) wr_nz[y].push_back(col);
      }
    for (int64_t row = 0; row < nw * ns0 * nt0; ++row)
      for (int64_t y = 0; y < ny; ++y) {
        const int64_t i0 = row * ny + y;          // Wl is [w][s1][t0][y] row-major
        const double ar = Wdn[0][i0 * ed], ai = ed == 2 ? Wdn[0][i0 * ed + 1] : 0.0;
        if (ar == 0.0 && ai == 0.0) continue;
        for (int64_t col : wr_nz[y]) {
          const int64_t i1 = y * wc_cols + col, o = row * wc_cols + col;
          const double br = Wdn[1][i1 * ed], bi = ed == 2 ? Wdn[1][i1 * ed + 1] : 0.0;
          Wc[o * ed] += ar * br - ai * bi;
          if (ed == 2) Wc[o * ed + 1] += ar * bi + ai * br;
        }
      }
  }
  pt.mark("s2-wc");
  // Per channel, independently and on host threads: its outputs, their row
  // descriptors and its combined operator.  The lists are then joined in
  // channel order, so the plan is the same for any thread count.
  struct ChanBuilt {
    std::vector<RowDesc> orows;
    std::vector<double> coef;
  };
  std::vector<ChanBuilt> built(ch.size());
  parallel_for((int64_t)ch.size(), [&](int64_t ci) {
    Chan &c = ch[(size_t)ci];
    ChanBuilt &cb = built[(size_t)ci];
    // structurally reachable outputs that feed an existing R block and output
    // block; they depend on an input row's sectors only, not its components
    std::vector<std::array<int, 3>> trip;   // (w sector, s0 sector, s1 sector)
    trip.reserve(c.in.size());
    for (const InRow &r : c.in) trip.push_back({r.wsec, r.ssec0, r.ssec1});
    sort_unique(trip);
    std::vector<Key7> outs;
    for (const auto &tr : trip) {
      std::vector<std::pair<std::array<int, 2>, int>> fr{{{-1, -1}, tr[0]}};
      const int ss[2] = {tr[1], tr[2]};
      for (int k = 0; k < nphys; ++k) {
        std::vector<std::pair<std::array<int, 2>, int>> nf;
        for (auto &f : fr) {
          auto it = W_by_in[k].find({f.second, ss[k]});
          if (it == W_by_in[k].end()) continue;
          for (int64_t wb : it->second) {
            auto t = f.first;
            t[k] = p.W[k].blk[wb][2];
            nf.push_back({t, p.W[k].blk[wb][3]});
          }
        }
        fr.swap(nf);
      }
      for (auto &f : fr) {
        const int z = f.second;
        auto rit = R_by_zb.find({z, c.b});
        if (rit == R_by_zb.end()) continue;
        for (int64_t rb : rit->second) {
          const int bp = R.blk[rb][0];
          Coords o{};
          o[0] = c.ap;
          o[1] = f.first[0];
          if (nphys == 2) o[2] = f.first[1];
          o[bax] = bp;
          if (!p.vec.find.count(o)) continue;
          const int64_t dt0 = p.W[0].ax[2].deg[f.first[0]];
          const int64_t dt1 = nphys == 2 ? p.W[1].ax[2].deg[f.first[1]] : 1;
          const int64_t dz = W1.ax[3].deg[z];
          for (int a0 = 0; a0 < dt0; ++a0)
            for (int a1 = 0; a1 < dt1; ++a1)
              for (int zz = 0; zz < dz; ++zz)
                outs.push_back(Key7{f.first[0], a0, nphys == 2 ? f.first[1] : 0, a1, z, zz, bp});
        }
      }
    }
    sort_unique(outs);
    c.out.swap(outs);
    if (c.out.empty() || c.in.empty()) return;
    const int64_t Db = psi.ax[bax].deg[c.b];
    cb.orows.reserve(c.out.size());
    for (const Key7 &o : c.out) {
      const int bp = o[6];
      Coords ob{};
      ob[0] = c.ap;
      ob[1] = o[0];
      if (nphys == 2) ob[2] = o[2];
      ob[bax] = bp;
      const auto vit = p.vec.find.find(ob);
      const auto yit = ypn.find(bp);
      const auto rit = rpn.find(bp);
      QLG_CHECK(vit != p.vec.find.end() && yit != ypn.end() && rit != rpn.end(), "planner: channel output lost");
      const int64_t vb = vit->second;
      const auto cit = rit->second.col.find({o[4], c.b});
      QLG_CHECK(yrow0[vb] >= 0 && cit != rit->second.col.end(), "planner: channel output lost");
      const YPanel &Y = yit->second;
      const int64_t row0 = yrow0[vb] + (int64_t)o[1] * ystep0[vb] + (int64_t)o[3] * ystep1[vb];
      RowDesc d;
      d.off = Y.base + row0 * Y.ld + cit->second + (int64_t)o[5] * Db;
      d.rs = Y.ld;
      cb.orows.push_back(d);
    }
    // combined operator rows (outputs) x columns (inputs)
    const int64_t K = (int64_t)c.in.size();
    cb.coef.assign((size_t)((int64_t)c.out.size() * K * ed), 0.0);
    for (size_t r = 0; r < c.out.size(); ++r) {
      const Key7 &o = c.out[r];
      const int64_t tg0 = so_t[0][o[0]] + o[1];
      const int64_t tg1 = nphys == 2 ? so_t[1][o[2]] + o[3] : 0;
      const int64_t zg = so_z[o[4]] + o[5];
      for (int64_t q = 0; q < K; ++q) {
        const InRow &in = c.in[q];
        double vr = 0, vi = 0;
        if (nphys == 1) {
          const int64_t i = Widx(0, in.wg, in.sg0, tg0, zg);
          vr = Wdn[0][i * ed];
          if (ed == 2) vi = Wdn[0][i * ed + 1];
        } else {
          const int64_t row = (in.wg * wdims[0][1] + in.sg0) * wdims[0][2] + tg0;
          const int64_t col = (in.sg1 * wdims[1][2] + tg1) * wdims[1][3] + zg;
          const int64_t i = row * wc_cols + col;
          vr = Wc[i * ed];
          if (ed == 2) vi = Wc[i * ed + 1];
        }
        const int64_t e = (int64_t)r * K + q;
        cb.coef[e * ed] = vr;
        if (ed == 2) cb.coef[e * ed + 1] = vi;
      }
    }
  });
  std::vector<double> Ch;
  int64_t coff = 0;
  for (size_t ci = 0; ci < ch.size(); ++ci) {
    const Chan &c = ch[ci];
    if (c.out.empty() || c.in.empty()) continue;
    const int64_t Dap = L.ax[0].deg[c.ap], Db = psi.ax[bax].deg[c.b];
    ChanTask t{};
    t.c_off = coff;
    t.kin = (int32_t)c.in.size();
    t.kout = (int32_t)c.out.size();
    t.P = (int32_t)(Dap * Db);
    t.Db = (int32_t)Db;
    t.in_rows = (int32_t)s2.rows.size();
    for (const InRow &r : c.in) s2.rows.push_back(r.d);
    t.out_rows = (int32_t)s2.rows.size();
    s2.rows.insert(s2.rows.end(), built[ci].orows.begin(), built[ci].orows.end());
    Ch.insert(Ch.end(), built[ci].coef.begin(), built[ci].coef.end());
    coff += (int64_t)c.out.size() * (int64_t)c.in.size();
    s2.chan.push_back(t);
    s2.flops += fl * double(t.kout) * double(t.kin) * double(t.P);
  }
  p.c_elems = coff;

  pt.mark("stage2");
  // ---- bounds of every task, then tiles and uploads --------------------------------
  {
    int64_t slot_elems[NSLOTS] = {0};
    slot_elems[SLOT_IN] = slot_elems[SLOT_OUT] = p.vec.size;
    slot_elems[SLOT_L] = lint_elems;
    slot_elems[SLOT_R] = rint_elems;
    slot_elems[SLOT_X] = p.x_elems;
    slot_elems[SLOT_Y] = p.y_elems;
    slot_elems[SLOT_C] = p.c_elems;
    for (int i = 0; i < 3; ++i) validate_stage(p.st[i], slot_elems);
    // the MPO stage gathers from the whole first intermediate through 32-bit
    // byte offsets, and reads each channel operator the same way
    if (p.x_elems * ed * 8 >= (int64_t(1) << 31))
      throw Error(QLG_ERR_UNSUPPORTED, "first H_eff intermediate exceeds 2^31 bytes");
    auto check_copies = [](const std::vector<CopyDesc> &v, int64_t src_n, int64_t dst_n, const char *what) {
      for (const CopyDesc &c : v)
        QLG_CHECK(c.rows > 0 && c.cols > 0 && c.src >= 0 && c.dst >= 0 && c.sld >= c.cols && c.dld >= c.cols &&
                      c.src + (int64_t)(c.rows - 1) * c.sld + c.cols <= src_n &&
                      c.dst + (int64_t)(c.rows - 1) * c.dld + c.cols <= dst_n,
                  std::string("copy list out of bounds: ") + what);
    };
    auto extent = [](const Ten &t) {
      int64_t e = 0;
      for (size_t b = 0; b < t.blk.size(); ++b) e = std::max(e, t.off[b] + t.blk_elems(b));
      return e;
    };
    check_copies(p.pack_in, p.vec.size, p.vec.size, "pack");
    check_copies(p.pack_out, p.vec.size, p.vec.size, "unpack");
    check_copies(lcopy, extent(L), lint_elems, "L repack");
    check_copies(rcopy, extent(R), rint_elems, "R repack");
  }
  const char *pin = std::getenv("QLG_TILES");
  int pinned[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  if (pin) std::sscanf(pin, "%dx%d,%dx%d,%dx%d", &pinned[0][0], &pinned[0][1], &pinned[1][0], &pinned[1][1],
                       &pinned[2][0], &pinned[2][1]);
  for (int i = 0; i < 3; ++i) {
    if (pinned[i][0] > 0) {
      p.st[i].make_tiles(pinned[i][0], pinned[i][1]);
      p.tiles_pinned = true;
    } else if (i == 1) {
      // MPO stage: HBM-bound.  X is re-read once per row tile, so the row
      // tile covers K_out where the MFMA padding stays cheap; the cost counts
      // X reads per row tile plus padded MFMA work at ~10 flops per byte
      double best = 1e300;
      int bm_best = 32;
      for (int bm : {16, 32, 64}) {
        double cost = 0;
        for (const ChanTask &c : s2.chan) {
          const double rt = (c.kout + bm - 1) / bm;
          cost += rt * c.kin * double(c.P) * 8.0 + rt * bm * double(c.kin) * c.P * 2.0 / 10.0;
        }
        if (cost < best) {
          best = cost;
          bm_best = bm;
        }
      }
      p.st[i].make_tiles(bm_best, bm_best == 64 ? 64 : 128);
    } else {
      model_tiles(p.st[i], p.dtype == QLG_C128);   // replaced by measured shapes on the first matvec
    }
    p.st[i].upload(ctx->stream);
  }
  pt.mark("tiles");
  p.d_pack_in.ensure(std::max<size_t>(1, p.pack_in.size()) * sizeof(CopyDesc));
  p.d_pack_out.ensure(std::max<size_t>(1, p.pack_out.size()) * sizeof(CopyDesc));
  if (!p.pack_in.empty()) {
    QLG_HIP(hipMemcpyAsync(p.d_pack_in.p, p.pack_in.data(), p.pack_in.size() * sizeof(CopyDesc),
                           hipMemcpyHostToDevice, ctx->stream));
    QLG_HIP(hipMemcpyAsync(p.d_pack_out.p, p.pack_out.data(), p.pack_out.size() * sizeof(CopyDesc),
                           hipMemcpyHostToDevice, ctx->stream));
  }
  const size_t eb = ed * sizeof(double);
  p.X.ensure(std::max<int64_t>(p.x_elems, 1) * eb);
  p.Y.ensure(std::max<int64_t>(p.y_elems, 1) * eb);
  p.C.ensure(std::max<int64_t>(p.c_elems, 1) * eb);
  p.Lint.ensure(std::max<int64_t>(lint_elems, 1) * eb);
  p.Rint.ensure(std::max<int64_t>(rint_elems, 1) * eb);
  p.lint_elems = lint_elems;
  p.rint_elems = rint_elems;
  // padding columns of the stacks and panels stay zero (the split-BF16
  // kernels read them)
  QLG_HIP(hipMemsetAsync(p.Lint.p, 0, std::max<int64_t>(lint_elems, 1) * eb, ctx->stream));
  QLG_HIP(hipMemsetAsync(p.Rint.p, 0, std::max<int64_t>(rint_elems, 1) * eb, ctx->stream));
  pt.mark("alloc");
  // Y entries no channel writes stay zero for the lifetime of the plan
  QLG_HIP(hipMemsetAsync(p.Y.p, 0, std::max<int64_t>(p.y_elems, 1) * eb, ctx->stream));
  if (p.c_elems > 0)
    QLG_HIP(hipMemcpyAsync(p.C.p, Ch.data(), p.c_elems * eb, hipMemcpyHostToDevice, ctx->stream));
  {
    DevBuf dl, dr;
    dl.ensure(std::max<size_t>(1, lcopy.size()) * sizeof(CopyDesc));
    dr.ensure(std::max<size_t>(1, rcopy.size()) * sizeof(CopyDesc));
    if (!lcopy.empty())
      QLG_HIP(hipMemcpyAsync(dl.p, lcopy.data(), lcopy.size() * sizeof(CopyDesc), hipMemcpyHostToDevice, ctx->stream));
    if (!rcopy.empty())
      QLG_HIP(hipMemcpyAsync(dr.p, rcopy.data(), rcopy.size() * sizeof(CopyDesc), hipMemcpyHostToDevice, ctx->stream));
    launch_block_copy(ed == 2, dl.as<CopyDesc>(), (int)lcopy.size(), static_cast<const double *>(L.data),
                      p.Lint.as<double>(), ctx->stream);
    launch_block_copy(ed == 2, dr.as<CopyDesc>(), (int)rcopy.size(), static_cast<const double *>(R.data),
                      p.Rint.as<double>(), ctx->stream);
    QLG_HIP(hipStreamSynchronize(ctx->stream));
  }
  pt.mark("repack+sync");
  {
    // algorithmic HBM bytes per launch: operands read once, results written once
    const double e8 = 8.0 * ed;
    p.st[0].algo_bytes = e8 * double(lint_alg + p.vec.size + p.x_elems);
    p.st[1].algo_bytes = e8 * double(p.x_elems + p.c_elems + y_alg);
    p.st[2].algo_bytes = e8 * double(y_alg + rint_alg + p.vec.size);
  }
  if (const char *v = std::getenv("QLG_VERBOSE"); v && v[0] == '1') {
    const char *names[3] = {"L.psi", "MPO", ".R^T"};
    for (int i = 0; i < 3; ++i) {
      const Stage &s = p.st[i];
      std::fprintf(stderr, "[qlg plan] stage %d (%s): tasks %zu tiles %zu tile %dx%d gflop %.3f\n", i + 1, names[i],
                   i == 1 ? s.chan.size() : s.tasks.size(), s.tiles.size(), s.bm, s.bn, s.flops / 1e9);
    }
    for (int i : {0, 2}) {
      // tile efficiency (useful / padded MFMA work) and the task shape spread
      const Stage &s = p.st[i];
      double useful = 0, padded = 0;
      std::vector<int> ms, ns, ks;
      for (const GemmTask &k : s.tasks) {
        int64_t K = 0, Kp = 0;
        for (int q = 0; q < k.seg_count; ++q) {
          K += s.segs[k.seg_begin + q].K;
          Kp += (s.segs[k.seg_begin + q].K + 15) / 16 * 16;
        }
        useful += 2.0 * k.M * k.N * K;
        padded += 2.0 * double((k.M + s.bm - 1) / s.bm * s.bm) * double((k.N + s.bn - 1) / s.bn * s.bn) * Kp;
        ms.push_back(k.M);
        ns.push_back(k.N);
        ks.push_back((int)K);
      }
      auto pct = [](std::vector<int> v, double q) {
        if (v.empty()) return 0;
        std::sort(v.begin(), v.end());
        return v[std::min(v.size() - 1, size_t(q * v.size()))];
      };
      std::fprintf(stderr, "[qlg plan] stage %d tile efficiency %.3f; M p10/50/90 %d/%d/%d, N %d/%d/%d, K %d/%d/%d\n",
                   i + 1, padded > 0 ? useful / padded : 0.0, pct(ms, .1), pct(ms, .5), pct(ms, .9), pct(ns, .1),
                   pct(ns, .5), pct(ns, .9), pct(ks, .1), pct(ks, .5), pct(ks, .9));
      if (s.split())
        std::fprintf(stderr, "[qlg plan] stage %d split-K: chunk %lld, %zu tasks split into %zu launch tasks, %zu tiles, "
                     "workspace %lld elements\n", i + 1, (long long)s.kchunk, s.red.size(), s.ktasks.size(),
                     s.tiles.size(), (long long)s.work_elems);
    }
    std::fprintf(stderr, "[qlg plan] vector %lld, T1 %lld, Y %lld, C %lld, L stacks %lld, R panels %lld, algo gflop %.3f\n",
                 (long long)p.vec.size, (long long)p.x_elems, (long long)p.y_elems, (long long)p.c_elems,
                 (long long)lint_elems, (long long)rint_elems, p.algo_flops / 1e9);
    int kin_max = 0, kout_max = 0, p_max = 0, nbig = 0;
    double xreads = 0, kin_avg = 0, kout_avg = 0, wsum = 0;
    for (const ChanTask &c : s2.chan) {
      kin_max = std::max(kin_max, c.kin);
      kout_max = std::max(kout_max, c.kout);
      p_max = std::max(p_max, c.P);
      nbig += c.kin > 256;
      xreads += double((c.kout + s2.bm - 1) / s2.bm) * c.kin * c.P;
      kin_avg += double(c.kin) * c.P;
      kout_avg += double(c.kout) * c.P;
      wsum += c.P;
    }
    std::fprintf(stderr,
                 "[qlg plan] MPO channels: kin max %d avg %.1f, kout max %d avg %.1f, P max %d avg %.0f, "
                 "kin>256: %d, X element reads %.0f (%.2fx T1)\n",
                 kin_max, kin_avg / std::max(wsum, 1.0), kout_max, kout_avg / std::max(wsum, 1.0), p_max,
                 wsum / std::max<size_t>(s2.chan.size(), 1), nbig, xreads, xreads / std::max<double>(p.x_elems, 1));
  }
  return P.release();
}

// ---------------------------------------------------------------------------
Plan::~Plan() {
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  drop_graphs();
  if (host_seg) (void)hipHostFree(host_seg);
  if (host_lz) (void)hipHostFree(host_lz);
}

void Plan::drop_graphs() {
  for (auto &g : step_graphs)
    if (g) (void)hipGraphExecDestroy(g);
  step_graphs.clear();
  for (auto &g : chunk_graphs)
    if (g) (void)hipGraphExecDestroy(g);
  chunk_graphs.clear();
  chunk_key.clear();
}

void Plan::to_fused(const double *canon, double *fused) {
  launch_block_copy(dtype == QLG_C128, d_pack_in.as<CopyDesc>(), (int)pack_in.size(), canon, fused, ctx->stream);
}

void Plan::to_canonical(const double *fused, double *canon) {
  launch_block_copy(dtype == QLG_C128, d_pack_out.as<CopyDesc>(), (int)pack_out.size(), fused, canon, ctx->stream);
}

bool Plan::planar() const {
  return dtype == QLG_F64 && (precision == QLG_PREC_BF16X3 || precision == QLG_PREC_BF16X6);
}

void Plan::prepare_reduced() {
  if (planar()) {
    // real split-BF16: L stacks and R panels split once into three bf16 piece
    // planes; Y kept as planes written by the MPO stage (padding and entries
    // no channel writes stay zero); the channel operators in FP32
    if (planar_ready) return;
    Lp.ensure(std::max<int64_t>(3 * lint_elems, 1) * 2);
    Rp.ensure(std::max<int64_t>(3 * rint_elems, 1) * 2);
    Yp.ensure(std::max<int64_t>(3 * y_elems, 1) * 2);
    C32.ensure(std::max<int64_t>(c_elems, 1) * sizeof(float));
    QLG_HIP(hipMemsetAsync(Yp.p, 0, std::max<int64_t>(3 * y_elems, 1) * 2, ctx->stream));
    launch_split_planes(Lint.as<double>(), Lp.p, lint_elems, ctx->stream);
    launch_split_planes(Rint.as<double>(), Rp.p, rint_elems, ctx->stream);
    launch_f64_to_f32(C.as<double>(), C32.as<float>(), c_elems, ctx->stream);
    planar_ready = true;
    return;
  }
  if (reduced_ready) return;
  const int64_t ed = elem_doubles(dtype);   // complex data: (re, im) pairs of floats
  Lint32.ensure(std::max<int64_t>(lint_elems * ed, 1) * sizeof(float));
  Rint32.ensure(std::max<int64_t>(rint_elems * ed, 1) * sizeof(float));
  C32.ensure(std::max<int64_t>(c_elems * ed, 1) * sizeof(float));
  Y32.ensure(std::max<int64_t>(y_elems * ed, 1) * sizeof(float));
  QLG_HIP(hipMemsetAsync(Y32.p, 0, std::max<int64_t>(y_elems * ed, 1) * sizeof(float), ctx->stream));
  launch_f64_to_f32(Lint.as<double>(), Lint32.as<float>(), lint_elems * ed, ctx->stream);
  launch_f64_to_f32(Rint.as<double>(), Rint32.as<float>(), rint_elems * ed, ctx->stream);
  launch_f64_to_f32(C.as<double>(), C32.as<float>(), c_elems * ed, ctx->stream);
  reduced_ready = true;
}

Slots Plan::slots(const double *in, double *out) {
  const bool reduced = precision != QLG_PREC_F64;
  Slots s{};
  s.p[SLOT_IN] = const_cast<double *>(in);
  s.p[SLOT_OUT] = out;
  s.p[SLOT_L] = reduced ? reinterpret_cast<double *>(Lint32.p) : Lint.as<double>();
  s.p[SLOT_R] = reduced ? reinterpret_cast<double *>(Rint32.p) : Rint.as<double>();
  s.p[SLOT_X] = X.as<double>();
  s.p[SLOT_Y] = reduced ? reinterpret_cast<double *>(Y32.p) : Y.as<double>();
  s.p[SLOT_C] = reduced ? reinterpret_cast<double *>(C32.p) : C.as<double>();
  if (planar()) {
    s.p[SLOT_L] = static_cast<double *>(Lp.p);
    s.p[SLOT_R] = static_cast<double *>(Rp.p);
    s.p[SLOT_Y] = static_cast<double *>(Yp.p);
    s.ps[SLOT_L] = lint_elems;
    s.ps[SLOT_R] = rint_elems;
    s.ps[SLOT_Y] = y_elems;
  }
  s.p[SLOT_W] = st[2].work.as<double>();
  s.crows = st[2].d_crows.as<int64_t>();
  s.skip = skip_flag;
  return s;
}

void Plan::launch_stage(int i, const Slots &s) {
  const bool reduced = precision != QLG_PREC_F64, cplx = dtype == QLG_C128;
  hipStream_t stream = ctx->stream;
  Stage &g = st[i];
  const int out_planes = planar() ? (precision == QLG_PREC_BF16X6 ? 3 : 2) : 0;
  if (i == 1) {
    launch_mpo_stage(g.bm, g.bn, cplx, reduced, out_planes, y_elems, g.d_ctiles.as<ChanTile>(),
                     g.d_rows.as<RowDesc>(), s.p[SLOT_C], X.p, s.p[SLOT_Y], s.skip, (int)g.ctiles.size(), stream);
    return;
  }
  // the parts write their panels through SLOT_W: it must name this stage's workspace as uploaded
  QLG_CHECK(!g.split() || (s.p[SLOT_W] == g.work.p && g.work.bytes >= (size_t)g.work_elems * g.ed * sizeof(double)),
            "split-K workspace does not match the stage's task list");
  if (out_planes) {
    launch_grouped_gemm_planar(precision, g.bm, g.bn, i == 2, g.d_tiles.as<GemmTile>(), g.d_tasks.as<GemmTask>(),
                               g.d_segs.as<GemmSeg>(), s, (int)g.tiles.size(), stream);
  } else if (reduced) {
    // stage 1: A = L stacks, B = the Krylov vector (FP64), C = T1;
    // stage 3: A = Y, B = R panels (transposed), C = the Krylov output (FP64, row table)
    launch_grouped_gemm_mixed(precision, g.bm, g.bn, i == 0, i == 2, cplx, g.d_tiles.as<GemmTile>(),
                              g.d_tasks.as<GemmTask>(), g.d_segs.as<GemmSeg>(), s, (int)g.tiles.size(), stream);
  } else {
    launch_grouped_gemm_f64(g.bm, g.bn, cplx, g.d_tiles.as<GemmTile>(), g.d_tasks.as<GemmTask>(),
                            g.d_segs.as<GemmSeg>(), s, (int)g.tiles.size(), stream);
  }
  // split-K parts wrote FP64 partial panels: sum them into the output
  if (g.split())
    launch_splitk_reduce(cplx, g.d_red.as<RedTask>(), g.d_redc.as<RedChunk>(), (int)g.redc.size(), s, stream);
}

namespace {
std::mutex g_tile_mu;
struct TileChoice {
  int bm = 0, bn = 0;
  bool split = false;   // split-K parts (stage 3)
  bool operator==(const TileChoice &o) const { return bm == o.bm && bn == o.bn && split == o.split; }
};
std::unordered_map<std::string, TileChoice> g_tile_cache;   // signature -> tile shape and split-K
int g_tune_hits = 0, g_tune_misses = 0;

bool autotune_enabled() {
  static const bool on = [] {
    const char *v = std::getenv("QLG_AUTOTUNE");
    return !(v && v[0] == '0');
  }();
  return on;
}
}  // namespace

// Tile autotuner of the two grouped-GEMM stages.  Every tile shape (nine for
// real FP64, seven for the real reduced modes, four for complex; the cost
// model's ranking only orders them) is timed on
// this plan's own task lists and buffers (one warm-up and one timed launch each, HIP events
// on the plan's stream); the fastest is kept.  Results are cached process-wide
// under the stage's size-histogram signature, so sites with the same sector
// profile reuse them.  The launches write only the plan's intermediates and
// the output vector, which the matvec that follows overwrites.
void Plan::autotune(const double *in, double *out) {
  if (tuned_precision == precision) return;
  tuned_precision = precision;
  if (tiles_pinned || !autotune_enabled() || kind != PLAN_HEFF) return;
  const bool cplx = dtype == QLG_C128;
  const int variant = precision * 2 + (cplx ? 1 : 0);
  hipStream_t stream = ctx->stream;
  const bool verbose = [] {
    const char *v = std::getenv("QLG_VERBOSE");
    return v && v[0] == '1';
  }();
  // choices cached for this precision's shape signatures (bm = 0: not tuned yet)
  TileChoice cached[3];
  bool unchanged = step_graphs.empty() && chunk_graphs.empty();
  {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    for (int i : {0, 2}) {
      const Stage &g = st[i];
      if (g.tasks.empty()) continue;
      auto it = g_tile_cache.find(tile_signature(g, i, variant));
      if (it != g_tile_cache.end()) {
        cached[i] = it->second;
        ++g_tune_hits;
      }
      if (!(cached[i] == TileChoice{g.bm, g.bn, g.split_on}) || g.tiles.empty()) unchanged = false;
    }
  }
  // both stages already hold the tiles chosen for this precision: nothing to
  // time or re-upload, and the stream need not drain
  if (unchanged) return;
  // slot pointers are taken again after every upload: a tile list with
  // split-K parts (re)allocates the stage's workspace, which SLOT_W names
  Slots s = slots(in, out);
  // tile lists are re-uploaded (and their buffers may grow) below: nothing
  // queued earlier on the stream may still read them
  QLG_HIP(hipStreamSynchronize(stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  QLG_HIP(hipEventCreate(&e0));
  QLG_HIP(hipEventCreate(&e1));
  try {
    for (int i : {0, 2}) {
      Stage &g = st[i];
      if (g.tasks.empty()) continue;
      TileChoice chosen = cached[i];
      if (chosen.bm == 0) {
        if (i == 2) {   // fill T1 and Y with this matvec's data before stage 3 is timed
          launch_stage(0, s);
          launch_stage(1, s);
        }
        const auto rank = rank_tile_shapes(g, cplx, precision == QLG_PREC_F64 && !cplx);
        double best_ms = 1e300;
        auto time_choice = [&](const TileChoice &tc, double model_ms) {
          g.split_on = tc.split;
          g.make_tiles(tc.bm, tc.bn);
          g.upload(stream);
          s = slots(in, out);
          double ms_min = 1e300;
          for (int rep = 0; rep < 3; ++rep) {   // one warm-up, best of two timed launches
            QLG_HIP(hipEventRecord(e0, stream));
            launch_stage(i, s);
            QLG_HIP(hipEventRecord(e1, stream));
            QLG_HIP(hipEventSynchronize(e1));
            float ms = 0;
            QLG_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) ms_min = std::min(ms_min, (double)ms);
          }
          if (verbose)
            std::fprintf(stderr, "[qlg tune] stage %d %dx%d: %.4f ms (model %.4f ms, split-K chunk %lld)\n", i + 1,
                         tc.bm, tc.bn, ms_min, model_ms, (long long)g.kchunk);
          if (ms_min < best_ms) {
            best_ms = ms_min;
            chosen = tc;
          }
          return ms_min;
        };
        // every shape unsplit (or, with QLG_SPLITK set, split as it says);
        // then the two fastest shapes with split-K parts
        const bool forced = std::getenv("QLG_SPLITK") != nullptr;
        std::vector<std::pair<double, TileChoice>> timed;
        for (size_t c = 0; c < rank.size(); ++c) {
          const auto sh = tile_shape(rank[c].second);
          const TileChoice tc{sh.first, sh.second, forced && g.splitk};
          timed.push_back({time_choice(tc, rank[c].first * 1e3), tc});
        }
        if (g.splitk && !forced) {
          std::stable_sort(timed.begin(), timed.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
          for (size_t c = 0; c < std::min<size_t>(2, timed.size()); ++c)
            time_choice(TileChoice{timed[c].second.bm, timed[c].second.bn, true}, 0.0);
        }
        std::lock_guard<std::mutex> lk(g_tile_mu);
        g_tile_cache[tile_signature(g, i, variant)] = chosen;
        ++g_tune_misses;
      }
      if (!(chosen == TileChoice{g.bm, g.bn, g.split_on}) || g.tiles.empty()) {
        g.split_on = chosen.split;
        g.make_tiles(chosen.bm, chosen.bn);
        g.upload(stream);
        s = slots(in, out);
      }
    }
    drop_graphs();   // captured launches name the old tile lists
    QLG_HIP(hipStreamSynchronize(stream));
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (verbose)
    std::fprintf(stderr, "[qlg tune] tiles %dx%d / %dx%d%s (cache hits %d, misses %d)\n", st[0].bm, st[0].bn,
                 st[2].bm, st[2].bn, st[2].split() ? " split-K" : "", g_tune_hits, g_tune_misses);
  const char *vv = std::getenv("QLG_VERBOSE");
  if (vv && vv[0] == '3') {
    // task shapes of the two GEMM stages: FLOP-weighted quantiles of M, N, K
    for (int i : {0, 2}) {
      const Stage &g = st[i];
      std::vector<std::array<double, 4>> v;   // flops, M, N, K
      double tot = 0;
      for (const GemmTask &k : g.tasks) {
        int64_t K = 0;
        for (int q = 0; q < k.seg_count; ++q) K += g.segs[k.seg_begin + q].K;
        const double f = 2.0 * k.M * k.N * K;
        v.push_back({f, double(k.M), double(k.N), double(K)});
        tot += f;
      }
      std::string line;
      for (int d = 1; d <= 3; ++d) {
        std::sort(v.begin(), v.end(), [d](const auto &a, const auto &b) { return a[d] < b[d]; });
        double acc = 0;
        int qi = 0;
        const double qs[3] = {0.1, 0.5, 0.9};
        line += std::string(" ") + "MNK"[d - 1] + ":";
        for (const auto &e : v) {
          acc += e[0];
          while (qi < 3 && acc >= qs[qi] * tot) {
            line += " " + std::to_string((int64_t)e[d]);
            ++qi;
          }
        }
      }
      std::fprintf(stderr, "[qlg shapes] stage %d: %zu tasks, %zu tiles of %dx%d, %.3f GFLOP;%s (FLOP-weighted 10/50/90%%)\n",
                   i + 1, g.tasks.size(), g.tiles.size(), g.bm, g.bn, tot * 1e-9, line.c_str());
    }
  }
}

void Plan::matvec_fused(const double *in, double *out) {
  const bool reduced = precision != QLG_PREC_F64;
  if (reduced) prepare_reduced();
  autotune(in, out);
  const Slots s = slots(in, out);
  hipStream_t stream = ctx->stream;
  const bool timed = timing && !capturing;
  if (timed) {
    for (auto &e : ev)
      if (!e) QLG_HIP(hipEventCreate(&e));
    QLG_HIP(hipEventRecord(ev[0], stream));
  }
  for (int i = 0; i < 3; ++i) {
    launch_stage(i, s);
    if (timed) QLG_HIP(hipEventRecord(ev[i + 1], stream));
  }
}

void Plan::matvec(const double *in, double *out) {
  const size_t bytes = std::max<int64_t>(vec.size, 1) * elem_doubles(dtype) * sizeof(double);
  vtmp_in.ensure(bytes);
  vtmp_out.ensure(bytes);
  to_fused(in, vtmp_in.as<double>());
  matvec_fused(vtmp_in.as<double>(), vtmp_out.as<double>());
  to_canonical(vtmp_out.as<double>(), out);
}

}  // namespace qlg
