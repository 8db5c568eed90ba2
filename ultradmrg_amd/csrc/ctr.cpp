// Planners for the contractions around the Krylov solve, on the same
// grouped-GEMM stages as the effective Hamiltonian:
//
//  * tensordot: C = A x B over A's last k axes and B's first k axes (result
//    axes: A's free axes then B's, as the reference's Contract(A, B, {...})
//    returns them).  Used for psi0 = A_l . A_r
//    (two_site_update_finite_vmps_impl.h:278-279) and for U.S / S.V^T
//    (:322-331).
//  * left environment  L'[b*, w1, b] = A*[a', s', b'] L[a', w, a] W[w, s, s', w1] A[a, s, b]
//    (UpdateSiteLenvs, include/qlmps/algorithm/vmps/vmps_init.h:359-373)
//  * right environment R'[a*, w0, a] = B*[a', s', b'] R[b', w, b] W[w0, s, s', w] B[a, s, b]
//    (UpdateSiteRenvs, vmps_init.h:343-357)
//  * transpose: the axes of a tensor in a new order, result axis i = source
//    axis perm[i] (qlten's Transpose, two_site_update_finite_tdvp_impl.h:36)
//
// Environment updates: stage 1 contracts the environment with the MPS tensor
// into per-channel chunks, stage 2 applies the MPO tensor as a small dense
// matrix per channel, stage 3 contracts with the conjugated MPS tensor.
#include <algorithm>
#include <map>
#include <memory>
#include <set>

#include "plan.h"
#include "transpose.h"

namespace qlg {

namespace {

int64_t extent(const Ten &t) {
  int64_t e = 0;
  for (size_t b = 0; b < t.blk.size(); ++b) e = std::max(e, t.off[b] + t.blk_elems(b));
  return e;
}

void finish_layout(Layout &L, const std::set<Coords> &blocks) {
  L.blk.assign(blocks.begin(), blocks.end());
  L.off.resize(L.blk.size());
  int64_t o = 0;
  for (size_t b = 0; b < L.blk.size(); ++b) {
    L.off[b] = o;
    L.find[L.blk[b]] = (int64_t)b;
    o += L.blk_elems(b);
  }
  L.size = o;
}

// dense host copy of an MPO tensor (index order kept)
std::vector<double> mpo_dense(const Ten &t, hipStream_t st, int ed, std::vector<int64_t> &dims) {
  dims.assign(4, 0);
  std::vector<std::vector<int64_t>> so(4);
  int64_t tot = 1;
  for (int i = 0; i < 4; ++i) {
    dims[i] = t.ax[i].dim();
    so[i] = sector_offsets(t.ax[i]);
    tot *= dims[i];
  }
  const int64_t raw = extent(t);
  std::vector<double> h(raw * ed);
  if (raw > 0) {
    QLG_HIP(hipMemcpyAsync(h.data(), t.data, raw * ed * sizeof(double), hipMemcpyDeviceToHost, st));
    QLG_HIP(hipStreamSynchronize(st));
  }
  std::vector<double> d(tot * ed, 0.0);
  for (size_t b = 0; b < t.blk.size(); ++b) {
    int64_t bd[4], n = 1;
    for (int i = 0; i < 4; ++i) { bd[i] = t.ax[i].deg[t.blk[b][i]]; n *= bd[i]; }
    for (int64_t e = 0; e < n; ++e) {
      int64_t r = e, loc[4];
      for (int i = 3; i >= 0; --i) { loc[i] = r % bd[i]; r /= bd[i]; }
      int64_t gi = 0;
      for (int i = 0; i < 4; ++i) gi = gi * dims[i] + so[i][t.blk[b][i]] + loc[i];
      for (int c = 0; c < ed; ++c) d[gi * ed + c] = h[(t.off[b] + e) * ed + c];
    }
  }
  return d;
}

void finalize(Plan &p, int nst, const int64_t *slot_elems, bool cplx) {
  for (int i = 0; i < nst; ++i) validate_stage(p.st[i], slot_elems);
  for (int i = 0; i < nst; ++i) {
    model_tiles(p.st[i], cplx);
    p.st[i].upload(p.ctx->stream);
  }
  p.nstages = nst;
}

}  // namespace

// ---------------------------------------------------------------------------
Plan *build_tensordot_plan(Ctx *ctx, const qlg_bsten *Ad, const qlg_bsten *Bd, int k, int conj_a, int conj_b) {
  QLG_CHECK(Ad && Bd, "null descriptor");
  QLG_CHECK(k >= 1 && k <= Ad->rank && k <= Bd->rank, "tensordot: bad number of contracted axes");
  std::unique_ptr<Plan> P(new Plan());
  Plan &p = *P;
  p.ctx = ctx;
  p.kind = PLAN_TENSORDOT;
  p.L = parse_tensor(Ad, Ad->rank, "A");
  p.R = parse_tensor(Bd, Bd->rank, "B");
  const Ten &A = p.L, &B = p.R;
  QLG_CHECK(A.dtype == B.dtype && A.qn_len == B.qn_len, "tensordot: operands differ in dtype or charges");
  p.dtype = A.dtype;
  const int ra = A.rank, rb = B.rank;
  QLG_CHECK(ra - k + rb - k >= 1 && ra - k + rb - k <= QLG_MAX_RANK, "tensordot: bad result rank");
  for (int i = 0; i < k; ++i)
    QLG_CHECK(A.ax[ra - k + i].same_sectors(B.ax[i]) && A.ax[ra - k + i].dir != B.ax[i].dir,
              "tensordot: contracted axes do not match");
  const double fl = p.dtype == QLG_C128 ? 8.0 : 2.0;
  p.vec.rank = ra - k + rb - k;
  for (int i = 0; i < ra - k; ++i) p.vec.ax.push_back(A.ax[i]);
  for (int i = k; i < rb; ++i) p.vec.ax.push_back(B.ax[i]);
  // bucket B by its leading sectors
  std::map<std::vector<int>, std::vector<int64_t>> Bb;
  for (size_t b = 0; b < B.blk.size(); ++b) Bb[std::vector<int>(B.blk[b].begin(), B.blk[b].begin() + k)].push_back(b);
  std::map<Coords, std::vector<std::pair<int64_t, int64_t>>> pairs;
  for (size_t a = 0; a < A.blk.size(); ++a) {
    auto it = Bb.find(std::vector<int>(A.blk[a].begin() + (ra - k), A.blk[a].begin() + ra));
    if (it == Bb.end()) continue;
    for (int64_t b : it->second) {
      Coords c{};
      for (int i = 0; i < ra - k; ++i) c[i] = A.blk[a][i];
      for (int i = k; i < rb; ++i) c[ra - k + i - k] = B.blk[b][i];
      pairs[c].push_back({(int64_t)a, b});
    }
  }
  std::set<Coords> outs;
  for (auto &kv : pairs) outs.insert(kv.first);
  finish_layout(p.vec, outs);
  Stage &s = p.st[0];
  for (size_t ob = 0; ob < p.vec.blk.size(); ++ob) {
    const Coords &c = p.vec.blk[ob];
    int64_t M = 1, N = 1;
    for (int i = 0; i < ra - k; ++i) M *= A.ax[i].deg[c[i]];
    for (int i = ra - k; i < p.vec.rank; ++i) N *= p.vec.ax[i].deg[c[i]];
    GemmTask tk{};
    tk.c_slot = SLOT_OUT;
    tk.c_off = p.vec.off[ob];
    tk.ldc = N;
    tk.M = (int32_t)M;
    tk.N = (int32_t)N;
    tk.seg_begin = (int32_t)s.segs.size();
    for (auto &pr : pairs[c]) {
      int64_t K = 1;
      for (int i = 0; i < k; ++i) K *= B.ax[i].deg[B.blk[pr.second][i]];
      GemmSeg g{};
      g.a_slot = SLOT_L;
      g.a_off = A.off[pr.first];
      g.lda = K;
      g.b_slot = SLOT_R;
      g.b_off = B.off[pr.second];
      g.ldb = N;
      g.K = (int32_t)K;
      g.flags = (conj_a ? SEG_CONJ_A : 0) | (conj_b ? SEG_CONJ_B : 0);
      s.segs.push_back(g);
      tk.seg_count++;
      s.flops += fl * double(M) * double(N) * double(K);
    }
    s.tasks.push_back(tk);
  }
  p.algo_flops = s.flops;
  int64_t slot_elems[NSLOTS] = {0};
  slot_elems[SLOT_L] = extent(A);
  slot_elems[SLOT_R] = extent(B);
  slot_elems[SLOT_OUT] = p.vec.size;
  finalize(p, 1, slot_elems, p.dtype == QLG_C128);
  return P.release();
}

// ---------------------------------------------------------------------------
// environment updates
// ---------------------------------------------------------------------------
namespace {

using Key4 = std::array<int32_t, 4>;  // (phys sector, phys comp, mpo sector, mpo comp)

struct EnvChunk {
  int64_t stage1_a_off, stage1_b_off, lda, ldb;
  int32_t M, N, K;
  uint8_t flags;
  int32_t wg, sg;   // global MPO / physical component of the input row
};

struct EnvChannel {
  int r, c;            // row sector (of the stage-1 result), column sector
  int64_t P;
  std::vector<EnvChunk> in;
  std::vector<Key4> out;
  std::map<Key4, int> out_index;
  int64_t x_off = 0, y_off = 0, c_off = 0;
};

}  // namespace

Plan *build_env_plan(Ctx *ctx, const qlg_bsten *envd, const qlg_bsten *mpsd, const qlg_bsten *mpod, char side) {
  QLG_CHECK(side == 'l' || side == 'r', "side must be 'l' or 'r'");
  std::unique_ptr<Plan> P(new Plan());
  Plan &p = *P;
  p.ctx = ctx;
  p.kind = side == 'l' ? PLAN_ENV_LEFT : PLAN_ENV_RIGHT;
  Ten env = parse_tensor(envd, 3, "env");
  Ten A = parse_tensor(mpsd, 3, "mps");
  Ten W = parse_tensor(mpod, 4, "mpo");
  QLG_CHECK(env.dtype == A.dtype && A.dtype == W.dtype, "env update: operands differ in dtype");
  p.dtype = A.dtype;
  const int ed = elem_doubles(p.dtype);
  const double fl = p.dtype == QLG_C128 ? 8.0 : 2.0;
  const uint8_t conj = p.dtype == QLG_C128 ? SEG_CONJ_A : 0;
  std::vector<int64_t> wd;
  std::vector<double> Wdn = mpo_dense(W, ctx->stream, ed, wd);
  auto Widx = [&](int64_t a, int64_t b, int64_t c, int64_t d) { return ((a * wd[1] + b) * wd[2] + c) * wd[3] + d; };
  std::vector<std::vector<int64_t>> wso(4);
  for (int i = 0; i < 4; ++i) wso[i] = sector_offsets(W.ax[i]);
  auto Aso1 = sector_offsets(A.ax[1]);

  std::map<std::pair<int, int>, int> chan_id;
  std::vector<EnvChannel> ch;
  auto channel = [&](int r, int c, int64_t P_) {
    auto key = std::make_pair(r, c);
    auto it = chan_id.find(key);
    if (it != chan_id.end()) return it->second;
    const int id = (int)ch.size();
    chan_id[key] = id;
    EnvChannel e;
    e.r = r;
    e.c = c;
    e.P = P_;
    ch.push_back(e);
    return id;
  };
  // MPS blocks indexed by the axis they meet the environment with
  std::map<int, std::vector<int64_t>> A_by0, A_by2;
  for (size_t b = 0; b < A.blk.size(); ++b) {
    A_by0[A.blk[b][0]].push_back(b);
    A_by2[A.blk[b][2]].push_back(b);
  }
  // MPO blocks by (in-bond, s_in) for the left update, by (s_in, out-bond) for the right update
  std::map<std::pair<int, int>, std::vector<int64_t>> W_key;
  for (size_t b = 0; b < W.blk.size(); ++b) {
    const Coords &w = W.blk[b];
    if (side == 'l')
      W_key[{w[0], w[1]}].push_back(b);
    else
      W_key[{w[1], w[3]}].push_back(b);
  }

  if (side == 'l') {
    // L[a', w, a] (bra sector a' = env axis 0), A[a, s, b], W[w, s, s', w1]
    check_pair(env.ax[2], A.ax[0], "env ket / mps left");
    check_pair(env.ax[1], W.ax[0], "env mpo / mpo left");
    check_pair(W.ax[1], A.ax[1], "mpo s_in / mps physical");
    QLG_CHECK(env.ax[0].same_sectors(A.ax[0]) && env.ax[0].dir == A.ax[0].dir, "env bra / mps left differ");
    QLG_CHECK(W.ax[2].same_sectors(A.ax[1]), "mpo s_out / mps physical differ");
    p.vec.rank = 3;
    p.vec.ax = {A.ax[2], W.ax[3], A.ax[2]};
    p.vec.ax[0].dir = -A.ax[2].dir;
    for (size_t lb = 0; lb < env.blk.size(); ++lb) {
      const Coords &l = env.blk[lb];
      auto it = A_by0.find(l[2]);
      if (it == A_by0.end()) continue;
      const int64_t Dap = env.ax[0].deg[l[0]], Da = env.ax[2].deg[l[2]], Wm = env.ax[1].deg[l[1]];
      for (int64_t ab : it->second) {
        const Coords &a = A.blk[ab];
        const int64_t Db = A.ax[2].deg[a[2]], ds = A.ax[1].deg[a[1]];
        const int cid = channel(l[0], a[2], Dap * Db);
        for (int w = 0; w < Wm; ++w)
          for (int s = 0; s < ds; ++s) {
            EnvChunk c{};
            c.stage1_a_off = env.off[lb] + w * Da;
            c.lda = Wm * Da;
            c.stage1_b_off = A.off[ab] + s * Db;
            c.ldb = ds * Db;
            c.M = (int32_t)Dap;
            c.N = (int32_t)Db;
            c.K = (int32_t)Da;
            c.flags = 0;
            c.wg = (int32_t)(wso[0][l[1]] + w);
            c.sg = (int32_t)(Aso1[a[1]] + s);
            ch[cid].in.push_back(c);
          }
      }
    }
  } else {
    // R[b', w, b] (bra sector b' = env axis 0), B[a, s, b], W[w0, s, s', w]
    check_pair(env.ax[2], A.ax[2], "env ket / mps right");
    check_pair(env.ax[1], W.ax[3], "env mpo / mpo right");
    check_pair(W.ax[1], A.ax[1], "mpo s_in / mps physical");
    QLG_CHECK(env.ax[0].same_sectors(A.ax[2]) && env.ax[0].dir == A.ax[2].dir, "env bra / mps right differ");
    QLG_CHECK(W.ax[2].same_sectors(A.ax[1]), "mpo s_out / mps physical differ");
    p.vec.rank = 3;
    p.vec.ax = {A.ax[0], W.ax[0], A.ax[0]};
    p.vec.ax[0].dir = -A.ax[0].dir;
    for (size_t rb = 0; rb < env.blk.size(); ++rb) {
      const Coords &r = env.blk[rb];
      auto it = A_by2.find(r[2]);
      if (it == A_by2.end()) continue;
      const int64_t Dbp = env.ax[0].deg[r[0]], Db = env.ax[2].deg[r[2]], Wr_ = env.ax[1].deg[r[1]];
      for (int64_t ab : it->second) {
        const Coords &a = A.blk[ab];
        const int64_t Da = A.ax[0].deg[a[0]], ds = A.ax[1].deg[a[1]];
        const int cid = channel(a[0], r[0], Da * Dbp);
        for (int w = 0; w < Wr_; ++w)
          for (int s = 0; s < ds; ++s) {
            EnvChunk c{};
            c.stage1_a_off = A.off[ab] + s * Db;   // B[a, s, :] rows a
            c.lda = ds * Db;
            c.stage1_b_off = env.off[rb] + w * Db;  // R[b', w, :] used transposed
            c.ldb = Wr_ * Db;
            c.M = (int32_t)Da;
            c.N = (int32_t)Dbp;
            c.K = (int32_t)Db;
            c.flags = SEG_TRANS_B;
            c.wg = (int32_t)(wso[3][r[1]] + w);
            c.sg = (int32_t)(Aso1[a[1]] + s);
            ch[cid].in.push_back(c);
          }
      }
    }
  }
  // stage-2 output rows: (s' sector, s' comp, mpo-out sector, comp), reachable through W and consumed
  // by a conjugated MPS block in stage 3
  auto so_s = sector_offsets(A.ax[1]);
  for (EnvChannel &c : ch) {
    std::set<Key4> outs;
    for (const EnvChunk &k : c.in) {
      // recover sectors of the input row from global components
      int wsec = 0, ssec = 0;
      const int wax = side == 'l' ? 0 : 3;
      while (wso[wax][wsec + 1] <= k.wg) ++wsec;
      while (so_s[ssec + 1] <= k.sg) ++ssec;
      auto it = W_key.find(side == 'l' ? std::make_pair(wsec, ssec) : std::make_pair(ssec, wsec));
      if (it == W_key.end()) continue;
      for (int64_t wb : it->second) {
        const Coords &w = W.blk[wb];
        const int sp = w[2];
        const int mo = side == 'l' ? w[3] : w[0];   // MPO bond of the new environment
        // consumer: conjugated MPS block with (row sector of the channel, s') for the left update,
        // (s', column sector) for the right update
        bool consumed = false;
        if (side == 'l') {
          for (int64_t ab : A_by0[c.r])
            if (A.blk[ab][1] == sp) { consumed = true; break; }
        } else {
          for (int64_t ab : A_by2[c.c])
            if (A.blk[ab][1] == sp) { consumed = true; break; }
        }
        if (!consumed) continue;
        const int64_t dsp = A.ax[1].deg[sp];
        const int64_t dm = W.ax[side == 'l' ? 3 : 0].deg[mo];
        for (int a = 0; a < dsp; ++a)
          for (int z = 0; z < dm; ++z) outs.insert(Key4{sp, a, mo, z});
      }
    }
    c.out.assign(outs.begin(), outs.end());
    for (size_t i = 0; i < c.out.size(); ++i) c.out_index[c.out[i]] = (int)i;
  }
  int64_t xo = 0, yo = 0, co = 0;
  for (EnvChannel &c : ch) {
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
  std::vector<double> Ch(std::max<int64_t>(co, 1) * ed, 0.0);
  for (EnvChannel &c : ch) {
    const int64_t K = (int64_t)c.in.size();
    for (size_t r = 0; r < c.out.size(); ++r) {
      const Key4 &o = c.out[r];
      const int64_t spg = so_s[o[0]] + o[1];
      const int64_t mg = wso[side == 'l' ? 3 : 0][o[2]] + o[3];
      for (int64_t q = 0; q < K; ++q) {
        const EnvChunk &k = c.in[q];
        const int64_t i = side == 'l' ? Widx(k.wg, k.sg, spg, mg) : Widx(mg, k.sg, spg, k.wg);
        const int64_t e = c.c_off + (int64_t)r * K + q;
        for (int z = 0; z < ed; ++z) Ch[e * ed + z] = Wdn[i * ed + z];
      }
    }
  }
  // ---- tasks: stages 1 and 2
  Stage &s1 = p.st[0], &s2 = p.st[1], &s3 = p.st[2];
  for (EnvChannel &c : ch) {
    for (size_t q = 0; q < c.in.size(); ++q) {
      const EnvChunk &k = c.in[q];
      GemmSeg g{};
      g.a_slot = side == 'l' ? SLOT_L : SLOT_IN;
      g.a_off = k.stage1_a_off;
      g.lda = k.lda;
      g.b_slot = side == 'l' ? SLOT_IN : SLOT_R;
      g.b_off = k.stage1_b_off;
      g.ldb = k.ldb;
      g.K = k.K;
      g.flags = k.flags;
      GemmTask t{};
      t.c_slot = SLOT_X;
      t.c_off = c.x_off + (int64_t)q * c.P;
      t.ldc = k.N;
      t.M = k.M;
      t.N = k.N;
      t.seg_begin = (int32_t)s1.segs.size();
      t.seg_count = 1;
      s1.segs.push_back(g);
      s1.tasks.push_back(t);
      s1.flops += fl * double(k.M) * double(k.N) * double(k.K);
    }
    if (!c.out.empty() && !c.in.empty()) {
      GemmSeg g{};
      g.a_slot = SLOT_C;
      g.a_off = c.c_off;
      g.lda = (int64_t)c.in.size();
      g.b_slot = SLOT_X;
      g.b_off = c.x_off;
      g.ldb = c.P;
      g.K = (int32_t)c.in.size();
      GemmTask t{};
      t.c_slot = SLOT_Y;
      t.c_off = c.y_off;
      t.ldc = c.P;
      t.M = (int32_t)c.out.size();
      t.N = (int32_t)c.P;
      t.seg_begin = (int32_t)s2.segs.size();
      t.seg_count = 1;
      s2.segs.push_back(g);
      s2.tasks.push_back(t);
      s2.flops += fl * double(c.out.size()) * double(c.in.size()) * double(c.P);
    }
  }
  // ---- stage 3: output blocks and their segments
  struct Seg3 {
    int64_t a_off, lda, b_off, ldb;
    int32_t K;
    uint8_t flags;
  };
  std::map<std::pair<Coords, int>, std::vector<Seg3>> out_tasks;  // (out block, mpo comp) -> segments
  for (EnvChannel &c : ch) {
    for (size_t r = 0; r < c.out.size(); ++r) {
      const Key4 &o = c.out[r];
      const int sp = o[0], sl = o[1], mo = o[2], mz = o[3];
      if (side == 'l') {
        // A*[a', s', b'] with a' in the channel's row sector: out block (b', mo, b = channel column)
        for (int64_t ab : A_by0[c.r]) {
          const Coords &a = A.blk[ab];
          if (a[1] != sp) continue;
          const int64_t Dap = A.ax[0].deg[a[0]], dsp = A.ax[1].deg[sp], Dbp = A.ax[2].deg[a[2]];
          Coords ob{};
          ob[0] = a[2];
          ob[1] = mo;
          ob[2] = c.c;
          // op(A) = A*_{s'}^T [D_b' x D_a'] from storage [D_a' x D_b'] (row stride d_s' D_b');
          // B = Y chunk [D_a' x D_b] (row stride D_b)
          Seg3 g{A.off[ab] + sl * Dbp, dsp * Dbp, c.y_off + (int64_t)r * c.P, c.P / Dap, (int32_t)Dap,
                 (uint8_t)(SEG_TRANS_A | conj)};
          out_tasks[{ob, mz}].push_back(g);
        }
      } else {
        // B*[a', s', b'] with b' in the channel's column sector: out block (a', mo, a = channel row)
        for (int64_t ab : A_by2[c.c]) {
          const Coords &a = A.blk[ab];
          if (a[1] != sp) continue;
          const int64_t dsp = A.ax[1].deg[sp], Dbp = A.ax[2].deg[a[2]];
          Coords ob{};
          ob[0] = a[0];
          ob[1] = mo;
          ob[2] = c.r;
          Seg3 g{A.off[ab] + sl * Dbp, dsp * Dbp, c.y_off + (int64_t)r * c.P, Dbp, (int32_t)Dbp,
                 (uint8_t)(SEG_TRANS_B | conj)};
          out_tasks[{ob, mz}].push_back(g);
        }
      }
    }
  }
  std::set<Coords> outs;
  for (auto &kv : out_tasks) outs.insert(kv.first.first);
  finish_layout(p.vec, outs);
  for (auto &kv : out_tasks) {
    const Coords &ob = kv.first.first;
    const int mz = kv.first.second;
    const int64_t D0 = p.vec.ax[0].deg[ob[0]], Wm = p.vec.ax[1].deg[ob[1]], D2 = p.vec.ax[2].deg[ob[2]];
    GemmTask t{};
    t.c_slot = SLOT_OUT;
    t.c_off = p.vec.off[p.vec.find[ob]] + mz * D2;
    t.ldc = Wm * D2;
    t.M = (int32_t)D0;
    t.N = (int32_t)D2;
    t.seg_begin = (int32_t)s3.segs.size();
    t.seg_count = (int32_t)kv.second.size();
    for (const Seg3 &g3 : kv.second) {
      GemmSeg g{};
      g.a_slot = SLOT_IN;
      g.a_off = g3.a_off;
      g.lda = g3.lda;
      g.b_slot = SLOT_Y;
      g.b_off = g3.b_off;
      g.ldb = g3.ldb;
      g.K = g3.K;
      g.flags = g3.flags;
      s3.segs.push_back(g);
      s3.flops += fl * double(D0) * double(D2) * double(g3.K);
    }
    s3.tasks.push_back(t);
  }
  // output blocks not reached by any mpo component keep zeros: every block has
  // W_m tasks only where segments exist, so add zero tasks for the rest
  {
    std::set<std::pair<Coords, int>> have;
    for (auto &kv : out_tasks) have.insert(kv.first);
    for (size_t b = 0; b < p.vec.blk.size(); ++b) {
      const Coords &ob = p.vec.blk[b];
      const int64_t Wm = p.vec.ax[1].deg[ob[1]], D0 = p.vec.ax[0].deg[ob[0]], D2 = p.vec.ax[2].deg[ob[2]];
      for (int z = 0; z < Wm; ++z) {
        if (have.count({ob, z})) continue;
        GemmTask t{};
        t.c_slot = SLOT_OUT;
        t.c_off = p.vec.off[b] + z * D2;
        t.ldc = Wm * D2;
        t.M = (int32_t)D0;
        t.N = (int32_t)D2;
        t.seg_begin = (int32_t)s3.segs.size();
        t.seg_count = 0;
        s3.tasks.push_back(t);
      }
    }
  }
  p.algo_flops = s1.flops + s2.flops + s3.flops;
  p.L = env;   // slot L or R binds the environment
  p.R = env;
  p.W[0] = A;  // the MPS tensor binds slot IN at execution
  p.X.ensure(std::max<int64_t>(p.x_elems, 1) * ed * sizeof(double));
  p.Y.ensure(std::max<int64_t>(p.y_elems, 1) * ed * sizeof(double));
  p.C.ensure(std::max<int64_t>(p.c_elems, 1) * ed * sizeof(double));
  QLG_HIP(hipMemcpyAsync(p.C.p, Ch.data(), std::max<int64_t>(co, 1) * ed * sizeof(double), hipMemcpyHostToDevice,
                         ctx->stream));
  int64_t slot_elems[NSLOTS] = {0};
  slot_elems[SLOT_L] = slot_elems[SLOT_R] = extent(env);
  slot_elems[SLOT_IN] = extent(A);
  slot_elems[SLOT_OUT] = p.vec.size;
  slot_elems[SLOT_X] = p.x_elems;
  slot_elems[SLOT_Y] = p.y_elems;
  slot_elems[SLOT_C] = p.c_elems;
  finalize(p, 3, slot_elems, p.dtype == QLG_C128);
  QLG_HIP(hipStreamSynchronize(ctx->stream));
  return P.release();
}

// ---------------------------------------------------------------------------
Plan *build_transpose_plan(Ctx *ctx, const qlg_bsten *td, const int32_t *perm) {
  QLG_CHECK(td && perm, "transpose: null argument");
  std::unique_ptr<Plan> P(new Plan());
  Plan &p = *P;
  p.ctx = ctx;
  p.kind = PLAN_TRANSPOSE;
  p.L = parse_tensor(td, td->rank, "t");
  const Ten &t = p.L;
  p.dtype = t.dtype;
  const int r = t.rank;
  std::vector<int> seen(r, 0);
  for (int i = 0; i < r; ++i) {
    QLG_CHECK(perm[i] >= 0 && perm[i] < r && !seen[perm[i]], "transpose: perm is not a permutation of the axes");
    seen[perm[i]] = 1;
  }
  p.vec.rank = r;
  for (int i = 0; i < r; ++i) p.vec.ax.push_back(t.ax[perm[i]]);
  std::set<Coords> outs;
  std::vector<Coords> dst_of(t.blk.size());
  for (size_t b = 0; b < t.blk.size(); ++b) {
    Coords c{};
    for (int i = 0; i < r; ++i) c[i] = t.blk[b][perm[i]];
    QLG_CHECK(outs.insert(c).second, "transpose: duplicate block");
    dst_of[b] = c;
  }
  finish_layout(p.vec, outs);
  std::vector<PermJob> jobs;
  for (size_t b = 0; b < t.blk.size(); ++b) {
    PermJob j{};
    j.src = t.off[b];
    j.dst = p.vec.off[p.vec.find[dst_of[b]]];
    j.n = t.blk_elems(b);
    j.rank = r;
    // destination strides of destination axes, then mapped to source axes
    int64_t ds[QLG_MAX_RANK], acc = 1;
    for (int i = r - 1; i >= 0; --i) {
      ds[i] = acc;
      acc *= p.vec.ax[i].deg[dst_of[b][i]];
    }
    for (int i = 0; i < r; ++i) {
      j.dim[i] = t.ax[i].deg[t.blk[b][i]];
      j.dstride[perm[i]] = ds[i];
    }
    if (j.n == 0) continue;
    p.perm_maxel = std::max(p.perm_maxel, j.n);
    jobs.push_back(j);
  }
  p.perm_n = (int64_t)jobs.size();
  p.perm_jobs.ensure(std::max<size_t>(1, jobs.size()) * sizeof(PermJob));
  if (!jobs.empty())
    QLG_HIP(hipMemcpyAsync(p.perm_jobs.p, jobs.data(), jobs.size() * sizeof(PermJob), hipMemcpyHostToDevice,
                           ctx->stream));
  QLG_HIP(hipStreamSynchronize(ctx->stream));
  p.nstages = 0;
  return P.release();
}

// Execute a tensordot / environment / transpose plan: in0, in1 are (A, B) for tensordot and
// (env, mps) for environment updates; null pointers use the plan-time data.
void Plan::execute(const void *in0, const void *in1, void *out) {
  Slots s{};
  auto dp = [](const void *p) { return const_cast<double *>(static_cast<const double *>(p)); };
  s.p[SLOT_OUT] = static_cast<double *>(out);
  if (kind == PLAN_TRANSPOSE) {
    launch_transpose(ctx->stream, perm_jobs.as<PermJob>(), perm_n, perm_maxel,
                     static_cast<const double *>(in0 ? in0 : L.data), static_cast<double *>(out), elem_doubles(dtype));
    return;
  }
  if (kind == PLAN_TENSORDOT) {
    s.p[SLOT_L] = dp(in0 ? in0 : L.data);
    s.p[SLOT_R] = dp(in1 ? in1 : R.data);
  } else {
    s.p[SLOT_L] = s.p[SLOT_R] = dp(in0 ? in0 : L.data);
    s.p[SLOT_IN] = dp(in1 ? in1 : W[0].data);
    s.p[SLOT_X] = X.as<double>();
    s.p[SLOT_Y] = Y.as<double>();
    s.p[SLOT_C] = C.as<double>();
  }
  for (int i = 0; i < nstages; ++i) {
    Stage &g = st[i];
    launch_grouped_gemm_f64(g.bm, g.bn, dtype == QLG_C128, g.d_tiles.as<GemmTile>(), g.d_tasks.as<GemmTask>(),
                            g.d_segs.as<GemmSeg>(), s, (int)g.tiles.size(), ctx->stream);
  }
}

}  // namespace qlg
