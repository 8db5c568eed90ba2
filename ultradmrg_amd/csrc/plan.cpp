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
// Launch order of the tiles (make_tiles_lpt): per-task tile counts and the
// padded K of each task feed the XCD-balanced placement.
void Stage::make_tiles(int bm_, int bn_) {
  bm = bm_;
  bn = bn_;
  std::vector<std::array<int64_t, 3>> mnk;
  if (!chan.empty()) {
    for (const ChanTask &c : chan) mnk.push_back({c.kout, c.P, c.kin});
  } else {
    for (const GemmTask &t : tasks) {
      int64_t kp = 0;
      for (int s = 0; s < t.seg_count; ++s) kp += (segs[t.seg_begin + s].K + 15) / 16 * 16;
      mnk.push_back({t.M, t.N, kp});
    }
  }
  kchunk = 0;
  ktasks.clear();
  ksegs.clear();
  red.clear();
  redc.clear();
  work_elems = 0;
  if (splitk && split_on && chan.empty()) kchunk = pick_kchunk(mnk, bm, bn);
  if (kchunk > 0) {
    // Cut every task whose K exceeds 3/2 of the chunk into parts of at most
    // kchunk.  A cut falls inside a segment only at a multiple of 64 of its
    // K, so every part's operand offsets keep the alignment of the segment's
    // own (16-byte rows of the bf16 planes).
    std::vector<std::array<int64_t, 3>> kmnk;
    for (size_t ti = 0; ti < tasks.size(); ++ti) {
      const GemmTask &t = tasks[ti];
      int64_t ktot = 0;
      for (int q = 0; q < t.seg_count; ++q) ktot += segs[t.seg_begin + q].K;
      // tasks too short to cut, and panels too large for the reduce
      // kernel's 32-bit element index, run unsplit
      if (ktot * 2 <= kchunk * 3 || t.M <= 0 || t.N <= 0 || (int64_t)t.M * t.N >= (int64_t(1) << 31)) {
        GemmTask c = t;
        c.seg_begin = (int32_t)ksegs.size();
        for (int q = 0; q < t.seg_count; ++q) ksegs.push_back(segs[t.seg_begin + q]);
        ktasks.push_back(c);
        kmnk.push_back(mnk[ti]);
        continue;
      }
      RedTask r{};
      r.c_off = t.c_off;
      r.ldc = t.ldc;
      r.c_row = t.c_row;
      r.c_slot = t.c_slot;
      r.c_rowtab = t.c_rowtab;
      r.M = t.M;
      r.N = t.N;
      r.wbase = work_elems;
      const int64_t panel = (int64_t)t.M * t.N;
      int parts = 0;
      GemmTask cur{};
      int64_t have = 0;   // K in the part being filled
      auto open_part = [&]() {
        cur = GemmTask{};
        cur.c_slot = SLOT_W;
        cur.c_off = r.wbase + (int64_t)parts * panel;
        cur.ldc = t.N;
        cur.M = t.M;
        cur.N = t.N;
        cur.seg_begin = (int32_t)ksegs.size();
        cur.seg_count = 0;
        have = 0;
      };
      auto close_part = [&]() {
        ktasks.push_back(cur);
        kmnk.push_back({cur.M, cur.N, (have + 15) / 16 * 16});
        ++parts;
      };
      open_part();
      for (int q = 0; q < t.seg_count; ++q) {
        const GemmSeg &g = segs[t.seg_begin + q];
        const bool ta = g.flags & SEG_TRANS_A, tb = g.flags & SEG_TRANS_B;
        int64_t k0 = 0;
        while (k0 < g.K) {
          if (have >= kchunk) {
            close_part();
            open_part();
          }
          int64_t take = std::min<int64_t>(g.K - k0, kchunk - have);
          if (k0 + take < g.K) {
            take = take / 64 * 64;   // cut at a multiple of 64 of the segment's K
            if (take == 0) {
              if (have == 0) take = std::min<int64_t>(g.K - k0, 64);   // kchunk >= 64: unreachable
              else {
                close_part();
                open_part();
                continue;
              }
            }
          }
          GemmSeg h = g;
          h.K = (int32_t)take;
          h.a_off = g.a_off + k0 * (ta ? g.lda : 1);
          h.b_off = g.b_off + k0 * (tb ? 1 : g.ldb);
          ksegs.push_back(h);
          ++cur.seg_count;
          have += take;
          k0 += take;
        }
      }
      if (cur.seg_count > 0) close_part();
      r.parts = parts;
      work_elems += (int64_t)parts * panel;
      for (int64_t e = 0; e < panel; e += kRedChunk) redc.push_back(RedChunk{(int32_t)red.size(), (int32_t)e});
      red.push_back(r);
    }
    // a workspace beyond 2^29 elements (4 GiB in FP64, 8 GiB complex) is not
    // worth its traffic: such a stage runs unsplit
    if (red.empty() || work_elems >= (int64_t(1) << 29)) {
      red.clear();
      redc.clear();
      work_elems = 0;
      kchunk = 0;
      ktasks.clear();
      ksegs.clear();
      make_tiles_lpt(tiles, mnk, bm, bn);
      return;
    }
    make_tiles_lpt(tiles, kmnk, bm, bn);
    return;
  }
  make_tiles_lpt(tiles, mnk, bm, bn);
}

// Split-K chunk for a task list and tile shape.  A grouped launch ends with
// its longest tiles: a tile's time grows with its K, and the stage-3 tasks
// of the matvec (one per bra sector, one long segment each) put K up to a
// few thousand on every tile of a task.  The chunk is the tile-K that one of
// 2048 resident workgroups (256 CUs x 8) would process if the work were
// spread evenly, halved, so the last tiles to start are short; at least 512
// (shorter parts cost more in partial panels than they gain in balance).
// QLG_SPLITK overrides it (0: off, else the chunk, rounded up to 64).
int64_t pick_kchunk(const std::vector<std::array<int64_t, 3>> &mnk, int bm, int bn) {
  if (const char *v = std::getenv("QLG_SPLITK")) {
    const int64_t k = std::atoll(v);
    return k <= 0 ? 0 : (k + 63) / 64 * 64;
  }
  double work = 0;
  for (const auto &t : mnk) work += double((t[0] + bm - 1) / bm) * double((t[1] + bn - 1) / bn) * double(t[2]);
  const int64_t k = (int64_t)(work / 2048.0 / 2.0);
  return std::max<int64_t>(512, (k + 63) / 64 * 64);
}

// Workgroups are dispatched to the 8 XCDs round-robin in blockIdx order, and
// each XCD has its own L2.  A task's tiles share its A row panels and B
// column panels, so a task stays on one XCD unless it is large: a task whose
// cost exceeds 1/32 of the stage is cut into bands of whole tile rows (or
// whole tile columns, if it has more column than row tiles), so each band
// fetches its own A rows (B columns) once and the other operand at most once
// per band.  Units are placed longest-first on the least-loaded XCD, tiles
// of a unit in row-major order, and XCD x's j-th tile is launched as block
// j * 8 + x.
void make_tiles_lpt(std::vector<GemmTile> &tiles, const std::vector<std::array<int64_t, 3>> &mnk, int bm, int bn) {
  constexpr int kXcd = 8, kUnitsPerXcd = 4;
  tiles.clear();
  struct Unit {
    int32_t task;
    int64_t r0, r1, c0, c1;   // tile-row / tile-column ranges
    double cost;
  };
  std::vector<double> tile_cost(mnk.size());
  std::vector<int64_t> ntm(mnk.size()), ntn(mnk.size());
  double total = 0;
  for (size_t t = 0; t < mnk.size(); ++t) {
    ntm[t] = (mnk[t][0] + bm - 1) / bm;
    ntn[t] = (mnk[t][1] + bn - 1) / bn;
    tile_cost[t] = double(bm) * bn * std::max<int64_t>(mnk[t][2], 16) + 4096.0;
    total += tile_cost[t] * double(ntm[t]) * double(ntn[t]);
  }
  const double target = total / (kXcd * kUnitsPerXcd);
  std::vector<Unit> units;
  for (size_t t = 0; t < mnk.size(); ++t) {
    const double cost = tile_cost[t] * double(ntm[t]) * double(ntn[t]);
    if (ntm[t] * ntn[t] == 0) continue;
    const int64_t pieces = std::max<int64_t>(1, (int64_t)std::ceil(cost / target));
    const bool by_rows = ntm[t] >= ntn[t];
    const int64_t len = by_rows ? ntm[t] : ntn[t];
    const int64_t np = std::min(pieces, len);
    for (int64_t q = 0; q < np; ++q) {
      const int64_t lo = len * q / np, hi = len * (q + 1) / np;
      Unit u{(int32_t)t, 0, ntm[t], 0, ntn[t], 0.0};
      if (by_rows) {
        u.r0 = lo;
        u.r1 = hi;
      } else {
        u.c0 = lo;
        u.c1 = hi;
      }
      u.cost = tile_cost[t] * double(u.r1 - u.r0) * double(u.c1 - u.c0);
      units.push_back(u);
    }
  }
  std::stable_sort(units.begin(), units.end(), [](const Unit &a, const Unit &b) { return a.cost > b.cost; });
  std::vector<std::vector<GemmTile>> per(kXcd);
  std::vector<double> load(kXcd, 0.0);
  for (const Unit &u : units) {
    const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += u.cost;
    for (int64_t r = u.r0; r < u.r1; ++r)
      for (int64_t c = u.c0; c < u.c1; ++c)
        per[x].push_back(GemmTile{u.task, (int32_t)(r * bm), (int32_t)(c * bn), 0});
  }
  size_t slots = 0;
  for (auto &v : per) slots = std::max(slots, v.size());
  tiles.reserve(slots * kXcd);
  for (size_t s = 0; s < slots; ++s)
    for (int x = 0; x < kXcd; ++x) tiles.push_back(s < per[x].size() ? per[x][s] : GemmTile{-1, 0, 0, 0});
  while (!tiles.empty() && tiles.back().task < 0) tiles.pop_back();
}

void Stage::upload(hipStream_t st) {
  if (!chan.empty()) {
    d_chan.ensure(chan.size() * sizeof(ChanTask));
    d_rows.ensure(std::max<size_t>(1, rows.size()) * sizeof(RowDesc));
    QLG_HIP(hipMemcpyAsync(d_chan.p, chan.data(), chan.size() * sizeof(ChanTask), hipMemcpyHostToDevice, st));
    if (!rows.empty())
      QLG_HIP(hipMemcpyAsync(d_rows.p, rows.data(), rows.size() * sizeof(RowDesc), hipMemcpyHostToDevice, st));
    ctiles.assign(tiles.size(), ChanTile{});
    for (size_t i = 0; i < tiles.size(); ++i) {
      const GemmTile &t = tiles[i];
      ChanTile &o = ctiles[i];
      if (t.task < 0) {
        o.m0 = -1;
        continue;
      }
      QLG_CHECK((size_t)t.task < chan.size(), "MPO-stage tile: channel index out of range");
      o.c = chan[t.task];
      o.m0 = t.m0;
      o.n0 = t.n0;
    }
    d_ctiles.ensure(std::max<size_t>(1, ctiles.size()) * sizeof(ChanTile));
    if (!ctiles.empty())
      QLG_HIP(hipMemcpyAsync(d_ctiles.p, ctiles.data(), ctiles.size() * sizeof(ChanTile), hipMemcpyHostToDevice, st));
  }
  if (!crows.empty()) {
    d_crows.ensure(crows.size() * sizeof(int64_t));
    QLG_HIP(hipMemcpyAsync(d_crows.p, crows.data(), crows.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
  }
  // the launch lists: the split ones when a task is split
  const std::vector<GemmTask> &lt = split() ? ktasks : tasks;
  const std::vector<GemmSeg> &ls = split() ? ksegs : segs;
  d_tiles.ensure(std::max<size_t>(1, tiles.size()) * sizeof(GemmTile));
  d_tasks.ensure(std::max<size_t>(1, lt.size()) * sizeof(GemmTask));
  d_segs.ensure(std::max<size_t>(1, ls.size()) * sizeof(GemmSeg));
  if (!tiles.empty())
    QLG_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), tiles.size() * sizeof(GemmTile), hipMemcpyHostToDevice, st));
  if (!lt.empty())
    QLG_HIP(hipMemcpyAsync(d_tasks.p, lt.data(), lt.size() * sizeof(GemmTask), hipMemcpyHostToDevice, st));
  if (!ls.empty())
    QLG_HIP(hipMemcpyAsync(d_segs.p, ls.data(), ls.size() * sizeof(GemmSeg), hipMemcpyHostToDevice, st));
  if (split()) {
    d_red.ensure(red.size() * sizeof(RedTask));
    d_redc.ensure(redc.size() * sizeof(RedChunk));
    work.ensure((size_t)std::max<int64_t>(work_elems, 1) * ed * sizeof(double));
    QLG_HIP(hipMemcpyAsync(d_red.p, red.data(), red.size() * sizeof(RedTask), hipMemcpyHostToDevice, st));
    QLG_HIP(hipMemcpyAsync(d_redc.p, redc.data(), redc.size() * sizeof(RedChunk), hipMemcpyHostToDevice, st));
  }
}

// Bounds of every access a task list can make: rows < M, cols < N, k < K of
// each segment, relative to the buffer behind its slot (in elements).
void validate_stage(const Stage &s, const int64_t *slot_elems) {
  // stage-2 channel tasks (gathered / scattered rows)
  for (const ChanTask &c : s.chan) {
    QLG_CHECK(c.kin > 0 && c.kout > 0 && c.Db > 0 && c.P % c.Db == 0, "task list: bad channel shape");
    QLG_CHECK(c.c_off >= 0 && c.c_off + (int64_t)c.kin * c.kout <= slot_elems[SLOT_C], "task list: channel operator");
    const int64_t na = c.P / c.Db;
    for (int r = 0; r < c.kin; ++r) {
      const RowDesc &d = s.rows[c.in_rows + r];
      QLG_CHECK(d.off >= 0 && d.rs >= c.Db && d.off + (na - 1) * d.rs + c.Db - 1 < slot_elems[SLOT_X],
                "task list: channel input row outside the first intermediate");
    }
    for (int r = 0; r < c.kout; ++r) {
      const RowDesc &d = s.rows[c.out_rows + r];
      QLG_CHECK(d.off >= 0 && d.rs >= c.Db && d.off + (na - 1) * d.rs + c.Db - 1 < slot_elems[SLOT_Y],
                "task list: channel output row outside the second intermediate");
      // the MPO-stage kernel keeps output row strides in 32 bits
      QLG_CHECK(d.rs < (int64_t(1) << 31), "task list: channel output row stride beyond 32 bits");
    }
  }
  auto inside = [&](int slot, int64_t first, int64_t last, const char *what) {
    QLG_CHECK(slot > SLOT_ABS && slot < NSLOTS, std::string("task list: bad slot for ") + what);
    QLG_CHECK(first >= 0 && last < slot_elems[slot],
              std::string("task list: ") + what + " extent [" + std::to_string(first) + ", " + std::to_string(last) +
                  "] outside buffer of " + std::to_string(slot_elems[slot]) + " elements");
  };
  // the kernels address an operand through 32-bit byte offsets from its
  // first element (complex elements: 16 bytes)
  auto span = [&](int64_t first, int64_t last, const char *what) {
    if ((last - first + 1) * 16 >= (int64_t(1) << 31))
      throw Error(QLG_ERR_UNSUPPORTED, std::string("task list: ") + what + " operand spans more than 2^31 bytes");
  };
  for (const GemmTask &t : s.tasks) {
    QLG_CHECK(t.M > 0 && t.N > 0 && t.seg_count >= 0, "task list: empty task shape");
    if (t.c_rowtab) {
      QLG_CHECK(t.c_row >= 0 && (size_t)t.c_row + t.M <= s.crows.size(), "task list: row table range");
      for (int r = 0; r < t.M; ++r) inside(t.c_slot, s.crows[t.c_row + r], s.crows[t.c_row + r] + t.N - 1, "C row");
    } else {
      QLG_CHECK(t.ldc >= t.N, "task list: ldc < N");
      inside(t.c_slot, t.c_off, t.c_off + (int64_t)(t.M - 1) * t.ldc + (t.N - 1), "C");
    }
    for (int q = 0; q < t.seg_count; ++q) {
      const GemmSeg &g = s.segs[t.seg_begin + q];
      if (g.K <= 0) continue;
      if (g.flags & SEG_TRANS_A) {
        QLG_CHECK(g.lda >= t.M, "task list: lda < M");
        inside(g.a_slot, g.a_off, g.a_off + (int64_t)(g.K - 1) * g.lda + (t.M - 1), "A^T");
        span(g.a_off, g.a_off + (int64_t)(g.K - 1) * g.lda + (t.M - 1), "A^T");
      } else {
        QLG_CHECK(g.lda >= g.K, "task list: lda < K");
        inside(g.a_slot, g.a_off, g.a_off + (int64_t)(t.M - 1) * g.lda + (g.K - 1), "A");
        span(g.a_off, g.a_off + (int64_t)(t.M - 1) * g.lda + (g.K - 1), "A");
      }
      if (g.flags & SEG_TRANS_B) {
        QLG_CHECK(g.ldb >= g.K, "task list: ldb < K");
        inside(g.b_slot, g.b_off, g.b_off + (int64_t)(t.N - 1) * g.ldb + (g.K - 1), "B^T");
        span(g.b_off, g.b_off + (int64_t)(t.N - 1) * g.ldb + (g.K - 1), "B^T");
      } else {
        QLG_CHECK(g.ldb >= t.N, "task list: ldb < N");
        inside(g.b_slot, g.b_off, g.b_off + (int64_t)(g.K - 1) * g.ldb + (t.N - 1), "B");
        span(g.b_off, g.b_off + (int64_t)(g.K - 1) * g.ldb + (t.N - 1), "B");
      }
    }
  }
}

// Tile shapes of the grouped GEMMs ranked by a cost model: MFMA time of the
// padded tiles at the FP64 peak, operand streaming at an L2-assisted rate,
// the longest single tile, plus a fixed latency per tile.  Returns (estimate,
// shape index) pairs, best first.
// shapes 0-6 serve every kernel; the thin shapes 7-8 (four waves along M)
// only the real FP64 one
static const int kShapes[][2] = {{32, 32}, {64, 32}, {32, 64}, {64, 64}, {128, 64}, {64, 128}, {128, 128},
                                 {64, 16}, {128, 16}};

std::vector<std::pair<double, int>> rank_tile_shapes(const Stage &s, bool cplx, bool thin) {
  // complex tiles keep two accumulator planes: cap at 64x64
  const int nshapes = cplx ? 4 : (thin ? 9 : 7);
  std::vector<int64_t> kpad(s.tasks.size(), 0);
  for (size_t t = 0; t < s.tasks.size(); ++t)
    for (int q = 0; q < s.tasks[t].seg_count; ++q) kpad[t] += (s.segs[s.tasks[t].seg_begin + q].K + 15) / 16 * 16;
  const double ebytes = cplx ? 16.0 : 8.0, fmul = cplx ? 4.0 : 1.0;
  std::vector<std::pair<double, int>> out;
  for (int si = 0; si < nshapes; ++si) {
    const int *sh = kShapes[si];
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
    out.push_back({std::max({flops / 78.6e12, bytes / 12e12, longest}) + ntiles * 1.5e-6 / 1024.0, si});
  }
  std::stable_sort(out.begin(), out.end());
  return out;
}

void model_tiles(Stage &s, bool cplx) {
  const auto r = rank_tile_shapes(s, cplx, false);
  s.make_tiles(kShapes[r[0].second][0], kShapes[r[0].second][1]);
}

std::pair<int, int> tile_shape(int index) { return {kShapes[index][0], kShapes[index][1]}; }

// Key of the tile autotuner's cache: the stage, the arithmetic variant, the
// log2 of the stage's FLOPs and the (log2 M, log2 N, log2 K) cell that holds
// the largest share of them.  Coarse on purpose: the sites of a sweep at a
// fixed bond dimension share a key, so each is tuned once.
std::string tile_signature(const Stage &s, int stage, int variant) {
  auto ilog2 = [](int64_t v) {
    int r = 0;
    while (v > 1) {
      v >>= 1;
      ++r;
    }
    return r;
  };
  std::map<int, double> cells;
  double tot = 0;
  for (const GemmTask &k : s.tasks) {
    int64_t K = 0;
    for (int q = 0; q < k.seg_count; ++q) K += s.segs[k.seg_begin + q].K;
    const double f = 2.0 * k.M * k.N * K;
    cells[ilog2(k.M) * 1024 + ilog2(k.N) * 32 + ilog2(K)] += f;
    tot += f;
  }
  int top = -1;
  double best = -1;
  for (const auto &c : cells)
    if (c.second > best) {
      best = c.second;
      top = c.first;
    }
  return std::to_string(stage) + "/" + std::to_string(variant) + "/" +
         std::to_string(tot > 0 ? (int)std::lround(std::log2(tot)) : -1) + "/" + std::to_string(top);
}

// ---------------------------------------------------------------------------
std::vector<int64_t> sector_offsets(const Axis &a) {
  std::vector<int64_t> o(a.deg.size() + 1, 0);
  for (size_t s = 0; s < a.deg.size(); ++s) o[s + 1] = o[s] + a.deg[s];
  return o;
}

// elements spanned by a tensor's blocks from its data pointer
int64_t raw_elems(const Ten &t) {
  int64_t raw = 0;
  for (size_t b = 0; b < t.blk.size(); ++b) raw = std::max(raw, t.off[b] + t.blk_elems(b));
  return raw;
}

// dense host array of a small tensor from a host copy of its block data
// (raw_elems(t) elements), index order kept
std::vector<double> densify(const Ten &t, const double *h, int ed) {
  std::vector<int64_t> dims(t.rank);
  std::vector<std::vector<int64_t>> so(t.rank);
  int64_t tot = 1;
  for (int i = 0; i < t.rank; ++i) {
    dims[i] = t.ax[i].dim();
    so[i] = sector_offsets(t.ax[i]);
    tot *= dims[i];
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

// dense host copy of a small tensor (MPO), index order kept
std::vector<double> dense_copy(const Ten &t, hipStream_t st, int ed) {
  const int64_t raw = raw_elems(t);
  std::vector<double> h(raw * ed);
  if (raw > 0) {
    QLG_HIP(hipMemcpyAsync(h.data(), t.data, raw * ed * sizeof(double), hipMemcpyDeviceToHost, st));
    QLG_HIP(hipStreamSynchronize(st));
  }
  return densify(t, h.data(), ed);
}


// ---------------------------------------------------------------------------
}  // namespace qlg
