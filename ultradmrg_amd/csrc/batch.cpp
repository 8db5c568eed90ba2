// Batched effective Hamiltonians: independent chains (each with its own
// L, W_l, W_r, R and state) solved together.  A batch plan concatenates the
// member plans' operators, intermediates and Krylov vectors into arenas and
// their task lists into one list per stage, with every (slot, offset)
// operand address rebased into the arena, so each stage of all the chains'
// matvecs is one grouped launch.  The batched Lanczos solver (krylov.cpp)
// runs the chains in lock step with per-chain scalars.
#include <algorithm>
#include <array>

#include "plan.h"

namespace qlg {

Plan *build_batch_plan(Ctx *ctx, int64_t n, Plan *const *members) {
  QLG_CHECK(n >= 1 && members, "batch: needs at least one plan");
  const int dtype = members[0]->dtype;
  for (int64_t m = 0; m < n; ++m) {
    QLG_CHECK(members[m] && members[m]->kind == PLAN_HEFF, "batch: members must be effective-Hamiltonian plans");
    QLG_CHECK(members[m]->ctx == ctx, "batch: members must belong to the batch's context");
    QLG_CHECK(members[m]->dtype == dtype, "batch: members must share the element type");
  }
  if (dtype != QLG_F64) throw Error(QLG_ERR_UNSUPPORTED, "batch: real (QLG_F64) chains only");
  auto p = std::make_unique<Plan>();
  p->ctx = ctx;
  p->kind = PLAN_HEFF;
  p->nstages = 3;
  p->dtype = dtype;
  p->precision = QLG_PREC_F64;
  p->nphys = members[0]->nphys;
  // arena bases of every member, per slot
  std::vector<std::array<int64_t, NSLOTS>> base(n);
  int64_t vs = 0, xs = 0, ys = 0, cs = 0, ls = 0, rs = 0;
  p->batch_off.assign(n + 1, 0);
  for (int64_t m = 0; m < n; ++m) {
    const Plan &q = *members[m];
    auto &b = base[m];
    b.fill(0);
    b[SLOT_IN] = b[SLOT_OUT] = vs;
    b[SLOT_X] = xs;
    b[SLOT_Y] = ys;
    b[SLOT_C] = cs;
    b[SLOT_L] = ls;
    b[SLOT_R] = rs;
    vs += q.vec.size;
    xs += q.x_elems;
    ys += q.y_elems;
    cs += q.c_elems;
    ls += q.lint_elems;
    rs += q.rint_elems;
    p->batch_off[m + 1] = vs;
    p->algo_flops += q.algo_flops;
  }
  if (xs * 8 >= (int64_t(1) << 31))
    throw Error(QLG_ERR_UNSUPPORTED, "batch: the chains' first intermediates exceed 2^31 bytes together");
  p->vec.size = vs;
  p->x_elems = xs;
  p->y_elems = ys;
  p->c_elems = cs;
  p->lint_elems = ls;
  p->rint_elems = rs;
  // task lists
  for (int i = 0; i < 3; ++i) {
    Stage &d = p->st[i];
    for (int64_t m = 0; m < n; ++m) {
      const Stage &s = members[m]->st[i];
      const auto &b = base[m];
      const int32_t seg0 = (int32_t)d.segs.size(), crow0 = (int32_t)d.crows.size();
      for (GemmSeg g : s.segs) {
        g.a_off += b[g.a_slot];
        g.b_off += b[g.b_slot];
        d.segs.push_back(g);
      }
      for (GemmTask t : s.tasks) {
        QLG_CHECK(!t.c_rowtab || t.c_slot == SLOT_OUT, "batch: row tables must address the Krylov output");
        if (t.c_rowtab)
          t.c_row += crow0;
        else
          t.c_off += b[t.c_slot];
        t.seg_begin += seg0;
        d.tasks.push_back(t);
      }
      // row-table entries address the task's output slot (the Krylov output)
      for (int64_t r : s.crows) d.crows.push_back(r + b[SLOT_OUT]);
      for (ChanTask c : s.chan) {
        for (int r = 0; r < c.kin; ++r) {
          RowDesc e = s.rows[c.in_rows + r];
          e.off += b[SLOT_X];
          d.rows.push_back(e);
        }
        for (int r = 0; r < c.kout; ++r) {
          RowDesc e = s.rows[c.out_rows + r];
          e.off += b[SLOT_Y];
          d.rows.push_back(e);
        }
        c.in_rows = (int32_t)d.rows.size() - c.kin - c.kout;
        c.out_rows = (int32_t)d.rows.size() - c.kout;
        c.c_off += b[SLOT_C];
        d.chan.push_back(c);
      }
      d.flops += s.flops;
      d.algo_flops += s.algo_flops;
    }
  }
  for (int64_t m = 0; m < n; ++m) {
    const int64_t vb = base[m][SLOT_IN];
    for (CopyDesc c : members[m]->pack_in) {
      c.src += vb;
      c.dst += vb;
      p->pack_in.push_back(c);
    }
    for (CopyDesc c : members[m]->pack_out) {
      c.src += vb;
      c.dst += vb;
      p->pack_out.push_back(c);
    }
  }
  int64_t slot_elems[NSLOTS] = {0, vs, vs, ls, rs, xs, ys, cs};
  for (int i = 0; i < 3; ++i) validate_stage(p->st[i], slot_elems);
  // arenas: operators copied from the members, intermediates fresh
  hipStream_t st = ctx->stream;
  const size_t eb = sizeof(double);
  p->X.ensure(std::max<int64_t>(xs, 1) * eb);
  p->Y.ensure(std::max<int64_t>(ys, 1) * eb);
  p->C.ensure(std::max<int64_t>(cs, 1) * eb);
  p->Lint.ensure(std::max<int64_t>(ls, 1) * eb);
  p->Rint.ensure(std::max<int64_t>(rs, 1) * eb);
  QLG_HIP(hipMemsetAsync(p->Y.p, 0, std::max<int64_t>(ys, 1) * eb, st));   // Y entries no channel writes stay zero
  for (int64_t m = 0; m < n; ++m) {
    const Plan &q = *members[m];
    const auto &b = base[m];
    if (q.c_elems)
      QLG_HIP(hipMemcpyAsync(p->C.as<double>() + b[SLOT_C], q.C.p, q.c_elems * eb, hipMemcpyDeviceToDevice, st));
    if (q.lint_elems)
      QLG_HIP(hipMemcpyAsync(p->Lint.as<double>() + b[SLOT_L], q.Lint.p, q.lint_elems * eb, hipMemcpyDeviceToDevice,
                             st));
    if (q.rint_elems)
      QLG_HIP(hipMemcpyAsync(p->Rint.as<double>() + b[SLOT_R], q.Rint.p, q.rint_elems * eb, hipMemcpyDeviceToDevice,
                             st));
  }
  p->d_pack_in.ensure(std::max<size_t>(1, p->pack_in.size()) * sizeof(CopyDesc));
  p->d_pack_out.ensure(std::max<size_t>(1, p->pack_out.size()) * sizeof(CopyDesc));
  if (!p->pack_in.empty()) {
    QLG_HIP(hipMemcpyAsync(p->d_pack_in.p, p->pack_in.data(), p->pack_in.size() * sizeof(CopyDesc),
                           hipMemcpyHostToDevice, st));
    QLG_HIP(hipMemcpyAsync(p->d_pack_out.p, p->pack_out.data(), p->pack_out.size() * sizeof(CopyDesc),
                           hipMemcpyHostToDevice, st));
  }
  for (int i = 0; i < 3; ++i) {
    if (i == 1)
      p->st[i].make_tiles(members[0]->st[1].bm, members[0]->st[1].bn);
    else
      model_tiles(p->st[i], false);   // replaced by measured shapes on the first matvec
    p->st[i].upload(st);
  }
  p->st[0].algo_bytes = 8.0 * double(ls + vs + xs);
  p->st[1].algo_bytes = 8.0 * double(xs + cs + ys);
  p->st[2].algo_bytes = 8.0 * double(ys + rs + vs);
  QLG_HIP(hipStreamSynchronize(st));
  return p.release();
}

}  // namespace qlg
