// Plans: sector pairing done once, grouped-GEMM task lists replayed.
//
// The effective-Hamiltonian plan (heff.cpp) runs three launches per matvec:
//  stage 1  per ket sector a:  T1_a = Lstack_a . Psi_a   (grouped GEMM)
//  stage 2  per channel (a', b): Y = C . X on gathered / scattered rows of
//           T1 and Y, C the combined two-site MPO operator  (mpo_stage.hip)
//  stage 3  per bra sector b': out = Y_b' . R_b'^T        (grouped GEMM)
// Tensordot and environment-update plans (ctr.cpp) reuse the grouped GEMM.
#pragma once
#include <memory>
#include <vector>

#include "common.h"
#include "gemm.h"

namespace qlg {

struct Stage {
  std::vector<GemmTile> tiles;
  std::vector<GemmTask> tasks;
  std::vector<GemmSeg> segs;
  std::vector<int64_t> crows;      // output-row address table (tasks with c_rowtab)
  std::vector<ChanTask> chan;      // MPO-stage channels (instead of tasks)
  std::vector<RowDesc> rows;       // their gathered / scattered rows
  std::vector<ChanTile> ctiles;    // MPO stage: tiles with their channels inlined (upload)
  DevBuf d_tiles, d_tasks, d_segs, d_crows, d_chan, d_rows, d_ctiles;
  int bm = 64, bn = 64;
  double flops = 0;      // useful 2*M*N*K issued by this stage
  double algo_flops = 0; // this stage's share of the reference schedule's FLOPs
  double algo_bytes = 0; // operands read once + results written once (bytes)
  // Split-K (the matvec's stage 3): with `splitk` and `split_on` set, make_tiles cuts tasks
  // whose K is long into parts of at most `kchunk` (a multiple of 64), which
  // run as the launch list ktasks / ksegs and write partial panels to `work`;
  // `red` / `redc` drive the reduce kernel.  tasks / segs stay the unsplit
  // lists (signatures, batch plans, the cost model).
  bool splitk = false;             // the stage may run split-K parts
  bool split_on = false;           // make_tiles splits (the tuner's choice, or QLG_SPLITK)
  int ed = 1;                      // doubles per element (complex: 2)
  int64_t kchunk = 0;              // 0: no task is split
  std::vector<GemmTask> ktasks;
  std::vector<GemmSeg> ksegs;
  std::vector<RedTask> red;
  std::vector<RedChunk> redc;
  int64_t work_elems = 0;
  DevBuf d_red, d_redc, work;
  void upload(hipStream_t st);
  void make_tiles(int bm_, int bn_);
  bool split() const { return kchunk > 0 && !red.empty(); }
};

struct Layout {
  int rank = 0;
  std::vector<Axis> ax;
  std::vector<Coords> blk;
  std::vector<int64_t> off;
  std::unordered_map<Coords, int64_t, CoordsHash> find;
  int64_t size = 0;
  int64_t blk_elems(int64_t b) const {
    int64_t n = 1;
    for (int i = 0; i < rank; ++i) n *= ax[i].deg[blk[b][i]];
    return n;
  }
};

enum PlanKind { PLAN_HEFF = 0, PLAN_TENSORDOT = 1, PLAN_ENV_LEFT = 2, PLAN_ENV_RIGHT = 3, PLAN_TRANSPOSE = 4 };

struct Plan {
  Ctx *ctx = nullptr;
  int kind = PLAN_HEFF;
  int nstages = 3;
  int nphys = 2;
  int dtype = QLG_F64;
  int precision = QLG_PREC_F64;
  Ten L, W[2], R;
  Layout vec;
  Stage st[3];
  DevBuf X, Y, C;
  int64_t x_elems = 0, y_elems = 0, c_elems = 0;
  double algo_flops = 0;
  // Krylov workspace
  DevBuf basis, coef_buf, ptr_buf, scal_buf;
  int64_t basis_vectors = 0;
  // timing
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // tile autotuning of the two GEMM stages (per arithmetic variant)
  bool tiles_pinned = false;
  int tuned_precision = -1;
  void autotune(const double *in, double *out);
  void launch_stage(int i, const Slots &s);
  Slots slots(const double *in, double *out);
  // Inside the library the Krylov vectors use the fused layout: for each left
  // sector a of the state, one row-major matrix [D_a x N_a] whose column
  // ranges are the state's blocks (a, ...).  pack_in / pack_out convert
  // between it and the canonical layout of the boundary.
  std::vector<CopyDesc> pack_in, pack_out;
  DevBuf d_pack_in, d_pack_out;
  DevBuf Lint, Rint;            // environments repacked (ket-major L stacks, bra-major R panels)
  int64_t lint_elems = 0, rint_elems = 0;
  // reduced-precision modes: FP32 copies of the L stacks, R panels and
  // channel operators (made on first use).  T1 reuses X (stage 1 rewrites
  // all of it every matvec); Y32 is a buffer of its own because stage 3
  // reads Y entries that no channel writes, which must stay zero.
  DevBuf Lint32, Rint32, C32, Y32;
  bool reduced_ready = false;
  // real split-BF16 modes: L stacks, R panels and Y as three bf16 piece
  // planes each (h, m, l of the FP32 value; the 3-term form reads h and m)
  DevBuf Lp, Rp, Yp;
  bool planar_ready = false;
  bool planar() const;
  void prepare_reduced();
  DevBuf vtmp_in, vtmp_out;     // staging for qlg_matvec on canonical vectors
  // batch plans: member m's vector occupies [batch_off[m], batch_off[m+1])
  // of the batch vector (canonical and fused layouts alike)
  std::vector<int64_t> batch_off;
  // single-chain Krylov solves run ahead of the host (krylov.cpp): device
  // control block and scalars, the skip flag their matvecs read (Slots::skip),
  // optional full reorthogonalisation, and hipGraphs of iteration chunks
  int reorth = 0;                  // classical Gram-Schmidt passes per iteration (0: off, the reference)
  const int *skip_flag = nullptr;  // set while a run-ahead solve is being queued
  DevBuf lz_ws, reorth_ws;
  double *host_lz = nullptr;       // pinned copies of the control block and the (a, b) scalars
  size_t host_lz_n = 0;
  std::vector<hipGraphExec_t> chunk_graphs;
  std::string chunk_key;
  // hipGraphs of the batched Lanczos steps (index m), captured on first use
  bool use_graphs = false;
  std::vector<hipGraphExec_t> step_graphs;
  int64_t graph_maxit = 0;
  const void *graph_basis = nullptr;
  bool capturing = false;
  void drop_graphs();
  DevBuf seg_ws;                 // batched Lanczos: chunk table, per-chain scalars, partials
  double *host_seg = nullptr;    // pinned per-chain scalars (2 x chains)
  size_t host_seg_n = 0;
  std::vector<std::unique_ptr<DevBuf>> owned;   // operators the plan built itself (super-block form)
  DevBuf perm_jobs;              // transpose plans: one PermJob per block
  int64_t perm_n = 0, perm_maxel = 0;
  ~Plan();
  void matvec(const double *in, double *out);        // canonical in / out
  void matvec_fused(const double *in, double *out);  // fused in / out
  void to_fused(const double *canon, double *fused);
  void to_canonical(const double *fused, double *canon);
  void execute(const void *in0, const void *in1, void *out);
};

Plan *build_plan(Ctx *ctx, int nphys, const qlg_bsten *L, const qlg_bsten *const *W, const qlg_bsten *R,
                 const qlg_bsten *psi, int precision);
Plan *build_transpose_plan(Ctx *ctx, const qlg_bsten *t, const int32_t *perm);
Plan *build_super_block_plan(Ctx *ctx, int64_t ngroups, const int64_t *nleft, const qlg_term *left,
                             const int64_t *nright, const qlg_term *right, const qlg_bsten *psi, int precision);
Plan *build_batch_plan(Ctx *ctx, int64_t n, Plan *const *members);
Plan *build_tensordot_plan(Ctx *ctx, const qlg_bsten *A, const qlg_bsten *B, int k, int conj_a, int conj_b);
Plan *build_env_plan(Ctx *ctx, const qlg_bsten *env, const qlg_bsten *mps, const qlg_bsten *mpo, char side);

void check_pair(const Axis &a, const Axis &b, const char *what);
std::vector<int64_t> sector_offsets(const Axis &a);
std::vector<double> dense_copy(const Ten &t, hipStream_t st, int ed);
int64_t raw_elems(const Ten &t);
std::vector<double> densify(const Ten &t, const double *host_raw, int ed);
void make_tiles_lpt(std::vector<GemmTile> &tiles, const std::vector<std::array<int64_t, 3>> &mnk, int bm, int bn);
void launch_mpo_stage(int bm, int bn, bool cplx, bool f32, int out_planes, int64_t yps, const ChanTile *tiles,
                      const RowDesc *rows, const void *C, const void *X, void *Y, const int *skip, int ntiles,
                      hipStream_t st);
// split-BF16 on pre-split operands (gemm_mixed.hip): stage 1 (A planes, B the
// FP64 vector) or stage 3 (A and B planes, FP64 output); FP64 -> 3 bf16 planes
void launch_grouped_gemm_planar(int mode, int bm, int bn, bool stage3, const GemmTile *tiles, const GemmTask *tasks,
                                const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st);
void launch_split_planes(const double *src, void *dst, int64_t n, hipStream_t st);
void launch_grouped_gemm_mixed(int mode, int bm, int bn, bool b_f64, bool c_f64, bool cplx, const GemmTile *tiles,
                               const GemmTask *tasks, const GemmSeg *segs, const Slots &slots, int ntiles,
                               hipStream_t st);
void launch_f64_to_f32(const double *src, float *dst, int64_t n, hipStream_t st);
void launch_block_copy(bool cplx, const CopyDesc *d, int n, const double *src, double *dst, hipStream_t st);
void validate_stage(const Stage &s, const int64_t *slot_elems);
std::vector<std::pair<double, int>> rank_tile_shapes(const Stage &s, bool cplx, bool thin);
void model_tiles(Stage &s, bool cplx);
std::pair<int, int> tile_shape(int index);
std::string tile_signature(const Stage &s, int stage, int variant);

void launch_grouped_gemm_f64(int bm, int bn, bool cplx, const GemmTile *tiles, const GemmTask *tasks,
                             const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st);
// sums the split-K partial panels of a stage into its outputs (FP64 / complex128)
void launch_splitk_reduce(bool cplx, const RedTask *red, const RedChunk *chunks, int nchunks, const Slots &slots,
                          hipStream_t st);
int64_t pick_kchunk(const std::vector<std::array<int64_t, 3>> &mnk, int bm, int bn);

}  // namespace qlg
