// Effective-Hamiltonian plan: sector pairing done once per site, three
// grouped-GEMM stages replayed for every matvec.
//
// Stage 1  (L . psi):   for every L block (a', w, a) and state block
//          (a, s1, s2, b) and every MPO / physical component, one GEMM
//          X[a', b] = L[a', w, :] . psi[:, s1, s2, b]   (M=D_a', K=D_a, N=D_b),
//          written as a contiguous chunk of the channel (a'-sector, b-sector).
// Stage 2  (MPO):       per channel, Y = C . X with C the combined two-site MPO
//          operator sum_y Wl[w,s1,t1,y] Wr[y,s2,t2,z] restricted to the
//          channel's input chunks (w,s1,s2) and output rows (t1,t2,z).
// Stage 3  (. R^T):     per output block (a', t1, t2, b'), the sum over
//          (z, b) segments of Y[a', b] . R[b', z, b]^T.
// The one-site form uses the same three stages with C = W.
#pragma once
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
  DevBuf d_tiles, d_tasks, d_segs, d_crows, d_chan, d_rows;
  int bm = 64, bn = 64;
  double flops = 0;      // useful 2*M*N*K issued by this stage
  double algo_flops = 0; // this stage's share of the reference schedule's FLOPs
  double algo_bytes = 0; // operands read once + results written once (bytes)
  void upload(hipStream_t st);
  void make_tiles(int bm_, int bn_);
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

enum PlanKind { PLAN_HEFF = 0, PLAN_TENSORDOT = 1, PLAN_ENV_LEFT = 2, PLAN_ENV_RIGHT = 3 };

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
  // Inside the library the Krylov vectors use the fused layout: for each left
  // sector a of the state, one row-major matrix [D_a x N_a] whose column
  // ranges are the state's blocks (a, ...).  pack_in / pack_out convert
  // between it and the canonical layout of the boundary.
  std::vector<CopyDesc> pack_in, pack_out;
  DevBuf d_pack_in, d_pack_out;
  DevBuf Lint, Rint;            // environments repacked (ket-major L stacks, bra-major R panels)
  DevBuf vtmp_in, vtmp_out;     // staging for qlg_matvec on canonical vectors
  ~Plan();
  void matvec(const double *in, double *out);        // canonical in / out
  void matvec_fused(const double *in, double *out);  // fused in / out
  void to_fused(const double *canon, double *fused);
  void to_canonical(const double *fused, double *canon);
  void execute(const void *in0, const void *in1, void *out);
};

Plan *build_plan(Ctx *ctx, int nphys, const qlg_bsten *L, const qlg_bsten *const *W, const qlg_bsten *R,
                 const qlg_bsten *psi, int precision);
Plan *build_tensordot_plan(Ctx *ctx, const qlg_bsten *A, const qlg_bsten *B, int k, int conj_a, int conj_b);
Plan *build_env_plan(Ctx *ctx, const qlg_bsten *env, const qlg_bsten *mps, const qlg_bsten *mpo, char side);

void check_pair(const Axis &a, const Axis &b, const char *what);
std::vector<int64_t> sector_offsets(const Axis &a);
std::vector<double> dense_copy(const Ten &t, hipStream_t st, int ed);
void make_tiles_lpt(std::vector<GemmTile> &tiles, const std::vector<std::array<int64_t, 3>> &mnk, int bm, int bn);
void launch_mpo_stage(int bm, int bn, bool cplx, const GemmTile *tiles, const ChanTask *tasks, const RowDesc *rows,
                      const double *C, const double *X, double *Y, int ntiles, hipStream_t st);
void launch_block_copy(bool cplx, const CopyDesc *d, int n, const double *src, double *dst, hipStream_t st);
void validate_stage(const Stage &s, const int64_t *slot_elems);
void autotune_tiles(Stage &s, bool cplx);

void launch_grouped_gemm_f64(int bm, int bn, bool cplx, const GemmTile *tiles, const GemmTask *tasks,
                             const GemmSeg *segs, const Slots &slots, int ntiles, hipStream_t st);

}  // namespace qlg
