// Block SVD of a block-sparse tensor with truncation (svd.cpp, kernels in
// svd.hip): the qlten SVD call of the reference's two-site updates.
#pragma once
#include "common.h"

namespace qlg {

// one source block -> its place in a sector matrix (elements, row-major):
// position = base + (tr ? cc * ld + rr : rr * ld + cc) with rr = r0 + r, cc = c0 + c
struct SvdBlk {
  int64_t src, rows, cols, base, r0, c0, ld;
  int32_t tr, pad;
};
// a matrix [rows x cols] at element off with leading dimension ld; aux: a
// per-matrix index (reflector offset, output offset)
struct SvdMat {
  int64_t off, rows, cols, ld, aux;
};
// k rows of cols elements: dst row i (leading dimension dld) = src row perm[perm + i]
// (leading dimension sld), scaled by f[perm + i]; diag_kernel reads (src, k, dst)
struct SvdRows {
  int64_t src, sld, dst, dld, k, cols, perm;
};
// output block [rows x cols]: dst[dst + r * cols + c] = src[base + r * rs + c * cs]
struct SvdOut {
  int64_t dst, rows, cols, base, rs, cs;
  int32_t conj, pad;
};

// blocked one-sided Jacobi job: k rows of R (length m, leading dimension rld)
// at roff, listed by index at idx[idx .. idx + k), and the same rows of E^H
// (length and leading dimension eld) at eoff; nblk = ceil(k / 16) blocks
struct SvdJacB {
  int64_t roff, rld, m, eoff, eld, k, idx;
  int32_t nblk, pad;
};
// one round of a sweep: pairs = (job, player) for every block pair slot of
// every job; flag[job] is set when a pair of that job needed a rotation
void svd_jacobi_round(hipStream_t st, const SvdJacB *jobs, const int32_t *pairs, int64_t npairs, const int32_t *idx,
                      int round, double *r, double *e, int32_t *flag, int ed);
// Per-matrix kernels run on (matrices x gy) workgroups; gy = svd_grid_y(largest
// element count of the launch's matrices).  maxel / maxrows: that largest count.
int svd_grid_y(int64_t maxel);
void svd_gather(hipStream_t st, const SvdBlk *tab, int64_t n, int64_t maxel, const double *src, double *dst, int ed);
// part: n * svd_grid_y(maxel) doubles of scratch
void svd_normalise(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, double *x, double *part,
                   double *scale, int ed);
void svd_pad_diag(hipStream_t st, const SvdMat *mats, int64_t n, double *g, int ed);
void svd_reflect(hipStream_t st, const SvdMat *mats, int64_t n, int64_t bmax, double *g, const double *u);
void svd_reflect_rows(hipStream_t st, const SvdMat *mats, int64_t n, double *e, const double *u);
void svd_row_norms(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxrows, const double *x, double *out,
                   int ed);
void svd_gather_rows(hipStream_t st, const SvdRows *jobs, int64_t n, int64_t maxel, const int32_t *perm,
                     const double *f, const double *src, double *dst, int ed);
// flag[mats[i].aux] = 1 if the rows x rows matrix at mats[i].off holds a NaN (flags cleared by the caller)
void svd_nan_flags(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, const double *g, int32_t *flag,
                   int ed);
// dst row perm[perm + i] = f[perm + i] * src row i (SvdRows: src / sld the packed rows, dst / dld the target)
void svd_scatter_rows(hipStream_t st, const SvdRows *jobs, int64_t n, int64_t maxel, const int32_t *perm,
                      const double *f, const double *src, double *dst, int ed);
// dev: n * svd_grid_y(maxel) values, the deviation of matrix i is the max of dev[i * gy .. i * gy + gy)
void svd_ns_dev(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, const double *a, const double *mask,
                double *dev, int ed);
void svd_ns_update(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, double *p, const double *t, int ed);
void svd_complete_rows(hipStream_t st, const SvdMat *mats, int64_t n, double *p, double *mask, int ed);
void svd_assemble(hipStream_t st, const SvdOut *jobs, int64_t n, int64_t maxel, const double *src, double *dst,
                  int ed);
void svd_diag(hipStream_t st, const SvdRows *jobs, int64_t n, int64_t maxel, const double *sv, double *dst);

// the result of qlg_block_svd: the new bond and the three tensors' layouts and data
struct SvdResult {
  Ctx *ctx = nullptr;
  int qn_len = 1, dtype = QLG_F64;
  double trunc_err = 0.0;
  int64_t D = 0;
  std::vector<QN> bond_qn;          // kept sectors of the new bond, ascending charge
  std::vector<int64_t> bond_deg;
  // 0: u (left axes + bond OUT), 1: s (bond IN, bond OUT; real), 2: vt (bond IN + right axes)
  std::vector<int32_t> coords[3];   // flattened block coordinates (rank entries per block)
  std::vector<int64_t> offsets[3];
  int64_t size[3] = {0, 0, 0};
  DevBuf data[3];
};

SvdResult *block_svd(Ctx &c, const qlg_bsten *t, int ldims, const int32_t *ldiv, double trunc_err, int64_t Dmin,
                     int64_t Dmax);

}  // namespace qlg
