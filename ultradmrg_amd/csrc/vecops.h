// Launchers of the Krylov vector kernels (vecops.hip).  Results of reductions
// are written to device memory (dev_out) on the context's stream.
#pragma once
#include "common.h"

namespace qlg {

constexpr int kMaxPartials = 1024;

// skip: optional device flag; when it is set the kernels return at once (a
// Krylov solve that runs ahead of the host and has stopped)
void vec_dot(Ctx &c, int dtype, int64_t n, const double *x, const double *y, double *dev_out,
             const int *skip = nullptr);
void vec_nrm2(Ctx &c, int dtype, int64_t n, const double *x, double *dev_out);
// the same reductions with caller-provided partials (kMaxPartials * 2 doubles) and stream
void dot_into(hipStream_t st, int dtype, int64_t n, const double *x, const double *y, double *partials,
              double *dev_out, const int *skip);
void nrm2_into(hipStream_t st, int dtype, int64_t n, const double *x, double *partials, double *dev_out);
// x <- x / *nrm with *nrm read on the device
void vec_scal_inv_dev(Ctx &c, int dtype, int64_t n, const double *nrm, double *x, const int *skip);
void vec_scal(Ctx &c, int dtype, int64_t n, double ar, double ai, double *x);
void vec_axpy(Ctx &c, int dtype, int64_t n, double ar, double ai, const double *x, double *y);
void vec_lincomb(Ctx &c, int dtype, int64_t n, int k, const double *dev_coefs, const double *const *dev_bases,
                 bool complex_coefs, double br, double bi, bool use_beta, double *y);
void vec_lanczos_residual(Ctx &c, int dtype, int64_t n, double *w, const double *v1, const double *a_dev,
                          const double *v0, const double *b_dev, double *dev_norm_out, const int *skip = nullptr);

// Segmented (batched-chain) forms, real data.  A SegTable lives in device
// memory: chunks[nchunks] of at most kSegChunk elements, each inside one
// segment, ordered by segment; chunk_begin[nseg + 1] indexes each segment's
// chunks.  Per-segment coefficients and results are device arrays of nseg.
constexpr int64_t kSegChunk = 8192;
struct SegChunk {
  int64_t begin, end;
  int32_t seg, pad;
};
struct SegTable {
  const SegChunk *chunks = nullptr;
  const int32_t *chunk_begin = nullptr;
  int nchunks = 0, nseg = 0;
};
void seg_dot(hipStream_t st, const SegTable &t, const double *x, const double *y, double *partials, double *out,
             bool sqrt_result);
void seg_scal(hipStream_t st, const SegTable &t, const double *alpha, double *x);
void seg_lanczos_residual(hipStream_t st, const SegTable &t, double *w, const double *v1, const double *a,
                          const double *v0, const double *b, double *partials, double *norm_out);
void seg_lincomb(hipStream_t st, const SegTable &t, int k, const double *coefs, const double *const *bases,
                 double *y);

}  // namespace qlg
