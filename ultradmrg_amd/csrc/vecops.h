// Launchers of the Krylov vector kernels (vecops.hip).  Results of reductions
// are written to device memory (dev_out) on the context's stream.
#pragma once
#include "common.h"

namespace qlg {

constexpr int kMaxPartials = 1024;

void vec_dot(Ctx &c, int dtype, int64_t n, const double *x, const double *y, double *dev_out);
void vec_nrm2(Ctx &c, int dtype, int64_t n, const double *x, double *dev_out);
void vec_scal(Ctx &c, int dtype, int64_t n, double ar, double ai, double *x);
void vec_axpy(Ctx &c, int dtype, int64_t n, double ar, double ai, const double *x, double *y);
void vec_lincomb(Ctx &c, int dtype, int64_t n, int k, const double *dev_coefs, const double *const *dev_bases,
                 bool complex_coefs, double br, double bi, bool use_beta, double *y);
void vec_lanczos_residual(Ctx &c, int dtype, int64_t n, double *w, const double *v1, const double *a_dev,
                          const double *v0, const double *b_dev, double *dev_norm_out);

}  // namespace qlg
