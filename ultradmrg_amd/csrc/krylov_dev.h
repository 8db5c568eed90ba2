// Device side of the single-chain Krylov solvers: the per-iteration
// tridiagonal work (FP64, one workgroup), the stop rule evaluated on the
// device, and the optional full reorthogonalisation kernels.
#pragma once
#include "common.h"

namespace qlg {

// Control block of one solve in device memory.  Once `stop` is set, every
// kernel of later (already queued) iterations returns at once.
struct LzCtl {
  int stop;        // the stop rule has fired
  int iters;       // Krylov dimension at the stop (LanczosRes::iters)
  double e_prev;   // ground state: lowest Ritz value of the previous iteration
  double energy;   // ground state: lowest Ritz value at the stop
  double dist;     // expmv: Distance of the last iteration
  double pad[3];
};

// largest Krylov dimension the device exponential handles (n x n eigenvectors in a workspace)
constexpr int kMaxExpmvDim = 128;

void lz_reset(hipStream_t st, LzCtl *ctl);
// Iteration m (>= 0) of LanczosSolver: ab holds (a_k, b_k) pairs for k <= m;
// the lowest eigenvalue of T_{m+1} by bisection and the stop rule of
// lanczos_vmps_solver_impl.h:230-233 (plus breakdown when b_m == 0).
void lz_gs_check(hipStream_t st, const double *ab, LzCtl *ctl, int m, int64_t dim, int64_t maxit, double err);
// Iteration m of LanczosExpmvSolver: coef_cur <- exp(-i tau T_{m+1}) e_1
// (interleaved complex, zero beyond m), its Distance to coef_prev, the stop
// rule of lanczos_expmv_solver_impl.h:162-167 (plus breakdown); coef_prev
// takes coef_cur when the solve goes on.  z: workspace of kMaxExpmvDim^2 doubles.
void lz_expmv_check(hipStream_t st, const double *ab, LzCtl *ctl, double *coef_cur, double *coef_prev, double *z,
                    int m, int64_t dim, int64_t maxit, double err, double tau);
// Classical Gram-Schmidt of g against bases[0..k): c = V^H g, g <- g - V c,
// then *nrm_out = |g|.  partials: k * kMaxPartials * 2 doubles; c: 2k doubles.
void lz_reorth(hipStream_t st, int dtype, int64_t n, const double *const *bases, int k, double *g,
               double *partials, double *c, double *nrm_out, const int *skip);

}  // namespace qlg
