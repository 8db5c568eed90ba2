// GPU Krylov drivers (Lanczos ground state, Lanczos exp(-i tau H) psi).
#pragma once
#include "plan.h"

namespace qlg {
void lanczos_gs(Plan &p, double *psi, const qlg_lanczos_params &prm, qlg_lanczos_result &res);
void lanczos_gs_batch(Plan &p, double *psi, const qlg_lanczos_params &prm, qlg_lanczos_result *res);
void lanczos_expmv(Plan &p, double *psi, double tau, const qlg_lanczos_params &prm, int64_t &iters);
}  // namespace qlg
