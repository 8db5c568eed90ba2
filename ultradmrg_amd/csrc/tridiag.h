// FP64 symmetric tridiagonal eigen-solver (implicit QL) and exp(-i tau T) e_1.
#pragma once
#include <cstdint>

namespace qlg {
bool tridiag_eig(int64_t n, const double *a, const double *b, double *d, double *z);
bool tridiag_gs(int64_t n, const double *a, const double *b, bool want_vec, double *eng, double *vec);
bool tridiag_expme1(int64_t n, const double *a, const double *b, double tau, double *res);
}  // namespace qlg
