// FP64 symmetric tridiagonal eigen-solver and the small-matrix exponential
// exp(-i tau T) e_1 used by the Krylov drivers.
//
// Method: implicit QL iteration with Wilkinson shifts and Givens rotations
// (Bowdler, Martin, Reinsch and Wilkinson, "The QR and QL algorithms for
// symmetric matrices", Numer. Math. 11 (1968); EISPACK tql2), followed by
// an ascending sort of the eigenpairs.  This replaces LAPACK dstev used at
// include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:326 and
// include/qlmps/algorithm/tdvp/lanczos_expmv_solver_impl.h:212.
#include "tridiag.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

namespace qlg {

// Eigenvalues d (ascending) and, if z != nullptr, eigenvectors in the columns
// of the row-major n x n matrix z (z[i*n + j] = component i of vector j).
bool tridiag_eig(int64_t n, const double *a, const double *b, double *d, double *z) {
  if (n <= 0) return true;
  std::vector<double> e(n, 0.0);
  for (int64_t i = 0; i < n; ++i) d[i] = a[i];
  for (int64_t i = 0; i + 1 < n; ++i) e[i] = b[i];
  if (z) {
    for (int64_t i = 0; i < n * n; ++i) z[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) z[i * n + i] = 1.0;
  }
  const double eps = std::numeric_limits<double>::epsilon();
  for (int64_t l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int64_t m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) return false;
      // Wilkinson shift from the leading 2x2 block
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int64_t i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        double f = s * e[i];
        const double bb = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * bb;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - bb;
        if (z) {
          for (int64_t k = 0; k < n; ++k) {
            f = z[k * n + i + 1];
            z[k * n + i + 1] = s * z[k * n + i] + c * f;
            z[k * n + i] = c * z[k * n + i] - s * f;
          }
        }
      }
      if (underflow) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // ascending order
  std::vector<int64_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return d[x] < d[y]; });
  std::vector<double> dd(d, d + n);
  for (int64_t i = 0; i < n; ++i) d[i] = dd[idx[i]];
  if (z) {
    std::vector<double> zz(z, z + n * n);
    for (int64_t r = 0; r < n; ++r)
      for (int64_t j = 0; j < n; ++j) z[r * n + j] = zz[r * n + idx[j]];
  }
  return true;
}

bool tridiag_gs(int64_t n, const double *a, const double *b, bool want_vec, double *eng, double *vec) {
  std::vector<double> d(n), z(want_vec ? n * n : 0);
  if (!tridiag_eig(n, a, b, d.data(), want_vec ? z.data() : nullptr)) return false;
  *eng = d[0];
  if (want_vec)
    for (int64_t i = 0; i < n; ++i) vec[i] = z[i * n];
  return true;
}

// res[i] = sum_j z_ij exp(-i tau lambda_j) z_0j, interleaved complex
bool tridiag_expme1(int64_t n, const double *a, const double *b, double tau, double *res) {
  std::vector<double> d(n), z(n * n);
  if (!tridiag_eig(n, a, b, d.data(), z.data())) return false;
  std::vector<double> tr(n), ti(n);
  for (int64_t j = 0; j < n; ++j) {
    const double ph = -tau * d[j];
    tr[j] = std::cos(ph) * z[j];   // z[0*n + j]
    ti[j] = std::sin(ph) * z[j];
  }
  for (int64_t i = 0; i < n; ++i) {
    double sr = 0, si = 0;
    for (int64_t j = 0; j < n; ++j) {
      sr += z[i * n + j] * tr[j];
      si += z[i * n + j] * ti[j];
    }
    res[2 * i] = sr;
    res[2 * i + 1] = si;
  }
  return true;
}

}  // namespace qlg
