// Device side of the single-chain Krylov solvers.
//
// Each Lanczos iteration ends with one single-wave kernel that does, in FP64
// on the GPU, what the reference does on the host after every matvec:
//  * ground state (lanczos_vmps_solver_impl.h:228-233): the lowest eigenvalue
//    of the tridiagonal T_{m+1}(a, b) and the stop rule E_prev - E < error,
//    m == dim or m == max_iterations - 1;
//  * exp(-i tau H) psi (lanczos_expmv_solver_impl.h:162-167): the
//    coefficients exp(-i tau T_{m+1}) e_1, their Distance to the previous
//    iteration's, and the stop rule Distance < error, m == dim or
//    m == max_iterations - 1.
// A zero residual norm b_m (an invariant subspace) also stops the solve, with
// the Krylov dimension the reference's breakdown branch returns.  The result
// goes to a control block in device memory; every kernel of the iterations
// the host has already queued behind it reads the block's stop flag first and
// returns at once, so the host only looks at the flag every few iterations.
//
// Methods: the lowest eigenvalue by multisection (64 Sturm counts per round,
// one per lane; Sturm sequences as in Barth, Martin and Wilkinson, Numer.
// Math. 9 (1967)); the full eigen-decomposition of the expmv tridiagonal by
// the implicit QL iteration with Wilkinson shifts (EISPACK tql2), the scalar
// recurrence evaluated by every lane alike and the eigenvector rotations
// spread over the lanes (one row of z per lane, z in LDS).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "krylov_dev.h"
#include "tile_io.h"
#include "vecops.h"

namespace qlg {

namespace {

constexpr int kMaxGsDim = 2048;   // tridiagonal order of the ground-state check (LDS: a and b^2)
constexpr int kZ = 64;            // eigenvector workspace of the expmv check: kZ x kZ doubles in LDS
static_assert(kMaxExpmvDim >= kZ, "");

__global__ void lz_reset_kernel(LzCtl *ctl) {
  if (threadIdx.x == 0) {
    ctl->stop = 0;
    ctl->iters = 0;
    ctl->e_prev = 0.0;
    ctl->energy = 0.0;
    ctl->dist = 0.0;
  }
}

// number of eigenvalues of T below x (Sturm count), with the pivot guard of
// LAPACK's dstebz
__device__ inline int sturm_count(const double *a, const double *b2, int n, double x, double pivmin) {
  double q = a[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int i = 1; i < n; ++i) {
    q = (a[i] - x) - b2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

__global__ __launch_bounds__(64) void lz_gs_check_kernel(const double *__restrict__ ab, LzCtl *ctl, int m,
                                                         int64_t dim, int64_t maxit, double err) {
  if (ctl->stop) return;
  __shared__ double a[kMaxGsDim], b2[kMaxGsDim];
  const int lane = threadIdx.x;
  const int n = m + 1;
  const double bm = ab[2 * m + 1];   // norm of the residual of this iteration
  if (m == 0) {
    if (lane == 0) {
      ctl->e_prev = ab[0];
      if (bm == 0.0) {
        ctl->stop = 1;
        ctl->iters = 1;
        ctl->energy = ab[0];
      }
    }
    return;
  }
  double bmax2 = 0.0, lo = DBL_MAX, hi = -DBL_MAX;
  for (int i = lane; i < n; i += 64) {
    a[i] = ab[2 * i];
    const double bl = i > 0 ? fabs(ab[2 * i - 1]) : 0.0, br = i + 1 < n ? fabs(ab[2 * i + 1]) : 0.0;
    if (i + 1 < n) {
      b2[i] = br * br;
      bmax2 = fmax(bmax2, br * br);
    }
    lo = fmin(lo, a[i] - bl - br);
    hi = fmax(hi, a[i] + bl + br);
  }
  for (int o = 32; o > 0; o >>= 1) {
    bmax2 = fmax(bmax2, __shfl_xor(bmax2, o, 64));
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  __syncthreads();
  const double pivmin = DBL_MIN * fmax(1.0, bmax2);
  const double pad = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin;
  lo -= pad;
  hi += pad;
  // invariant: count(lo) == 0, count(hi) >= 1
  for (int round = 0; round < 40; ++round) {
    if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    const bool hit = sturm_count(a, b2, n, x, pivmin) >= 1;
    const unsigned long long mask = __ballot(hit);
    double nlo, nhi;
    if (mask == 0ull) {
      nlo = __shfl(x, 63, 64);
      nhi = hi;
    } else {
      const int j = __ffsll((long long)mask) - 1;
      nhi = __shfl(x, j, 64);
      nlo = j > 0 ? __shfl(x, j - 1, 64) : lo;
    }
    if (nlo == lo && nhi == hi) break;   // no representable progress left
    lo = nlo;
    hi = nhi;
  }
  const double e = 0.5 * (lo + hi);
  if (lane == 0) {
    const bool stop = (ctl->e_prev - e) < err || (int64_t)m == dim || (int64_t)m == maxit - 1 || bm == 0.0;
    if (stop) {
      ctl->stop = 1;
      ctl->iters = n;
      ctl->energy = e;
    } else {
      ctl->e_prev = e;
    }
  }
}

__global__ __launch_bounds__(64) void lz_expmv_check_kernel(const double *__restrict__ ab, LzCtl *ctl,
                                                            double *__restrict__ cur, double *__restrict__ prev,
                                                            int m, int64_t dim, int64_t maxit, double err,
                                                            double tau) {
  if (ctl->stop) return;
  __shared__ double z[kZ * kZ];
  __shared__ double d[kZ], e[kZ], tr[kZ], ti[kZ];
  __shared__ int perm[kZ];
  const int lane = threadIdx.x;
  const int n = m + 1;
  const int ncoef = 2 * (int)(maxit + 1);
  const double bm = ab[2 * m + 1];
  if (m == 0) {
    // coefficients of the breakdown case exp(-i tau a_0) e_1; the previous
    // coefficients start as zeros
    for (int i = lane; i < ncoef; i += 64) {
      cur[i] = 0.0;
      prev[i] = 0.0;
    }
    if (lane == 0) {
      cur[0] = cos(-tau * ab[0]);
      cur[1] = sin(-tau * ab[0]);
      if (bm == 0.0) {
        ctl->stop = 1;
        ctl->iters = 1;
      }
    }
    return;
  }
  // T_{m+1}: diagonal a_0..a_m, off-diagonal b_0..b_{m-1}
  if (lane < n) {
    d[lane] = ab[2 * lane];
    e[lane] = lane + 1 < n ? ab[2 * lane + 1] : 0.0;
  }
  for (int k = lane; k < n * n; k += 64) z[k] = (k / n == k % n) ? 1.0 : 0.0;
  __syncthreads();
  // implicit QL (tql2); every lane runs the scalar recurrence on the shared
  // d and e, lane k < n applies the rotations to row k of z
  bool ok = true;
  for (int l = 0; l < n && ok; ++l) {
    int iter = 0;
    while (true) {
      int mm = l;
      for (; mm < n - 1; ++mm) {
        const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
        if (fabs(e[mm]) <= DBL_EPSILON * dd) break;
      }
      if (mm == l) break;
      if (++iter > 60) {
        ok = false;
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      bool underflow = false;
      for (int i = mm - 1; i >= l; --i) {
        double f = s * e[i];
        const double bb = c * e[i];
        r = hypot(f, g);
        __syncthreads();
        if (lane == 0) e[i + 1] = r;
        if (r == 0.0) {
          if (lane == 0) {
            d[i + 1] -= p;
            e[mm] = 0.0;
          }
          __syncthreads();
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * bb;
        p = s * r;
        __syncthreads();
        if (lane == 0) d[i + 1] = g + p;
        g = c * r - bb;
        if (lane < n) {
          const double zf = z[lane * n + i + 1], zi = z[lane * n + i];
          z[lane * n + i + 1] = s * zi + c * zf;
          z[lane * n + i] = c * zi - s * zf;
        }
        __syncthreads();
      }
      if (underflow) continue;
      __syncthreads();
      if (lane == 0) {
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
      __syncthreads();
    }
  }
  if (!ok) {   // not converged: stop the solve and report it through iters = -1
    if (lane == 0) {
      ctl->stop = 1;
      ctl->iters = -1;
    }
    return;
  }
  // ascending order, ties by index (a stable sort)
  if (lane < n) {
    int rank = 0;
    const double dj = d[lane];
    for (int i = 0; i < n; ++i) rank += (d[i] < dj || (d[i] == dj && i < lane)) ? 1 : 0;
    perm[rank] = lane;
  }
  __syncthreads();
  if (lane < n) {
    const int jo = perm[lane];
    const double ph = -tau * d[jo];
    tr[lane] = cos(ph) * z[jo];   // z[0 * n + jo]
    ti[lane] = sin(ph) * z[jo];
  }
  __syncthreads();
  double dist2 = 0.0;
  if (lane < n) {
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < n; ++j) {
      const double zz = z[lane * n + perm[j]];
      sr += zz * tr[j];
      si += zz * ti[j];
    }
    cur[2 * lane] = sr;
    cur[2 * lane + 1] = si;
    const double dr = sr - prev[2 * lane], di = si - prev[2 * lane + 1];
    dist2 = dr * dr + di * di;
  }
  for (int o = 32; o > 0; o >>= 1) dist2 += __shfl_xor(dist2, o, 64);
  const double dist = sqrt(dist2);
  const bool stop = dist < err || (int64_t)m == dim || (int64_t)m == maxit - 1 || bm == 0.0;
  if (!stop && lane < n) {
    prev[2 * lane] = cur[2 * lane];
    prev[2 * lane + 1] = cur[2 * lane + 1];
  }
  if (lane == 0) {
    ctl->dist = dist;
    if (stop) {
      ctl->stop = 1;
      ctl->iters = n;
    }
  }
}

// ---- classical Gram-Schmidt against the basis --------------------------------
constexpr int VT = 256;

template <int NC>
__device__ inline void block_sum_to(double (&v)[NC], double (*sh)[VT / 64]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o, 64);
    if (lane == 0) sh[c][wave] = v[c];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double s = 0.0;
      for (int w = 0; w < VT / 64; ++w) s += sh[c][w];
      v[c] = s;
    }
  }
}

// partials[j][block] of <v_j, g> = sum conj(v_j) g; grid (G, k)
template <bool CPLX>
__global__ __launch_bounds__(VT) void multidot_kernel(int64_t n, const double *const *__restrict__ bases,
                                                      const double *__restrict__ g, double *__restrict__ partials,
                                                      const int *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int NC = CPLX ? 2 : 1;
  __shared__ double sh[NC][VT / 64];
  const int j = blockIdx.y;
  const auto *v = global_ptr(bases[j]);
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += stride) {
    if constexpr (CPLX) {
      const double2 x = global_c128(bases[j], i);
      const double2 y = reinterpret_cast<const double2 *>(g)[i];
      acc[0] += x.x * y.x + x.y * y.y;
      acc[1] += x.x * y.y - x.y * y.x;
    } else {
      acc[0] += v[i] * g[i];
    }
  }
  block_sum_to<NC>(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) partials[((int64_t)j * gridDim.x + blockIdx.x) * NC + c] = acc[c];
}

// c[j] = sum over blocks of partials[j][*]; one workgroup per j
template <int NC>
__global__ __launch_bounds__(VT) void reduce_rows_kernel(const double *__restrict__ partials, int G,
                                                         double *__restrict__ c, const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sh[NC][VT / 64];
  const int j = blockIdx.x;
  double acc[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) acc[q] = 0.0;
  for (int b = threadIdx.x; b < G; b += VT)
#pragma unroll
    for (int q = 0; q < NC; ++q) acc[q] += partials[((int64_t)j * G + b) * NC + q];
  block_sum_to<NC>(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NC; ++q) c[j * NC + q] = acc[q];
}

// g <- g - sum_j c_j v_j, partial sums of |g|^2
template <bool CPLX>
__global__ __launch_bounds__(VT) void reorth_update_kernel(int64_t n, const double *const *__restrict__ bases, int k,
                                                           const double *__restrict__ c, double *__restrict__ g,
                                                           double *__restrict__ partials,
                                                           const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sh[1][VT / 64];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += stride) {
    if constexpr (CPLX) {
      double2 y = reinterpret_cast<double2 *>(g)[i];
#pragma unroll 8
      for (int j = 0; j < k; ++j) {
        const double2 x = global_c128(bases[j], i);
        const double cr = c[2 * j], ci = c[2 * j + 1];
        y.x -= cr * x.x - ci * x.y;
        y.y -= cr * x.y + ci * x.x;
      }
      reinterpret_cast<double2 *>(g)[i] = y;
      acc[0] += y.x * y.x + y.y * y.y;
    } else {
      double y = g[i];
#pragma unroll 8
      for (int j = 0; j < k; ++j) y -= c[j] * global_ptr(bases[j])[i];
      g[i] = y;
      acc[0] += y * y;
    }
  }
  block_sum_to<1>(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(VT) void reduce_norm_kernel(const double *__restrict__ partials, int G,
                                                         double *__restrict__ out, const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sh[1][VT / 64];
  double acc[1] = {0.0};
  for (int b = threadIdx.x; b < G; b += VT) acc[0] += partials[b];
  block_sum_to<1>(acc, sh);
  if (threadIdx.x == 0) *out = sqrt(acc[0]);
}

int reorth_grid(int64_t n) {
  int64_t g = (n + 4 * VT - 1) / (4 * VT);
  if (g < 1) g = 1;
  if (g > kMaxPartials) g = kMaxPartials;
  return (int)g;
}

}  // namespace

void lz_reset(hipStream_t st, LzCtl *ctl) {
  hipLaunchKernelGGL(lz_reset_kernel, dim3(1), dim3(64), 0, st, ctl);
  QLG_HIP(hipGetLastError());
}

void lz_gs_check(hipStream_t st, const double *ab, LzCtl *ctl, int m, int64_t dim, int64_t maxit, double err) {
  QLG_CHECK(m + 1 <= kMaxGsDim, "device Lanczos: Krylov dimension beyond the device tridiagonal solver");
  hipLaunchKernelGGL(lz_gs_check_kernel, dim3(1), dim3(64), 0, st, ab, ctl, m, dim, maxit, err);
  QLG_HIP(hipGetLastError());
}

void lz_expmv_check(hipStream_t st, const double *ab, LzCtl *ctl, double *coef_cur, double *coef_prev, double *z,
                    int m, int64_t dim, int64_t maxit, double err, double tau) {
  (void)z;
  QLG_CHECK(m + 1 <= kZ, "device expmv: Krylov dimension beyond the device exponential");
  hipLaunchKernelGGL(lz_expmv_check_kernel, dim3(1), dim3(64), 0, st, ab, ctl, coef_cur, coef_prev, m, dim, maxit,
                     err, tau);
  QLG_HIP(hipGetLastError());
}

void lz_reorth(hipStream_t st, int dtype, int64_t n, const double *const *bases, int k, double *g, double *partials,
               double *c, double *nrm_out, const int *skip) {
  const bool cplx = dtype == QLG_C128;
  const int G = reorth_grid(n);
  if (cplx) {
    hipLaunchKernelGGL(multidot_kernel<true>, dim3(G, k), dim3(VT), 0, st, n, bases, g, partials, skip);
    hipLaunchKernelGGL(reduce_rows_kernel<2>, dim3(k), dim3(VT), 0, st, partials, G, c, skip);
    hipLaunchKernelGGL(reorth_update_kernel<true>, dim3(G), dim3(VT), 0, st, n, bases, k, c, g, partials, skip);
  } else {
    hipLaunchKernelGGL(multidot_kernel<false>, dim3(G, k), dim3(VT), 0, st, n, bases, g, partials, skip);
    hipLaunchKernelGGL(reduce_rows_kernel<1>, dim3(k), dim3(VT), 0, st, partials, G, c, skip);
    hipLaunchKernelGGL(reorth_update_kernel<false>, dim3(G), dim3(VT), 0, st, n, bases, k, c, g, partials, skip);
  }
  hipLaunchKernelGGL(reduce_norm_kernel, dim3(1), dim3(VT), 0, st, partials, G, nrm_out, skip);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
