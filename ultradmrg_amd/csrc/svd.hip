// Device kernels of the block SVD (svd.cpp): gathering sector matrices into
// padded batches, per-matrix norms and scaling, the Gram matrices' padding,
// the complex batches' Householder rotation, row norms, Newton-Schulz
// updates, completion of zero rows, and the assembly of u, s and vt blocks.
// Complex data is interleaved (re, im); every kernel takes the element size
// in doubles (ed = 1 real, 2 complex).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "svd.h"

namespace qlg {

namespace {
constexpr int VT = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double block_sum(double v) {
  __shared__ double sh[VT / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < VT / 64; ++w) s += sh[w];
  return s;
}
__device__ inline double block_max(double v) {
  __shared__ double sh[VT / 64];
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < VT / 64; ++w) s = fmax(s, sh[w]);
  return s;
}

// Every per-matrix kernel below runs on a grid (matrices, gy): workgroup
// (x, y) handles elements y * VT + t, (y + gy) * VT + t, ... of matrix x, so a
// large matrix is spread over gy workgroups (gy from svd_grid_y).
__device__ inline int64_t first_el() { return (int64_t)blockIdx.y * VT + threadIdx.x; }
__device__ inline int64_t el_stride() { return (int64_t)gridDim.y * VT; }

// element (r, c) of a block's [rows x cols] row-major data goes to row
// rr = r0 + r, column cc = c0 + c of its sector matrix, stored transposed
// (short side first) when tr is set
__global__ __launch_bounds__(VT) void gather_kernel(const SvdBlk *__restrict__ tab, const double *__restrict__ src,
                                                    double *__restrict__ dst, int ed) {
  const SvdBlk b = tab[blockIdx.x];
  const int64_t tot = b.rows * b.cols;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t r = e / b.cols, c = e % b.cols;
    const int64_t rr = b.r0 + r, cc = b.c0 + c;
    const int64_t pos = b.base + (b.tr ? cc * b.ld + rr : rr * b.ld + cc);
    for (int q = 0; q < ed; ++q) dst[pos * ed + q] = src[(b.src + e) * ed + q];
  }
}

// Frobenius norms in two passes: each workgroup's sum of squares of its share
// of matrix i (rows x cols within leading dimension ld) goes to
// part[i * gy + y]; then every workgroup adds matrix i's gy partial sums in
// the same order, divides its share by the norm (norm 0: left as is) and
// workgroup (i, 0) writes scale[i] (1 for norm 0)
__global__ __launch_bounds__(VT) void norm_partial_kernel(const SvdMat *__restrict__ mats, const double *__restrict__ x,
                                                          double *__restrict__ part, int ed) {
  const SvdMat m = mats[blockIdx.x];
  double s = 0.0;
  const int64_t tot = m.rows * m.cols;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t p = m.off + (e / m.cols) * m.ld + e % m.cols;
    for (int q = 0; q < ed; ++q) s += x[p * ed + q] * x[p * ed + q];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) part[(int64_t)blockIdx.x * gridDim.y + blockIdx.y] = s;
}
__global__ __launch_bounds__(VT) void normalise_kernel(const SvdMat *__restrict__ mats, double *__restrict__ x,
                                                       const double *__restrict__ part, double *__restrict__ scale,
                                                       int ed) {
  const SvdMat m = mats[blockIdx.x];
  double s = 0.0;
  for (unsigned y = 0; y < gridDim.y; ++y) s += part[(int64_t)blockIdx.x * gridDim.y + y];
  const double nrm = sqrt(s);
  const double inv = nrm > 0.0 ? 1.0 / nrm : 1.0;
  const int64_t tot = m.rows * m.cols;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t p = m.off + (e / m.cols) * m.ld + e % m.cols;
    for (int q = 0; q < ed; ++q) x[p * ed + q] *= inv;
  }
  if (threadIdx.x == 0 && blockIdx.y == 0) scale[blockIdx.x] = nrm > 0.0 ? nrm : 1.0;
}

// Gram matrix padding: diagonal entries j >= n of the b x b matrix get
// -1 - j / b (distinct values below the spectrum)
__global__ __launch_bounds__(VT) void pad_diag_kernel(const SvdMat *__restrict__ mats, double *__restrict__ g,
                                                      int ed) {
  const SvdMat m = mats[blockIdx.x];   // rows = b, cols = n (valid order)
  for (int64_t j = m.cols + threadIdx.x; j < m.rows; j += VT) {
    const int64_t p = m.off + j * m.ld + j;
    g[p * ed] = -1.0 - (double)j / (double)m.rows;
    if (ed == 2) g[p * ed + 1] = 0.0;
  }
}

// complex Hermitian G (b x b) <- H G H with H = I - 2 u u^H (u unit):
// w = G u, G <- G - 2 u w^H - 2 w u^H + 4 (u^H w) u u^H
__global__ __launch_bounds__(VT) void reflect_kernel(const SvdMat *__restrict__ mats, double2 *__restrict__ g,
                                                     const double2 *__restrict__ u) {
  extern __shared__ double2 w[];
  const SvdMat m = mats[blockIdx.x];
  const int64_t b = m.rows;
  const double2 *ub = u + m.aux;   // this order's reflector vector
  for (int64_t i = threadIdx.x; i < b; i += VT) {
    double2 s{0.0, 0.0};
    for (int64_t j = 0; j < b; ++j) {
      const double2 a = g[m.off + i * m.ld + j], v = ub[j];
      s.x += a.x * v.x - a.y * v.y;
      s.y += a.x * v.y + a.y * v.x;
    }
    w[i] = s;
  }
  __syncthreads();
  double sc = 0.0;   // Re(u^H w)
  for (int64_t i = threadIdx.x; i < b; i += VT) sc += ub[i].x * w[i].x + ub[i].y * w[i].y;
  sc = block_sum(sc);
  for (int64_t e = threadIdx.x; e < b * b; e += VT) {
    const int64_t i = e / b, j = e % b;
    const double2 ui = ub[i], uj = ub[j], wi = w[i], wj = w[j];
    // u_i conj(w_j), w_i conj(u_j), u_i conj(u_j)
    const double2 uw{ui.x * wj.x + ui.y * wj.y, ui.y * wj.x - ui.x * wj.y};
    const double2 wu{wi.x * uj.x + wi.y * uj.y, wi.y * uj.x - wi.x * uj.y};
    const double2 uu{ui.x * uj.x + ui.y * uj.y, ui.y * uj.x - ui.x * uj.y};
    double2 &a = g[m.off + i * m.ld + j];
    a.x += -2.0 * uw.x - 2.0 * wu.x + 4.0 * sc * uu.x;
    a.y += -2.0 * uw.y - 2.0 * wu.y + 4.0 * sc * uu.y;
  }
}

// rows r of the b x b matrix E^H (row j = e_j^H): r <- r H = r - 2 (r u) u^H
__global__ __launch_bounds__(VT) void reflect_rows_kernel(const SvdMat *__restrict__ mats, double2 *__restrict__ e,
                                                          const double2 *__restrict__ u) {
  const SvdMat m = mats[blockIdx.x];
  const int64_t b = m.rows;
  const double2 *ub = u + m.aux;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t r = wave; r < b; r += VT / 64) {
    double2 *row = e + m.off + r * m.ld;
    double tr = 0.0, ti = 0.0;
    for (int64_t j = lane; j < b; j += 64) {
      tr += row[j].x * ub[j].x - row[j].y * ub[j].y;
      ti += row[j].x * ub[j].y + row[j].y * ub[j].x;
    }
    tr = wave_sum(tr);
    ti = wave_sum(ti);
    for (int64_t j = lane; j < b; j += 64) {
      // - 2 t conj(u_j)
      row[j].x -= 2.0 * (tr * ub[j].x + ti * ub[j].y);
      row[j].y -= 2.0 * (ti * ub[j].x - tr * ub[j].y);
    }
  }
}

// norms of rows: one wave per row (rows y * 4 + wave, stepping by 4 gy); out[m.aux + r]
__global__ __launch_bounds__(VT) void row_norms_kernel(const SvdMat *__restrict__ mats, const double *__restrict__ x,
                                                       double *__restrict__ out, int ed) {
  const SvdMat m = mats[blockIdx.x];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t r = (int64_t)blockIdx.y * (VT / 64) + wave; r < m.rows; r += (int64_t)gridDim.y * (VT / 64)) {
    double s = 0.0;
    for (int64_t c = lane; c < m.cols; c += 64) {
      const int64_t p = m.off + r * m.ld + c;
      for (int q = 0; q < ed; ++q) s += x[p * ed + q] * x[p * ed + q];
    }
    s = wave_sum(s);
    if (lane == 0) out[m.aux + r] = sqrt(s);
  }
}

// gather rows with a permutation and a scale: dst row i (of k) = src row perm[i] * f[i]
__global__ __launch_bounds__(VT) void gather_rows_kernel(const SvdRows *__restrict__ jobs, const int32_t *__restrict__ perm,
                                                         const double *__restrict__ f, const double *__restrict__ src,
                                                         double *__restrict__ dst, int ed) {
  const SvdRows j = jobs[blockIdx.x];
  const int64_t tot = j.k * j.cols;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t i = e / j.cols, c = e % j.cols;
    const int64_t sr = perm[j.perm + i];
    const double fi = f ? f[j.perm + i] : 1.0;
    for (int q = 0; q < ed; ++q)
      dst[(j.dst + i * j.dld + c) * ed + q] = src[(j.src + sr * j.sld + c) * ed + q] * fi;
  }
}

// flag[m.aux] = 1 if matrix m (rows x rows at off) holds a NaN; flags are
// cleared by the caller, workgroups of one matrix may store the same 1
__global__ __launch_bounds__(VT) void nan_flags_kernel(const SvdMat *__restrict__ mats, const double *__restrict__ g,
                                                       int32_t *__restrict__ flag, int ed) {
  const SvdMat m = mats[blockIdx.x];
  const int64_t tot = m.rows * m.rows * ed;
  bool bad = false;
  for (int64_t e = first_el(); e < tot; e += el_stride()) bad |= isnan(g[m.off * ed + e]);
  if (__any(bad) && (threadIdx.x & 63) == 0) flag[m.aux] = 1;
}

// the inverse of gather_rows: dst row perm[perm + i] (leading dimension dld)
// = src row i (leading dimension sld), scaled by f[perm + i]
__global__ __launch_bounds__(VT) void scatter_rows_kernel(const SvdRows *__restrict__ jobs,
                                                          const int32_t *__restrict__ perm,
                                                          const double *__restrict__ f, const double *__restrict__ src,
                                                          double *__restrict__ dst, int ed) {
  const SvdRows j = jobs[blockIdx.x];
  const int64_t tot = j.k * j.cols;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t i = e / j.cols, c = e % j.cols;
    const int64_t dr = perm[j.perm + i];
    const double fi = f ? f[j.perm + i] : 1.0;
    for (int q = 0; q < ed; ++q)
      dst[(j.dst + dr * j.dld + c) * ed + q] = src[(j.src + i * j.sld + c) * ed + q] * fi;
  }
}

// Newton-Schulz deviation: max over (i, j) of |A_ij - mask_i delta_ij| for each
// k x k matrix, per workgroup share: dev[i * gy + y] (the caller takes the max over y)
__global__ __launch_bounds__(VT) void ns_dev_kernel(const SvdMat *__restrict__ mats, const double *__restrict__ a,
                                                    const double *__restrict__ mask, double *__restrict__ dev, int ed) {
  const SvdMat m = mats[blockIdx.x];
  double mx = 0.0;
  const int64_t k = m.rows;
  for (int64_t e = first_el(); e < k * k; e += el_stride()) {
    const int64_t i = e / k, j = e % k;
    const int64_t p = m.off + i * m.ld + j;
    double re = a[p * ed], im = ed == 2 ? a[p * ed + 1] : 0.0;
    if (i == j) re -= mask[m.aux + i];
    mx = fmax(mx, sqrt(re * re + im * im));
  }
  mx = block_max(mx);
  if (threadIdx.x == 0) dev[(int64_t)blockIdx.x * gridDim.y + blockIdx.y] = mx;
}

// p <- 1.5 p - 0.5 t over each listed k x cols matrix (contiguous rows)
__global__ __launch_bounds__(VT) void ns_update_kernel(const SvdMat *__restrict__ mats, double *__restrict__ p,
                                                       const double *__restrict__ t, int ed) {
  const SvdMat m = mats[blockIdx.x];
  const int64_t tot = m.rows * m.cols * ed;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t i = m.off * ed + e;
    p[i] = 1.5 * p[i] - 0.5 * t[m.aux * ed + e];
  }
}

// Rows of a k x cols matrix whose mask is 0 (zero singular values) are replaced,
// one after another, by unit vectors orthogonal to every other row: a start
// vector (unit vector on a cycling coordinate) is orthogonalised twice
// against all rows with mask 1 or already completed, then normalised.  Each
// pass takes the rows in chunks of CK: the chunk's coefficients <r_i, r_z>
// are formed in parallel (one wave per row, lanes over the columns) and then
// subtracted at once (block Gram-Schmidt).  One workgroup per matrix.
__global__ __launch_bounds__(VT) void complete_rows_kernel(const SvdMat *__restrict__ mats, double *__restrict__ p,
                                                           double *__restrict__ mask, int ed) {
  constexpr int CK = 512;
  const SvdMat m = mats[blockIdx.x];
  __shared__ double cf[2 * CK];
  const int64_t k = m.rows, n = m.cols;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int64_t start = 0;
  for (int64_t z = 0; z < k; ++z) {
    if (mask[m.aux + z] != 0.0) continue;
    double *rz = p + (m.off + z * m.ld) * ed;
    // try start coordinates until one leaves a nonzero residual
    for (int64_t attempt = 0; attempt < n; ++attempt, ++start) {
      for (int64_t c = threadIdx.x; c < n * ed; c += VT) rz[c] = 0.0;
      __syncthreads();
      if (threadIdx.x == 0) rz[(start % n) * ed] = 1.0;
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass)
        for (int64_t i0 = 0; i0 < k; i0 += CK) {
          const int nc = (int)min<int64_t>(CK, k - i0);
          for (int ii = wave; ii < nc; ii += VT / 64) {
            const int64_t i = i0 + ii;
            double sr = 0.0, si = 0.0;   // <r_i, r_z> = sum conj(r_i) r_z
            if (i != z && mask[m.aux + i] != 0.0) {
              const double *ri = p + (m.off + i * m.ld) * ed;
              for (int64_t c = lane; c < n; c += 64) {
                if (ed == 2) {
                  sr += ri[2 * c] * rz[2 * c] + ri[2 * c + 1] * rz[2 * c + 1];
                  si += ri[2 * c] * rz[2 * c + 1] - ri[2 * c + 1] * rz[2 * c];
                } else {
                  sr += ri[c] * rz[c];
                }
              }
              sr = wave_sum(sr);
              si = wave_sum(si);
            }
            if (lane == 0) {
              cf[2 * ii] = sr;
              cf[2 * ii + 1] = si;
            }
          }
          __syncthreads();
          for (int64_t c = threadIdx.x; c < n; c += VT) {
            double ar = 0.0, ai = 0.0;
            for (int ii = 0; ii < nc; ++ii) {
              const double cr = cf[2 * ii], ci = cf[2 * ii + 1];
              if (cr == 0.0 && ci == 0.0) continue;
              const double *ri = p + (m.off + (i0 + ii) * m.ld) * ed;
              if (ed == 2) {
                ar += cr * ri[2 * c] - ci * ri[2 * c + 1];
                ai += cr * ri[2 * c + 1] + ci * ri[2 * c];
              } else {
                ar += cr * ri[c];
              }
            }
            if (ed == 2) {
              rz[2 * c] -= ar;
              rz[2 * c + 1] -= ai;
            } else {
              rz[c] -= ar;
            }
          }
          __syncthreads();
        }
      double s = 0.0;
      for (int64_t c = threadIdx.x; c < n * ed; c += VT) s += rz[c] * rz[c];
      const double nrm = sqrt(block_sum(s));
      if (nrm > 1e-6) {
        for (int64_t c = threadIdx.x; c < n * ed; c += VT) rz[c] /= nrm;
        ++start;
        break;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) mask[m.aux + z] = 1.0;
    __syncthreads();
  }
}

// Blocked one-sided (Hestenes) Jacobi on rows of R, with every row
// transformation applied to the same rows of E^H as well, so X = E R is kept
// exactly.  A job's rows (idx[job.idx .. + k), largest values first) are cut
// into blocks of JB = 16; a round pairs the blocks round-robin and one
// workgroup takes one block pair (32 rows):
//   1. g = R_pq R_pq^H, the 32 x 32 Gram of the pair's rows, accumulated over
//      the row length in LDS-staged chunks;
//   2. two-sided cyclic Jacobi on g in LDS: the rotation of rows (i, j) with
//      alpha = g_ii, beta = g_jj, gamma = g_ij = |gamma| e^{i phi} is
//        row_i <- c row_i - s e^{i phi} row_j,  row_j <- s row_i + c e^{i phi} row_j,
//      t = s / c the smaller root of t^2 + 2 zeta t - 1 = 0,
//      zeta = (beta - alpha) / (2 |gamma|); g <- T g T^H and the pair
//      transformation T <- T_ij T, until no pair has |gamma| > tol sqrt(alpha beta)
//      (at most 8 inner sweeps), tol = 8 eps sqrt(m): an inner product of two
//      length-m rows is known to about eps sqrt(m) of |r_i| |r_j|;
//   3. the pair's rows of R and of E^H <- T (rows), chunk by chunk.
// A pair that needed a rotation in its first inner sweep flags its job as
// not yet converged (flag[job] = 1); the host repeats sweeps until no flag is
// set.  Relative accuracy: each rotation is decided on |gamma| relative to
// sqrt(alpha beta), as in the scalar method, so small rows are resolved
// against large ones to ~eps of their own norm.
constexpr int JB = 16, JP = 2 * JB, JC = 32;   // block rows, pair rows, staged columns

template <bool CPLX>
struct JElem;
template <>
struct JElem<false> {
  using T = double;
};
template <>
struct JElem<true> {
  using T = double2;
};

__device__ inline double jre(double x) { return x; }
__device__ inline double jre(double2 x) { return x.x; }
__device__ inline double jabs2(double x) { return x * x; }
__device__ inline double jabs2(double2 x) { return x.x * x.x + x.y * x.y; }
__device__ inline double jmul_conj(double a, double b) { return a * b; }
__device__ inline double2 jmul_conj(double2 a, double2 b) {   // a conj(b)
  return double2{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
__device__ inline double jmul(double a, double b) { return a * b; }
__device__ inline double2 jmul(double2 a, double2 b) { return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline double jadd(double a, double b) { return a + b; }
__device__ inline double2 jadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ inline double jscale(double a, double f) { return a * f; }
__device__ inline double2 jscale(double2 a, double f) { return double2{a.x * f, a.y * f}; }
__device__ inline double jconj(double a) { return a; }
__device__ inline double2 jconj(double2 a) { return double2{a.x, -a.y}; }
template <class T>
__device__ inline T jzero();
template <>
__device__ inline double jzero<double>() { return 0.0; }
template <>
__device__ inline double2 jzero<double2>() { return double2{0.0, 0.0}; }
template <class T>
__device__ inline T jone();
template <>
__device__ inline double jone<double>() { return 1.0; }
template <>
__device__ inline double2 jone<double2>() { return double2{1.0, 0.0}; }
template <class T>
__device__ inline T jphase(T g, double ag);
template <>
__device__ inline double jphase<double>(double g, double ag) { return g / ag; }
template <>
__device__ inline double2 jphase<double2>(double2 g, double ag) { return double2{g.x / ag, g.y / ag}; }

template <bool CPLX>
__global__ __launch_bounds__(VT) void jacobi_round_kernel(const SvdJacB *__restrict__ jobs,
                                                          const int32_t *__restrict__ pairs,
                                                          const int32_t *__restrict__ idx, int round,
                                                          typename JElem<CPLX>::T *__restrict__ r,
                                                          typename JElem<CPLX>::T *__restrict__ e,
                                                          int32_t *__restrict__ flag) {
  using T = typename JElem<CPLX>::T;
  const int job = pairs[2 * blockIdx.x], pi = pairs[2 * blockIdx.x + 1];
  const SvdJacB j = jobs[job];
  const int nbp = j.nblk + (j.nblk & 1);
  if (round >= nbp - 1) return;   // this job's sweep is shorter
  int a, b;
  if (pi == 0) {
    a = nbp - 1;
    b = round;
  } else {
    a = (round + pi) % (nbp - 1);
    b = (round - pi + nbp - 1) % (nbp - 1);
  }
  // a block paired with the dummy block (odd block count) is processed alone:
  // the dummy's rows are absent (rowid -1), its own pairs still rotate
  __shared__ T g[JP][JP + 1], tr[JP][JP + 1], st[JP][JC + 1];
  __shared__ int rowid[JP];
  __shared__ double rc[JP / 2], rs[JP / 2];
  __shared__ T rph[JP / 2];
  __shared__ int rp[JP / 2], rq[JP / 2];
  __shared__ int any_rot, any_first;
  const int tid = threadIdx.x;
  if (tid < JP) {
    const int blk = tid < JB ? a : b, slot = blk * JB + (tid % JB);
    rowid[tid] = slot < j.k ? idx[j.idx + slot] : -1;
  }
  if (tid == 0) any_first = 0;
  __syncthreads();
  const double tol = fmax(1e-15, 8.0 * 2.220446049250313e-16 * sqrt((double)j.m));
  // ---- 1. Gram of the pair's rows ------------------------------------------------
  T acc[JP * JP / VT];
#pragma unroll
  for (int u = 0; u < JP * JP / VT; ++u) acc[u] = jzero<T>();
  for (int64_t c0 = 0; c0 < j.m; c0 += JC) {
    for (int q = tid; q < JP * JC; q += VT) {
      const int i = q / JC, c = q % JC;
      const int row = rowid[i];
      st[i][c] = (row >= 0 && c0 + c < j.m) ? r[j.roff + (int64_t)row * j.rld + c0 + c] : jzero<T>();
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < JP * JP / VT; ++u) {
      const int q = tid + u * VT, i = q / JP, k = q % JP;
      T s = acc[u];
      for (int c = 0; c < JC; ++c) s = jadd(s, jmul_conj(st[i][c], st[k][c]));
      acc[u] = s;
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < JP * JP / VT; ++u) {
    const int q = tid + u * VT, i = q / JP, k = q % JP;
    g[i][k] = acc[u];
    tr[i][k] = i == k ? jone<T>() : jzero<T>();
  }
  __syncthreads();
  // exact Hermitian symmetry: the lower triangle mirrors the upper one
  for (int q = tid; q < JP * JP; q += VT) {
    const int i = q / JP, k = q % JP;
    if (i > k) g[i][k] = jconj(g[k][i]);
  }
  __syncthreads();
  // ---- 2. cyclic Jacobi on g, accumulating T -------------------------------------------
  for (int sweep = 0; sweep < 8; ++sweep) {
    if (tid == 0) any_rot = 0;
    __syncthreads();
    for (int step = 0; step < JP - 1; ++step) {
      if (tid < JP / 2) {
        int p, q;
        if (tid == 0) {
          p = JP - 1;
          q = step;
        } else {
          p = (step + tid) % (JP - 1);
          q = (step - tid + JP - 1) % (JP - 1);
        }
        const double al = jre(g[p][p]), be = jre(g[q][q]);
        const T gam = g[p][q];
        const double ag = sqrt(jabs2(gam));
        double cs = 1.0, sn = 0.0;
        T ph = jone<T>();
        if (al > 0.0 && be > 0.0 && ag > tol * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ag);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          cs = 1.0 / sqrt(1.0 + t * t);
          sn = cs * t;
          ph = jphase<T>(gam, ag);
          any_rot = 1;
        }
        rp[tid] = p;
        rq[tid] = q;
        rc[tid] = cs;
        rs[tid] = sn;
        rph[tid] = ph;
      }
      __syncthreads();
      // rows of g and T: row_p <- c row_p - s ph row_q, row_q <- s row_p + c ph row_q
      for (int w = tid; w < (JP / 2) * JP; w += VT) {
        const int pr = w / JP, col = w % JP;
        if (rs[pr] == 0.0) continue;
        const int p = rp[pr], q = rq[pr];
        const double cs = rc[pr], sn = rs[pr];
        const T ph = rph[pr];
        const T gp = g[p][col], gq = jmul(ph, g[q][col]);
        g[p][col] = jadd(jscale(gp, cs), jscale(gq, -sn));
        g[q][col] = jadd(jscale(gp, sn), jscale(gq, cs));
        const T tp = tr[p][col], tq = jmul(ph, tr[q][col]);
        tr[p][col] = jadd(jscale(tp, cs), jscale(tq, -sn));
        tr[q][col] = jadd(jscale(tp, sn), jscale(tq, cs));
      }
      __syncthreads();
      // columns of g: col_p <- c col_p - s conj(ph) col_q, col_q <- s col_p + c conj(ph) col_q
      for (int w = tid; w < (JP / 2) * JP; w += VT) {
        const int pr = w / JP, row = w % JP;
        if (rs[pr] == 0.0) continue;
        const int p = rp[pr], q = rq[pr];
        const double cs = rc[pr], sn = rs[pr];
        const T ph = jconj(rph[pr]);
        const T gp = g[row][p], gq = jmul(ph, g[row][q]);
        g[row][p] = jadd(jscale(gp, cs), jscale(gq, -sn));
        g[row][q] = jadd(jscale(gp, sn), jscale(gq, cs));
      }
      __syncthreads();
    }
    const int rot = any_rot;
    if (sweep == 0 && tid == 0) any_first = rot;
    __syncthreads();
    if (!rot) break;
  }
  if (!any_first) return;   // the pair was already orthogonal: rows unchanged
  if (tid == 0) flag[job] = 1;
  // ---- 3. rows of R and E^H <- T rows --------------------------------------------------
  for (int which = 0; which < 2; ++which) {
    T *base = which == 0 ? r + j.roff : e + j.eoff;
    const int64_t ld = which == 0 ? j.rld : j.eld, len = which == 0 ? j.m : j.eld;
    for (int64_t c0 = 0; c0 < len; c0 += JC) {
      for (int q = tid; q < JP * JC; q += VT) {
        const int i = q / JC, c = q % JC;
        const int row = rowid[i];
        st[i][c] = (row >= 0 && c0 + c < len) ? base[(int64_t)row * ld + c0 + c] : jzero<T>();
      }
      __syncthreads();
      for (int q = tid; q < JP * JC; q += VT) {
        const int i = q / JC, c = q % JC;
        const int row = rowid[i];
        if (row < 0 || c0 + c >= len) continue;
        T s = jzero<T>();
        for (int l = 0; l < JP; ++l) s = jadd(s, jmul(tr[i][l], st[l][c]));
        base[(int64_t)row * ld + c0 + c] = s;
      }
      __syncthreads();
    }
  }
}

// output assembly: dst[o + r * cols + c] = src value at base + r * rs + c * cs
// (conjugated if conj), one row of the grid per output block
__global__ __launch_bounds__(VT) void assemble_kernel(const SvdOut *__restrict__ jobs, const double *__restrict__ src,
                                                      double *__restrict__ dst, int ed) {
  const SvdOut j = jobs[blockIdx.x];
  const int64_t tot = j.rows * j.cols;
  for (int64_t e = first_el(); e < tot; e += el_stride()) {
    const int64_t r = e / j.cols, c = e % j.cols;
    const int64_t s = j.base + r * j.rs + c * j.cs;
    dst[(j.dst + e) * ed] = src[s * ed];
    if (ed == 2) dst[(j.dst + e) * ed + 1] = j.conj ? -src[s * ed + 1] : src[s * ed + 1];
  }
}

// s blocks: diagonal k x k matrices from the kept singular values
__global__ __launch_bounds__(VT) void diag_kernel(const SvdRows *__restrict__ jobs, const double *__restrict__ sv,
                                                  double *__restrict__ dst) {
  const SvdRows j = jobs[blockIdx.x];
  for (int64_t e = first_el(); e < j.k * j.k; e += el_stride()) {
    const int64_t r = e / j.k, c = e % j.k;
    dst[j.dst + e] = r == c ? sv[j.src + r] : 0.0;
  }
}
}  // namespace

#define QLG_LAUNCH(kern, n, ...)                                                        \
  do {                                                                                  \
    if ((n) > 0) {                                                                      \
      hipLaunchKernelGGL(kern, dim3((unsigned)(n)), dim3(VT), 0, st, __VA_ARGS__);    \
      QLG_HIP(hipGetLastError());                                                       \
    }                                                                                   \
  } while (0)
#define QLG_LAUNCH2(kern, n, gy, ...)                                                           \
  do {                                                                                          \
    if ((n) > 0) {                                                                              \
      hipLaunchKernelGGL(kern, dim3((unsigned)(n), (unsigned)(gy)), dim3(VT), 0, st, __VA_ARGS__); \
      QLG_HIP(hipGetLastError());                                                               \
    }                                                                                           \
  } while (0)

int svd_grid_y(int64_t maxel) {
  // about 16 elements per thread, at most 256 workgroups per matrix
  const int64_t gy = (maxel + 16 * VT - 1) / (16 * VT);
  return (int)std::max<int64_t>(1, std::min<int64_t>(gy, 256));
}

void svd_gather(hipStream_t st, const SvdBlk *tab, int64_t n, int64_t maxel, const double *src, double *dst, int ed) {
  QLG_LAUNCH2(gather_kernel, n, svd_grid_y(maxel), tab, src, dst, ed);
}
void svd_normalise(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, double *x, double *part,
                   double *scale, int ed) {
  const int gy = svd_grid_y(maxel);
  QLG_LAUNCH2(norm_partial_kernel, n, gy, mats, x, part, ed);
  QLG_LAUNCH2(normalise_kernel, n, gy, mats, x, part, scale, ed);
}
void svd_pad_diag(hipStream_t st, const SvdMat *mats, int64_t n, double *g, int ed) {
  QLG_LAUNCH(pad_diag_kernel, n, mats, g, ed);
}
void svd_reflect(hipStream_t st, const SvdMat *mats, int64_t n, int64_t bmax, double *g, const double *u) {
  if (n <= 0) return;
  hipLaunchKernelGGL(reflect_kernel, dim3((unsigned)n), dim3(VT), (size_t)bmax * sizeof(double2), st, mats,
                     reinterpret_cast<double2 *>(g), reinterpret_cast<const double2 *>(u));
  QLG_HIP(hipGetLastError());
}
void svd_reflect_rows(hipStream_t st, const SvdMat *mats, int64_t n, double *e, const double *u) {
  QLG_LAUNCH(reflect_rows_kernel, n, mats, reinterpret_cast<double2 *>(e), reinterpret_cast<const double2 *>(u));
}
void svd_row_norms(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxrows, const double *x, double *out,
                   int ed) {
  const int64_t gy = std::max<int64_t>(1, std::min<int64_t>((maxrows + VT / 64 - 1) / (VT / 64), 1024));
  QLG_LAUNCH2(row_norms_kernel, n, gy, mats, x, out, ed);
}
void svd_nan_flags(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, const double *g, int32_t *flag,
                   int ed) {
  QLG_LAUNCH2(nan_flags_kernel, n, svd_grid_y(maxel), mats, g, flag, ed);
}
void svd_scatter_rows(hipStream_t st, const SvdRows *jobs, int64_t n, int64_t maxel, const int32_t *perm,
                      const double *f, const double *src, double *dst, int ed) {
  QLG_LAUNCH2(scatter_rows_kernel, n, svd_grid_y(maxel), jobs, perm, f, src, dst, ed);
}
void svd_gather_rows(hipStream_t st, const SvdRows *jobs, int64_t n, int64_t maxel, const int32_t *perm,
                     const double *f, const double *src, double *dst, int ed) {
  QLG_LAUNCH2(gather_rows_kernel, n, svd_grid_y(maxel), jobs, perm, f, src, dst, ed);
}
void svd_ns_dev(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, const double *a, const double *mask,
                double *dev, int ed) {
  QLG_LAUNCH2(ns_dev_kernel, n, svd_grid_y(maxel), mats, a, mask, dev, ed);
}
void svd_ns_update(hipStream_t st, const SvdMat *mats, int64_t n, int64_t maxel, double *p, const double *t, int ed) {
  QLG_LAUNCH2(ns_update_kernel, n, svd_grid_y(maxel), mats, p, t, ed);
}
void svd_complete_rows(hipStream_t st, const SvdMat *mats, int64_t n, double *p, double *mask, int ed) {
  QLG_LAUNCH(complete_rows_kernel, n, mats, p, mask, ed);
}
void svd_assemble(hipStream_t st, const SvdOut *jobs, int64_t n, int64_t maxel, const double *src, double *dst,
                  int ed) {
  QLG_LAUNCH2(assemble_kernel, n, svd_grid_y(maxel), jobs, src, dst, ed);
}
void svd_diag(hipStream_t st, const SvdRows *jobs, int64_t n, int64_t maxel, const double *sv, double *dst) {
  QLG_LAUNCH2(diag_kernel, n, svd_grid_y(maxel), jobs, sv, dst);
}
void svd_jacobi_round(hipStream_t st, const SvdJacB *jobs, const int32_t *pairs, int64_t npairs, const int32_t *idx,
                      int round, double *r, double *e, int32_t *flag, int ed) {
  if (npairs <= 0) return;
  if (ed == 2)
    hipLaunchKernelGGL(jacobi_round_kernel<true>, dim3((unsigned)npairs), dim3(VT), 0, st, jobs, pairs, idx, round,
                       reinterpret_cast<double2 *>(r), reinterpret_cast<double2 *>(e), flag);
  else
    hipLaunchKernelGGL(jacobi_round_kernel<false>, dim3((unsigned)npairs), dim3(VT), 0, st, jobs, pairs, idx, round, r,
                       e, flag);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
