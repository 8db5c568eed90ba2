// Krylov vector kernels: deterministic two-pass reductions (fixed grid,
// per-block partials reduced in index order by one workgroup), streaming
// axpy / scale / linear-combination, and the fused Lanczos residual update
// gamma = w - a*v1 - b*v0 with its squared norm.  All are HBM-bound.
#include <hip/hip_runtime.h>

#include "common.h"
#include "tile_io.h"
#include "vecops.h"

namespace qlg {

constexpr int VT = 256;  // threads per block

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum of NC components; result valid in thread 0
template <int NC>
__device__ inline void block_sum(double (&v)[NC], double (*sh)[VT / 64]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    v[c] = wave_sum(v[c]);
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

// sum_i x_i * conj(y_i); CPLX: vectors of n complex values stored as 2n doubles
template <bool CPLX>
__global__ __launch_bounds__(VT) void dot_partial_kernel(int64_t n, const double *__restrict__ x,
                                                         const double *__restrict__ y,
                                                         double *__restrict__ partials, const int *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int NC = CPLX ? 2 : 1;
  __shared__ double sh[NC][VT / 64];
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += stride) {
    if constexpr (CPLX) {
      const double2 a = reinterpret_cast<const double2 *>(x)[i];
      const double2 b = reinterpret_cast<const double2 *>(y)[i];
      acc[0] += a.x * b.x + a.y * b.y;
      acc[1] += a.y * b.x - a.x * b.y;
    } else {
      acc[0] += x[i] * y[i];
    }
  }
  block_sum<NC>(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) partials[(int64_t)blockIdx.x * NC + c] = acc[c];
}

template <int NC>
__global__ __launch_bounds__(VT) void reduce_partials_kernel(const double *__restrict__ partials, int g,
                                                             double *__restrict__ out, int sqrt_first,
                                                             const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sh[NC][VT / 64];
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (int i = threadIdx.x; i < g; i += VT)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += partials[(int64_t)i * NC + c];
  block_sum<NC>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = acc[c];
    if (sqrt_first) out[0] = sqrt(acc[0]);
  }
}

// x <- x / *nrm (real scale; complex data as 2n doubles), skipped once the solve has stopped
__global__ __launch_bounds__(VT) void scal_inv_dev_kernel(int64_t nd, const double *__restrict__ nrm,
                                                          double *__restrict__ x, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const double alpha = 1.0 / *nrm;
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nd; i += stride) x[i] *= alpha;
}

__global__ __launch_bounds__(VT) void scal_kernel(int64_t nd, double alpha, double *__restrict__ x) {
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nd; i += stride) x[i] *= alpha;
}

// x <- (ar + i ai) * x for complex x (n complex values)
__global__ __launch_bounds__(VT) void zscal_kernel(int64_t n, double ar, double ai, double *__restrict__ x) {
  const int64_t stride = (int64_t)gridDim.x * VT;
  double2 *z = reinterpret_cast<double2 *>(x);
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += stride) {
    double2 v = z[i];
    z[i] = double2{ar * v.x - ai * v.y, ar * v.y + ai * v.x};
  }
}

__global__ __launch_bounds__(VT) void axpy_kernel(int64_t nd, double alpha, const double *__restrict__ x,
                                                  double *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nd; i += stride) y[i] += alpha * x[i];
}

__global__ __launch_bounds__(VT) void zaxpy_kernel(int64_t n, double ar, double ai, const double *__restrict__ x,
                                                   double *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * VT;
  const double2 *a = reinterpret_cast<const double2 *>(x);
  double2 *b = reinterpret_cast<double2 *>(y);
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += stride) {
    double2 v = a[i], w = b[i];
    b[i] = double2{w.x + ar * v.x - ai * v.y, w.y + ar * v.y + ai * v.x};
  }
}

// y <- beta*y + sum_j c_j v_j  (real coefficients, real or complex data as doubles)
__global__ __launch_bounds__(VT) void lincomb_kernel(int64_t nd, int k, const double *__restrict__ coefs,
                                                     const double *const *__restrict__ bases, double beta,
                                                     int use_beta, double *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nd; i += stride) {
    double s = use_beta ? beta * y[i] : 0.0;
    // unrolled so that eight bases' loads are in flight together (the sum keeps its order)
#pragma unroll 8
    for (int j = 0; j < k; ++j) s += coefs[j] * global_ptr(bases[j])[i];
    y[i] = s;
  }
}

// complex coefficients (c_j = coefs[2j] + i coefs[2j+1]) over complex vectors
__global__ __launch_bounds__(VT) void zlincomb_kernel(int64_t n, int k, const double *__restrict__ coefs,
                                                      const double *const *__restrict__ bases, double br,
                                                      double bi, int use_beta, double *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * VT;
  double2 *yz = reinterpret_cast<double2 *>(y);
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += stride) {
    double sr = 0.0, si = 0.0;
    if (use_beta) {
      double2 v = yz[i];
      sr = br * v.x - bi * v.y;
      si = br * v.y + bi * v.x;
    }
#pragma unroll 8
    for (int j = 0; j < k; ++j) {
      const double2 v = global_c128(bases[j], i);
      const double cr = coefs[2 * j], ci = coefs[2 * j + 1];
      sr += cr * v.x - ci * v.y;
      si += cr * v.y + ci * v.x;
    }
    yz[i] = double2{sr, si};
  }
}

// w <- w - a*v1 - b*v0 (v0 may be null), partial sums of |w|^2.  The
// coefficients are read from device memory so that no host round trip sits
// between the dot product that produced a and this update.
__global__ __launch_bounds__(VT) void lanczos_residual_kernel(int64_t nd, double *__restrict__ w,
                                                              const double *__restrict__ v1,
                                                              const double *__restrict__ a_ptr,
                                                              const double *__restrict__ v0,
                                                              const double *__restrict__ b_ptr,
                                                              double *__restrict__ partials,
                                                              const int *__restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sh[1][VT / 64];
  double acc[1] = {0.0};
  const double a = *a_ptr;
  const double b = v0 ? *b_ptr : 0.0;
  const int64_t stride = (int64_t)gridDim.x * VT;
  for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < nd; i += stride) {
    double g = w[i] - a * v1[i];
    if (v0) g -= b * v0[i];
    w[i] = g;
    acc[0] += g * g;
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// ---------------------------------------------------------------------------
static int grid_for(int64_t n) {
  int64_t g = (n + 4 * VT - 1) / (4 * VT);
  if (g < 1) g = 1;
  if (g > kMaxPartials) g = kMaxPartials;
  return (int)g;
}

void dot_into(hipStream_t st, int dtype, int64_t n, const double *x, const double *y, double *part,
              double *dev_out, const int *skip) {
  const bool cplx = dtype == QLG_C128;
  const int g = grid_for(n);
  if (cplx) {
    hipLaunchKernelGGL(dot_partial_kernel<true>, dim3(g), dim3(VT), 0, st, n, x, y, part, skip);
    hipLaunchKernelGGL(reduce_partials_kernel<2>, dim3(1), dim3(VT), 0, st, part, g, dev_out, 0, skip);
  } else {
    hipLaunchKernelGGL(dot_partial_kernel<false>, dim3(g), dim3(VT), 0, st, n, x, y, part, skip);
    hipLaunchKernelGGL(reduce_partials_kernel<1>, dim3(1), dim3(VT), 0, st, part, g, dev_out, 0, skip);
  }
  QLG_HIP(hipGetLastError());
}

void nrm2_into(hipStream_t st, int dtype, int64_t n, const double *x, double *part, double *dev_out) {
  const int64_t nd = n * elem_doubles(dtype);
  const int g = grid_for(nd);
  hipLaunchKernelGGL(dot_partial_kernel<false>, dim3(g), dim3(VT), 0, st, nd, x, x, part, nullptr);
  hipLaunchKernelGGL(reduce_partials_kernel<1>, dim3(1), dim3(VT), 0, st, part, g, dev_out, 1, nullptr);
  QLG_HIP(hipGetLastError());
}

void vec_dot(Ctx &c, int dtype, int64_t n, const double *x, const double *y, double *dev_out, const int *skip) {
  dot_into(c.stream, dtype, n, x, y, c.partials.as<double>(), dev_out, skip);
}

void vec_nrm2(Ctx &c, int dtype, int64_t n, const double *x, double *dev_out) {
  nrm2_into(c.stream, dtype, n, x, c.partials.as<double>(), dev_out);
}

void vec_scal_inv_dev(Ctx &c, int dtype, int64_t n, const double *nrm, double *x, const int *skip) {
  const int64_t nd = n * elem_doubles(dtype);
  hipLaunchKernelGGL(scal_inv_dev_kernel, dim3(grid_for(nd)), dim3(VT), 0, c.stream, nd, nrm, x, skip);
  QLG_HIP(hipGetLastError());
}

void vec_scal(Ctx &c, int dtype, int64_t n, double ar, double ai, double *x) {
  if (dtype == QLG_C128 && ai != 0.0) {
    hipLaunchKernelGGL(zscal_kernel, dim3(grid_for(n)), dim3(VT), 0, c.stream, n, ar, ai, x);
  } else {
    const int64_t nd = n * elem_doubles(dtype);
    hipLaunchKernelGGL(scal_kernel, dim3(grid_for(nd)), dim3(VT), 0, c.stream, nd, ar, x);
  }
  QLG_HIP(hipGetLastError());
}

void vec_axpy(Ctx &c, int dtype, int64_t n, double ar, double ai, const double *x, double *y) {
  if (dtype == QLG_C128 && ai != 0.0) {
    hipLaunchKernelGGL(zaxpy_kernel, dim3(grid_for(n)), dim3(VT), 0, c.stream, n, ar, ai, x, y);
  } else {
    const int64_t nd = n * elem_doubles(dtype);
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(nd)), dim3(VT), 0, c.stream, nd, ar, x, y);
  }
  QLG_HIP(hipGetLastError());
}

void vec_lincomb(Ctx &c, int dtype, int64_t n, int k, const double *dev_coefs, const double *const *dev_bases,
                 bool complex_coefs, double br, double bi, bool use_beta, double *y) {
  if (complex_coefs) {
    hipLaunchKernelGGL(zlincomb_kernel, dim3(grid_for(n)), dim3(VT), 0, c.stream, n, k, dev_coefs, dev_bases, br, bi,
                       (int)use_beta, y);
  } else {
    const int64_t nd = n * elem_doubles(dtype);
    hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(nd)), dim3(VT), 0, c.stream, nd, k, dev_coefs, dev_bases, br,
                       (int)use_beta, y);
  }
  QLG_HIP(hipGetLastError());
}

void vec_lanczos_residual(Ctx &c, int dtype, int64_t n, double *w, const double *v1, const double *a_dev,
                          const double *v0, const double *b_dev, double *dev_norm_out, const int *skip) {
  const int64_t nd = n * elem_doubles(dtype);
  const int g = grid_for(nd);
  double *part = c.partials.as<double>();
  hipLaunchKernelGGL(lanczos_residual_kernel, dim3(g), dim3(VT), 0, c.stream, nd, w, v1, a_dev, v0, b_dev, part,
                     skip);
  hipLaunchKernelGGL(reduce_partials_kernel<1>, dim3(1), dim3(VT), 0, c.stream, part, g, dev_norm_out, 1, skip);
  QLG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Segmented forms for batches of independent chains: the vector is a
// concatenation of segments (one per chain), cut into chunks that never
// straddle a segment.  Each workgroup handles one chunk; reductions write one
// partial per chunk and a second kernel sums each segment's partials in chunk
// order (deterministic).  Coefficients are per-segment device arrays.

__global__ __launch_bounds__(VT) void seg_dot_kernel(const SegChunk *__restrict__ ch, const double *__restrict__ x,
                                                     const double *__restrict__ y, double *__restrict__ partials) {
  __shared__ double sh[1][VT / 64];
  const SegChunk c = ch[blockIdx.x];
  double acc[1] = {0.0};
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += VT) acc[0] += x[i] * y[i];
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(VT) void seg_reduce_kernel(const int32_t *__restrict__ chunk_begin,
                                                        const double *__restrict__ partials,
                                                        double *__restrict__ out, int sqrt_result) {
  __shared__ double sh[1][VT / 64];
  const int s = blockIdx.x;
  double acc[1] = {0.0};
  for (int i = chunk_begin[s] + threadIdx.x; i < chunk_begin[s + 1]; i += VT) acc[0] += partials[i];
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) out[s] = sqrt_result ? sqrt(acc[0]) : acc[0];
}

__global__ __launch_bounds__(VT) void seg_scal_kernel(const SegChunk *__restrict__ ch,
                                                      const double *__restrict__ alpha, double *__restrict__ x) {
  const SegChunk c = ch[blockIdx.x];
  const double a = alpha[c.seg];
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += VT) x[i] *= a;
}

__global__ __launch_bounds__(VT) void seg_residual_kernel(const SegChunk *__restrict__ ch, double *__restrict__ w,
                                                          const double *__restrict__ v1,
                                                          const double *__restrict__ a,
                                                          const double *__restrict__ v0,
                                                          const double *__restrict__ b,
                                                          double *__restrict__ partials) {
  __shared__ double sh[1][VT / 64];
  const SegChunk c = ch[blockIdx.x];
  const double ac = a[c.seg], bc = v0 ? b[c.seg] : 0.0;
  double acc[1] = {0.0};
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += VT) {
    double g = w[i] - ac * v1[i];
    if (v0) g -= bc * v0[i];
    w[i] = g;
    acc[0] += g * g;
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(VT) void seg_lincomb_kernel(const SegChunk *__restrict__ ch, int k, int nseg,
                                                         const double *__restrict__ coefs,
                                                         const double *const *__restrict__ bases,
                                                         double *__restrict__ y) {
  const SegChunk c = ch[blockIdx.x];
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += VT) {
    double s = 0.0;
#pragma unroll 8
    for (int j = 0; j < k; ++j) s += coefs[(int64_t)j * nseg + c.seg] * global_ptr(bases[j])[i];
    y[i] = s;
  }
}

void seg_dot(hipStream_t st, const SegTable &t, const double *x, const double *y, double *partials, double *out,
             bool sqrt_result) {
  hipLaunchKernelGGL(seg_dot_kernel, dim3(t.nchunks), dim3(VT), 0, st, t.chunks, x, y, partials);
  hipLaunchKernelGGL(seg_reduce_kernel, dim3(t.nseg), dim3(VT), 0, st, t.chunk_begin, partials, out,
                     (int)sqrt_result);
  QLG_HIP(hipGetLastError());
}

void seg_scal(hipStream_t st, const SegTable &t, const double *alpha, double *x) {
  hipLaunchKernelGGL(seg_scal_kernel, dim3(t.nchunks), dim3(VT), 0, st, t.chunks, alpha, x);
  QLG_HIP(hipGetLastError());
}

void seg_lanczos_residual(hipStream_t st, const SegTable &t, double *w, const double *v1, const double *a,
                          const double *v0, const double *b, double *partials, double *norm_out) {
  hipLaunchKernelGGL(seg_residual_kernel, dim3(t.nchunks), dim3(VT), 0, st, t.chunks, w, v1, a, v0, b, partials);
  hipLaunchKernelGGL(seg_reduce_kernel, dim3(t.nseg), dim3(VT), 0, st, t.chunk_begin, partials, norm_out, 1);
  QLG_HIP(hipGetLastError());
}

void seg_lincomb(hipStream_t st, const SegTable &t, int k, const double *coefs, const double *const *bases,
                 double *y) {
  hipLaunchKernelGGL(seg_lincomb_kernel, dim3(t.nchunks), dim3(VT), 0, st, t.chunks, k, t.nseg, coefs, bases, y);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
