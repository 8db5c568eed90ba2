// Axis permutation of a block-sparse tensor (qlten's Transpose, e.g.
// two_site_update_finite_tdvp_impl.h:36): every block is rewritten with its
// axes in the new order.  One row of the grid per block; workgroup (b, y)
// walks the block's elements in source order from y * VT in steps of
// gy * VT, so reads are coalesced and each element is written once.
#include <hip/hip_runtime.h>

#include "transpose.h"

namespace qlg {

namespace {
constexpr int VT = 256;

__global__ __launch_bounds__(VT) void transpose_blocks_kernel(const PermJob *__restrict__ jobs,
                                                              const double *__restrict__ src,
                                                              double *__restrict__ dst, int ed) {
  const PermJob j = jobs[blockIdx.x];
  for (int64_t e = (int64_t)blockIdx.y * VT + threadIdx.x; e < j.n; e += (int64_t)gridDim.y * VT) {
    int64_t r = e, d = 0;
    for (int k = j.rank - 1; k >= 0; --k) {
      const int64_t i = r % j.dim[k];
      r /= j.dim[k];
      d += i * j.dstride[k];
    }
    for (int q = 0; q < ed; ++q) dst[(j.dst + d) * ed + q] = src[(j.src + e) * ed + q];
  }
}
}  // namespace

void launch_transpose(hipStream_t st, const PermJob *jobs, int64_t n, int64_t maxel, const double *src, double *dst,
                      int ed) {
  if (n <= 0) return;
  const int64_t gy = std::max<int64_t>(1, std::min<int64_t>((maxel + 16 * VT - 1) / (16 * VT), 1024));
  hipLaunchKernelGGL(transpose_blocks_kernel, dim3((unsigned)n, (unsigned)gy), dim3(VT), 0, st, jobs, src, dst, ed);
  QLG_HIP(hipGetLastError());
}

}  // namespace qlg
