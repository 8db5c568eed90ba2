// Block-wise axis permutation (transpose.hip) and its plan (ctr.cpp).
#pragma once
#include <algorithm>

#include "common.h"

namespace qlg {

// one block: n elements read in row-major order of the source block (extents
// dim[0..rank)); source index (i_0, ..., i_{rank-1}) goes to dst + sum_k i_k dstride[k]
struct PermJob {
  int64_t src, dst, n;
  int64_t dim[QLG_MAX_RANK], dstride[QLG_MAX_RANK];
  int32_t rank, pad;
};

void launch_transpose(hipStream_t st, const PermJob *jobs, int64_t n, int64_t maxel, const double *src, double *dst,
                      int ed);

}  // namespace qlg
