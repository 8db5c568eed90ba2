// Batched Hermitian eigen-decomposition (rocSOLVER divide and conquer over
// host-threaded lanes); used by the block SVD.
#pragma once

#include "common.h"

namespace qlg {

void eigh_batched(Ctx &c, int dtype, int64_t ngroups, const qlg_eigh_group *groups);

}  // namespace qlg
