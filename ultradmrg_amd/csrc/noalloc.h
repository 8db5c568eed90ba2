// Fixed device staging memory of the entry points that allocate nothing.
#pragma once
#include <cstdint>

namespace qlg {

// qlg_linear_combine_noalloc: at most this many basis vectors per call
constexpr int64_t kNoAllocMaxBases = 1024;
// kNoAllocMaxBases complex coefficients (2 doubles each), then kNoAllocMaxBases pointers
constexpr int64_t kNoAllocScratchDoubles = 3 * kNoAllocMaxBases;

// Look up the staging array's address on `device` (which must be current);
// once per device, from qlg_ctx_create.
void noalloc_init(int device);
// Its address on `device` (after noalloc_init).
double *noalloc_scratch(int device);

}  // namespace qlg
