// Staging memory of the entry points that allocate nothing.
//
// qlg_linear_combine stages its coefficient table and its table of basis
// pointers in two device buffers it allocates per call.
// qlg_linear_combine_noalloc stages them in one fixed array in the library's
// device image instead, `scratch` below: kNoAllocMaxBases complex
// coefficients followed by kNoAllocMaxBases pointers (24 KiB per GPU, held
// until the process ends).  Host code never passes `scratch` itself to a
// kernel or a copy: the array's device address is looked up once per device
// with hipGetSymbolAddress, when a context is created on it, and that address
// is what every later call uses.  All calls on one GPU share its array; two
// calls that overlap in time on one GPU (from different threads or streams)
// race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "common.h"
#include "noalloc.h"

namespace qlg {

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::unordered_map<int, double *> g_scratch_addr;   // device -> address of `scratch` on it
}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_addr.count(device)) return;
  int cur = -1;
  QLG_HIP(hipGetDevice(&cur));
  QLG_CHECK(cur == device, "noalloc_init: the device must be current");
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw Error(QLG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_addr[device] = static_cast<double *>(ptr);
}

double *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  QLG_CHECK(it != g_scratch_addr.end(), "no staging memory on this device (create a context on it first)");
  return it->second;
}

}  // namespace qlg
