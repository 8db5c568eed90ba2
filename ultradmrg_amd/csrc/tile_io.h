// Device helpers shared by the tiled MFMA kernels: raw buffer loads with
// hardware range checking.  A buffer resource spans 2^31 bytes from an
// operand's base address; an element that lies outside the tile's logical
// operand is given the byte offset kOob, which the buffer unit answers with
// zeros.  The loads therefore need neither a branch nor a select on the
// loaded value, so the compiler can keep them in flight across the MFMA
// work of the previous K slice.  Every in-range operand offset is checked on
// the host to stay below 2^31 bytes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace qlg {

constexpr uint32_t kOob = 0x80000000u;
constexpr int kRsrcWord3 = 0x00020000;  // raw (untyped) buffer, gfx9 family

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)kOob, kRsrcWord3);
}

// A pointer read from a pointer table has no known address space, so loads
// through it compile to generic (flat) loads, which count on the scalar-memory
// counter too: the wait for the next table entry then also waits for every
// element load issued before it.  The tables hold device allocations, so the
// pointer is global memory.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *global_ptr(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}

// complex element i (re, im) through such a pointer
__device__ inline double2 global_c128(const double *p, int64_t i) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = global_ptr(reinterpret_cast<const d2 *>(p))[i];
  return double2{v.x, v.y};
}

__device__ inline double load_f64(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}

__device__ inline double2 load_c128(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0);
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  const u2 lo = {v.x, v.y}, hi = {v.z, v.w};
  return double2{__builtin_bit_cast(double, lo), __builtin_bit_cast(double, hi)};
}

// complex float (re, im): two 4-byte loads (kOob + 4 lies outside the range too)
__device__ inline float2 load_c64(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  float2 v;
  v.x = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
  v.y = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_off + 4u, 0, 0));
  return v;
}

}  // namespace qlg
