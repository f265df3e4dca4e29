// Storage precision of the engine, chosen at build time (types.hpp:12-16):
// binary32 by default, binary64 when BEADS_REAL64 is defined. Positions,
// velocities, forces, the box, the neighbour decisions and the integrator run
// in real_t; pair values, listed terms and every reduction are FP64 in both
// builds. The two builds are separate libraries (libbeads_amd.so,
// libbeads_amd_real64.so) with the same C ABI over beads_real_t.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace beads_amd {

#ifdef BEADS_REAL64
using real_t = double;
using real2 = double2;
using real3 = double3;
using real4 = double4;
#else
using real_t = float;
using real2 = float2;
using real3 = float3;
using real4 = float4;
#endif

constexpr bool kReal64 = sizeof(real_t) == sizeof(double);
constexpr const char* kRealPrecision = kReal64 ? "double" : "single";
constexpr const char* kRealDtype = kReal64 ? "float64" : "float32";

__host__ __device__ __forceinline__ real2 make_real2(real_t x, real_t y) {
  real2 v;
  v.x = x;
  v.y = y;
  return v;
}
__host__ __device__ __forceinline__ real3 make_real3(real_t x, real_t y, real_t z) {
  real3 v;
  v.x = x;
  v.y = y;
  v.z = z;
  return v;
}
__host__ __device__ __forceinline__ real4 make_real4(real_t x, real_t y, real_t z, real_t w) {
  real4 v;
  v.x = x;
  v.y = y;
  v.z = z;
  v.w = w;
  return v;
}

// A 32-bit id (type or slot) carried in the bits of a record's .w component:
// the whole word in binary32, the low word in binary64. Only ever moved and
// compared as bits, never used in arithmetic.
__host__ __device__ __forceinline__ uint32_t id_of(real_t w) {
#ifdef BEADS_REAL64
  return static_cast<uint32_t>(__builtin_bit_cast(uint64_t, w));
#else
  return __builtin_bit_cast(uint32_t, w);
#endif
}
__host__ __device__ __forceinline__ real_t real_of_id(uint32_t id) {
#ifdef BEADS_REAL64
  return __builtin_bit_cast(double, static_cast<uint64_t>(id));
#else
  return __builtin_bit_cast(float, id);
#endif
}

}  // namespace beads_amd
