// Device helpers shared by the per-step kernels (kernels.hip) and the batched
// replica kernel (replicas.hip): minimum image, binary32 decision arithmetic,
// cell indexing, wavefront/workgroup reductions, listed-term algebra.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace beads_amd {
namespace dev {

constexpr float kPiF = 3.14159265358979323846f;

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
// contiguous logical range [x * nb / 8, (x + 1) * nb / 8).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7u;
  const uint32_t xcd = b & 7u, idx = b >> 3;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// box_geometry_view.cuh:50-62 (minimum image of an already-formed difference).
__device__ __forceinline__ float mic(float d, float L, float hL) {
  return d > hL ? d - L : (d < -hL ? d + L : d);
}

// Minimum-image displacement xi - xj with exact intermediates: when the raw
// difference crosses half the box, xi - L (xi > L/2) or xj - L (xj > L/2) is
// exact by Sterbenz, so only the final small difference rounds. mic(xi - xj)
// instead rounds at the scale of L before folding back (about 1e-6 nm at
// L ~ 50 nm), which the r^-13 force amplifies past 1e-5.
__device__ __forceinline__ float mic_diff(float xi, float xj, float L, float hL) {
  const float d = xi - xj;
  return d > hL ? (xi - L) - xj : (d < -hL ? xi - (xj - L) : d);
}

// Squared distance exactly as written in the reference source (no FMA), so the
// r^2 <= r_s^2 and r^2 < r_c^2 decisions match a binary32 host evaluation.
__device__ __forceinline__ float r2_strict(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_xor(v, o, 64);
  }
  return v;
}

// Sum over a BLOCK-wide workgroup (BLOCK a multiple of 64, at most 1024);
// result valid in thread 0. lds needs BLOCK / 64 doubles.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    lds[wave] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    v = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) v += lds[w];
  }
  return v;
}

__device__ __forceinline__ double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Last-workgroup election after each workgroup has published its partials.
__device__ __forceinline__ bool arrive_last(uint32_t* counter) {
  __shared__ bool last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t prev = atomicAdd(counter, 1u);
    last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (last) {
    __threadfence();
  }
  return last;
}

// cell_list_build.cu:110-123: binary32 cell index, clamped.
__device__ __forceinline__ uint32_t axis_cell(float x, float lo, float inv, uint32_t count) {
  const float s = (x - lo) * inv;
  uint32_t i = s <= 0.f ? 0u : static_cast<uint32_t>(s);
  return i >= count ? count - 1 : i;
}

__device__ __forceinline__ uint32_t cell_of_position(float4 x, const BoxView& box,
                                                     const CellGridView& g) {
  const uint32_t ix = axis_cell(x.x, box.lo[0], g.inv_cell[0], g.nx);
  const uint32_t iy = axis_cell(x.y, box.lo[1], g.inv_cell[1], g.ny);
  const uint32_t iz = axis_cell(x.z, box.lo[2], g.inv_cell[2], g.nz);
  return ix + g.nx * (iy + g.ny * iz);
}

// regular_cell_neighbor_builder.cu:40-53.
__device__ __forceinline__ uint32_t wrapped_axis(uint32_t c, int d, uint32_t count) {
  if (d < 0) return c == 0 ? count - 1 : c - 1;
  if (d > 0) return c + 1 == count ? 0 : c + 1;
  return c;
}

// small_dimension_neighbor_builder.cu:56-76.
__device__ __forceinline__ int unique_axis(uint32_t c, uint32_t count, uint32_t out[3]) {
  if (count <= 1) {
    out[0] = 0;
    return 1;
  }
  if (count == 2) {
    out[0] = c == 0 ? 1u : 0u;
    out[1] = c;
    return 2;
  }
  out[0] = wrapped_axis(c, -1, count);
  out[1] = c;
  out[2] = wrapped_axis(c, 1, count);
  return 3;
}

// Candidate cells of the regular and small-dimension builders, z outermost.
struct CellNeighborhood {
  uint32_t xs[3], ys[3], zs[3];
  int nx, ny, nz;
};

__device__ __forceinline__ CellNeighborhood neighborhood(uint32_t c, const CellGridView& g) {
  CellNeighborhood h;
  const uint32_t cx = c % g.nx;
  const uint32_t cyz = c / g.nx;
  const uint32_t cy = cyz % g.ny;
  const uint32_t cz = cyz / g.ny;
  if (g.path == 1) {
    for (int d = -1; d <= 1; ++d) {
      h.xs[d + 1] = wrapped_axis(cx, d, g.nx);
      h.ys[d + 1] = wrapped_axis(cy, d, g.ny);
      h.zs[d + 1] = wrapped_axis(cz, d, g.nz);
    }
    h.nx = h.ny = h.nz = 3;
  } else {
    h.nx = unique_axis(cx, g.nx, h.xs);
    h.ny = unique_axis(cy, g.ny, h.ys);
    h.nz = unique_axis(cz, g.nz, h.zs);
  }
  return h;
}

// LJ pair value in FP64 from an exact binary32 displacement (FP64 VALU runs at
// the non-packed FP32 rate on CDNA4; binary32 would amplify the rounding of r^2
// six-fold through (sigma^2/r^2)^6). Returns the force scale s with
// f_i = s * d; energy gets 4 eps (sr12 - sr6) - shift.
__device__ __forceinline__ double lj_pair(float dx, float dy, float dz, float4 c, double& rr,
                                          double& energy) {
  const double ddx = dx, ddy = dy, ddz = dz;
  rr = ddx * ddx + ddy * ddy + ddz * ddz;
  double inv = __builtin_amdgcn_rcp(rr);
  inv = inv * (2.0 - rr * inv);
  inv = inv * (2.0 - rr * inv);
  const double sr2 = static_cast<double>(c.x) * inv;
  const double sr6 = sr2 * sr2 * sr2;
  const double sr12 = sr6 * sr6;
  energy = static_cast<double>(c.z) * (sr12 - sr6) - static_cast<double>(c.w);
  return static_cast<double>(c.y) * inv * (2.0 * sr12 - sr6);
}

}  // namespace dev
}  // namespace beads_amd
