// gfx950 kernels of the per-step MD path.
//
// Design (CDNA4): one work-item per atom, 256-wide workgroups (4 wave64s),
// float4 SoA streams, FP64 only for global reductions. Per-atom kernels that
// gather neighbour positions map workgroups onto atoms XCD by XCD so each XCD's
// L2 holds one contiguous (spatially sorted) slab of positions. Every kernel of
// the neighbour-rebuild sequence reads the device control block first and exits
// when no rebuild is due, so a step never waits on the host.
//
// Reference counterparts (paths under native/src):
//   control/bin/scan/scatter/sort  simulation/neighbor/cell_list_build.cu:110-219,347-414,
//                                  neighbor_system.cpp:79-146,198-216
//   build                          simulation/neighbor/full/*.cu,
//                                  full_neighbor_path_common.cuh:49-94,
//                                  rebuild_tracker.cu:26-53 (reference capture)
//   force                          forcefield/pair/pair_lj.cu:21-167,
//                                  bond/bond_harmonic.cu:24-99, angle/angle_harmonic.cu:47-178,
//                                  dihedral/dihedral_harmonic.cu:65-227,
//                                  dynamics/integrator/velocity_verlet.cu:97-111,
//                                  dynamics/thermostat/berendsen_thermostat.cu:104-163
//   integrate                      velocity_verlet.cu:59-95, rebuild_tracker.cu:55-94
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <cmath>

namespace beads_amd {
namespace {

constexpr float kPiF = 3.14159265358979323846f;

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
// contiguous logical range [x * nb / 8, (x + 1) * nb / 8).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7u;
  const uint32_t xcd = b & 7u, idx = b >> 3;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// box_geometry_view.cuh:50-62.
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

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_xor(v, o, 64);
  }
  return v;
}

// Sum over a 256-wide workgroup; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    lds[wave] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    v = lds[0] + lds[1] + lds[2] + lds[3];
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

// ---------------------------------------------------------------- rebuild --

// Rebuild bookkeeping of NeighborSystem::update_after_position_change /
// rebuild_now (neighbor_system.cpp:79-146, 198-207), on the device.
__global__ void control_kernel(Ctrl* ctrl, NeighborControl nc, int forced) {
  if (threadIdx.x != 0 || blockIdx.x != 0) {
    return;
  }
  if (forced) {
    ctrl->rebuild = 1;
  } else {
    const bool first = ctrl->checked == 0;
    ctrl->checked = 1;
    if (ctrl->rebuild && nc.check_every > 1 && first) {
      ctrl->dangerous += 1;
    }
  }
  if (ctrl->rebuild) {
    ctrl->build_count += 1;
    const bool reorder = nc.path != 0 && (ctrl->build_count % nc.sort_every) == 0;
    ctrl->reorder = reorder ? 1u : 0u;
    ctrl->reorder_count += reorder ? 1u : 0u;
    ctrl->checked = 0;
    ctrl->neighbor_total = 0;
  }
}

// cell_list_build.cu:110-166: binary32 cell index, clamped, x fastest.
__device__ __forceinline__ uint32_t axis_cell(float x, float lo, float inv, uint32_t count) {
  const float s = (x - lo) * inv;
  uint32_t i = s <= 0.f ? 0u : static_cast<uint32_t>(s);
  return i >= count ? count - 1 : i;
}

__global__ __launch_bounds__(kBlock) void bin_kernel(BinParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) {
    return;
  }
  const float4 x = p.pos[i];
  const uint32_t ix = axis_cell(x.x, p.box.lo[0], p.grid.inv_cell[0], p.grid.nx);
  const uint32_t iy = axis_cell(x.y, p.box.lo[1], p.grid.inv_cell[1], p.grid.ny);
  const uint32_t iz = axis_cell(x.z, p.box.lo[2], p.grid.inv_cell[2], p.grid.nz);
  const uint32_t c = ix + p.grid.nx * (iy + p.grid.ny * iz);
  p.cell_of[i] = c;
  atomicAdd(&p.cell_count[c], 1u);
}

// One workgroup: exclusive scan of cell counts into cell_start[0..n_cells],
// copy to the scatter cursor and clear the counts for the next rebuild.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void cell_scan_kernel(CellScanParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  __shared__ uint32_t sums[kScanBlock];
  const uint32_t t = threadIdx.x;
  const uint32_t chunk = (p.n_cells + kScanBlock - 1) / kScanBlock;
  const uint32_t begin = min(p.n_cells, t * chunk);
  const uint32_t end = min(p.n_cells, begin + chunk);
  uint32_t local = 0;
  for (uint32_t c = begin; c < end; ++c) {
    local += p.cell_count[c];
  }
  sums[t] = local;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 1024 chunk sums.
  for (uint32_t off = 1; off < kScanBlock; off <<= 1) {
    const uint32_t v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  uint32_t run = sums[t] - local;
  for (uint32_t c = begin; c < end; ++c) {
    const uint32_t cnt = p.cell_count[c];
    p.cell_start[c] = run;
    p.cell_cursor[c] = run;
    p.cell_count[c] = 0;
    run += cnt;
  }
  if (t == kScanBlock - 1) {
    p.cell_start[p.n_cells] = sums[t];
  }
}

__global__ __launch_bounds__(kBlock) void cell_scatter_kernel(CellScanParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) {
    return;
  }
  const uint32_t slot = atomicAdd(&p.cell_cursor[p.cell_of[i]], 1u);
  p.cell_atoms[slot] = i;
}

// Order each cell's members by slot: together with the counting sort this is a
// stable sort by (cell, slot), deterministic unlike the reference's atomic
// scatter (cell_list_build.cu:168-181).
__global__ __launch_bounds__(kBlock) void cell_sort_kernel(CellScanParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= p.n_cells) {
    return;
  }
  const uint32_t b = p.cell_start[c], e = p.cell_start[c + 1];
  uint32_t* a = p.cell_atoms;
  for (uint32_t k = b + 1; k < e; ++k) {
    const uint32_t v = a[k];
    uint32_t m = k;
    while (m > b && a[m - 1] > v) {
      a[m] = a[m - 1];
      --m;
    }
    a[m] = v;
  }
}

// cell_list_build.cu:183-219: gather every per-atom array into cell order
// (new slot s takes old slot cell_atoms[s]) and record the inverse map.
__global__ __launch_bounds__(kBlock) void reorder_gather_kernel(ReorderParams p) {
  if (p.ctrl->rebuild == 0 || p.ctrl->reorder == 0) {
    return;
  }
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= p.n) {
    return;
  }
  const uint32_t o = p.cell_atoms[s];
  p.new_of_old[o] = s;
  p.s_pos[s] = p.pos[o];
  p.s_vel[s] = p.vel[o];
  p.s_image[s] = p.image[o];
  p.s_tag[s] = p.tag[o];
  p.s_molid[s] = p.molid[o];
  p.s_cell_of[s] = p.cell_of[o];
  p.s_topo_count[s] = p.topo_count[o];
  const size_t n = p.n;
  for (uint32_t k = 0; k < p.bcap; ++k) p.s_bond[k * n + s] = p.bond[k * n + o];
  for (uint32_t k = 0; k < p.acap; ++k) p.s_angle[k * n + s] = p.angle[k * n + o];
  for (uint32_t k = 0; k < p.dcap; ++k) p.s_dih[k * n + s] = p.dih[k * n + o];
  for (uint32_t k = 0; k < p.ecap; ++k) p.s_excl[k * n + s] = p.excl[k * n + o];
}

// Copy back and translate partner slots through the inverse permutation.
__global__ __launch_bounds__(kBlock) void reorder_commit_kernel(ReorderParams p) {
  if (p.ctrl->rebuild == 0 || p.ctrl->reorder == 0) {
    return;
  }
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= p.n) {
    return;
  }
  p.pos[s] = p.s_pos[s];
  p.vel[s] = p.s_vel[s];
  p.image[s] = p.s_image[s];
  p.tag[s] = p.s_tag[s];
  p.molid[s] = p.s_molid[s];
  p.cell_of[s] = p.s_cell_of[s];
  p.cell_atoms[s] = s;
  const uint32_t cw = p.s_topo_count[s];
  p.topo_count[s] = cw;
  const size_t n = p.n;
  const uint32_t* m = p.new_of_old;
  const uint32_t nb = cw & 255u, na = (cw >> 8) & 255u, nd = (cw >> 16) & 255u, ne = cw >> 24;
  for (uint32_t k = 0; k < nb; ++k) {
    uint2 r = p.s_bond[k * n + s];
    r.x = m[r.x];
    p.bond[k * n + s] = r;
  }
  for (uint32_t k = 0; k < na; ++k) {
    uint4 r = p.s_angle[k * n + s];
    r.x = m[r.x];
    r.y = m[r.y];
    p.angle[k * n + s] = r;
  }
  for (uint32_t k = 0; k < nd; ++k) {
    uint4 r = p.s_dih[k * n + s];
    r.x = m[r.x];
    r.y = m[r.y];
    r.z = m[r.z];
    p.dih[k * n + s] = r;
  }
  for (uint32_t k = 0; k < ne; ++k) {
    p.excl[k * n + s] = m[p.s_excl[k * n + s]];
  }
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

// Full-list build with the append rule of full_neighbor_path_common.cuh:49-94:
// skip self, skip excluded, keep r2 <= r_s^2, overflow past max_neighbors.
// EXCL: 0 = no exclusions, 1 = up to 8 partners held in registers,
// 2 = any count read from memory. Also captures the rebuild reference.
template <int EXCL>
__global__ __launch_bounds__(kBlock) void build_kernel(BuildParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  __shared__ double red_lds[4];
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  uint32_t count = 0;
  bool overflowed = false;
  if (i < p.n) {
    const float4 xi = p.pos[i];
    const size_t n = p.n;
    uint32_t ex[8];
    uint32_t ne = 0;
    if (EXCL != 0) {
      ne = p.topo_count[i] >> 24;
      if (EXCL == 1) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          ex[k] = k < static_cast<int>(ne) ? p.excl[k * n + i] : 0xffffffffu;
        }
      }
    }
    const float L0 = p.box.L[0], L1 = p.box.L[1], L2 = p.box.L[2];
    const float h0 = p.box.hL[0], h1 = p.box.hL[1], h2 = p.box.hL[2];
    const uint32_t maxn = p.max_neighbors;
    auto visit = [&](uint32_t j) {
      const float4 xj = p.pos[j];
      const float dx = mic(xi.x - xj.x, L0, h0);
      const float dy = mic(xi.y - xj.y, L1, h1);
      const float dz = mic(xi.z - xj.z, L2, h2);
      const float r2 = r2_strict(dx, dy, dz);
      if (r2 > p.rs2 || j == i) {
        return;
      }
      if (EXCL == 1) {
        bool hit = false;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          hit |= ex[k] == j;
        }
        if (hit) return;
      } else if (EXCL == 2) {
        for (uint32_t k = 0; k < ne; ++k) {
          if (p.excl[k * n + i] == j) return;
        }
      }
      if (count < maxn) {
        p.nbr[static_cast<size_t>(count) * n + i] = j;
        ++count;
      } else {
        overflowed = true;
      }
    };
    if (p.grid.path == 0) {
      for (uint32_t j = 0; j < p.n; ++j) {
        visit(j);
      }
    } else {
      const uint32_t c = p.cell_of[i];
      const uint32_t cx = c % p.grid.nx;
      const uint32_t cyz = c / p.grid.nx;
      const uint32_t cy = cyz % p.grid.ny;
      const uint32_t cz = cyz / p.grid.ny;
      uint32_t xs[3], ys[3], zs[3];
      int nxs, nys, nzs;
      if (p.grid.path == 1) {
        for (int d = -1; d <= 1; ++d) {
          xs[d + 1] = wrapped_axis(cx, d, p.grid.nx);
          ys[d + 1] = wrapped_axis(cy, d, p.grid.ny);
          zs[d + 1] = wrapped_axis(cz, d, p.grid.nz);
        }
        nxs = nys = nzs = 3;
      } else {
        nxs = unique_axis(cx, p.grid.nx, xs);
        nys = unique_axis(cy, p.grid.ny, ys);
        nzs = unique_axis(cz, p.grid.nz, zs);
      }
      for (int a = 0; a < nzs; ++a) {
        for (int b = 0; b < nys; ++b) {
          const uint32_t row = p.grid.nx * (ys[b] + p.grid.ny * zs[a]);
          for (int d = 0; d < nxs; ++d) {
            const uint32_t nc = xs[d] + row;
            const uint32_t s0 = p.cell_start[nc], s1 = p.cell_start[nc + 1];
            for (uint32_t k = s0; k < s1; ++k) {
              visit(p.cell_atoms[k]);
            }
          }
        }
      }
    }
    p.nbr_count[i] = count;
    p.ref_pos[i] = xi;
    p.ref_image[i] = p.image[i];
  }
  const double total = block_sum(static_cast<double>(count), red_lds);
  if (threadIdx.x == 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&p.ctrl->neighbor_total),
              static_cast<unsigned long long>(total));
  }
  if (overflowed) {
    atomicOr(&p.ctrl->overflow, 1u);
  }
}

// ------------------------------------------------------------------ force --

struct LjCoeff {
  float sigma2, eps24, eps4, shift;
};

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 scale3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Minimum-image displacement p - q (wrapped coordinates), see mic_diff.
__device__ __forceinline__ float3 mic3(float3 p, float3 q, const BoxView& b) {
  return make_float3(mic_diff(p.x, q.x, b.L[0], b.hL[0]), mic_diff(p.y, q.y, b.L[1], b.hL[1]),
                     mic_diff(p.z, q.z, b.L[2], b.hL[2]));
}
__device__ __forceinline__ float clamp_cos(float c) { return c > 1.f ? 1.f : (c < -1.f ? -1.f : c); }

// Harmonic angle U = k/2 (theta - theta0)^2 (angle_harmonic.cu:61-158) for one
// role. Same mathematics as the reference, evaluated in a binary32-stable form:
// theta = atan2(|a x b|, a.b) and sin(theta) = |a x b| / (|a||b|) instead of
// acos(c) and sqrt(1 - c^2), and the force directions as a x (a x b), b x (a x b)
// instead of differences of nearly parallel vectors. The reference's floor
// (sin theta replaced by 0.001 at or below 0.001) is kept.
__device__ __forceinline__ float3 angle_role_force(float3 xi, float3 xj, float3 xk, float2 kc,
                                                   int role, const BoxView& box, float& energy,
                                                   float& virial) {
  const float3 a = mic3(xi, xj, box);
  const float3 b = mic3(xk, xj, box);
  const float ra2 = dot3(a, a), rb2 = dot3(b, b);
  if (ra2 == 0.f || rb2 == 0.f) {
    const float d = 0.f - kc.y;
    energy = 0.5f * kc.x * d * d;
    virial = 0.f;
    return make_float3(0.f, 0.f, 0.f);
  }
  const float ra = sqrtf(ra2), rb = sqrtf(rb2);
  const float3 nrm = cross3(a, b);
  const float nn = sqrtf(dot3(nrm, nrm));
  const float theta = atan2f(nn, dot3(a, b));
  const float dth = theta - kc.y;
  energy = 0.5f * kc.x * dth * dth;
  const float sn_true = nn / (ra * rb);
  const float sn = sn_true <= 0.001f ? 0.001f : sn_true;
  const float A = -kc.x * dth / sn;
  const float3 fi = scale3(cross3(a, nrm), A / (ra2 * ra * rb));
  const float3 fk = scale3(cross3(b, nrm), -A / (rb2 * rb * ra));
  virial = dot3(a, fi) + dot3(b, fk);
  if (role == 0) return fi;
  if (role == 2) return fk;
  return scale3(add3(fi, fk), -1.f);
}

// Harmonic dihedral k[1 + d cos(n phi)] (dihedral_harmonic.cu:83-206), one role.
// phi = atan2(|v1| v0.c1, c0.c1) carries the reference's sign convention
// (negative when v0.c1 < 0) without the ill-conditioned acos near 0 and pi.
__device__ __forceinline__ float3 dihedral_role_force(float3 xi, float3 xj, float3 xk, float3 xl,
                                                      float4 kc, int role, const BoxView& box,
                                                      float& energy, float& virial) {
  const float3 v0 = mic3(xi, xj, box);
  const float3 v1 = mic3(xk, xj, box);
  const float3 v2 = mic3(xk, xl, box);
  const float3 c0 = cross3(v0, v1), c1 = cross3(v1, v2);
  const float n1 = dot3(v1, v1), nc0 = dot3(c0, c0), nc1 = dot3(c1, c1);
  const int nn = static_cast<int>(kc.z);
  if (n1 == 0.f || nc0 == 0.f || nc1 == 0.f) {
    energy = nn == 0 ? kc.x * (1.f + kc.y) : kc.x;
    virial = 0.f;
    return make_float3(0.f, 0.f, 0.f);
  }
  const float phi = atan2f(sqrtf(n1) * dot3(v0, c1), dot3(c0, c1));
  const float arg = kc.z * phi;
  float sn, cs;
  sincosf(arg, &sn, &cs);
  energy = kc.x * (1.f + kc.y * cs);
  const float dE = -kc.x * kc.y * kc.z * sn;
  const float nb = sqrtf(n1);
  const float3 fi = scale3(c0, -dE * nb / nc0);
  const float3 fl = scale3(c1, dE * nb / nc1);
  const float inv_n1 = 1.f / n1;
  const float3 s = sub3(scale3(fi, dot3(v0, v1) * inv_n1), scale3(fl, dot3(v2, v1) * inv_n1));
  const float3 fj = sub3(s, fi);
  const float3 fk = scale3(add3(s, fl), -1.f);
  virial = dot3(v0, fi) + dot3(v1, fk) + dot3(sub3(v1, v2), fl);
  if (role == 0) return fi;
  if (role == 1) return fj;
  if (role == 2) return fk;
  return fl;
}

// Fused force evaluation: LJ over the stored full list (pair_lj.cu:21-167),
// listed terms gathered per atom (each atom evaluates every term it belongs to
// and keeps its own role's force: no atomics, deterministic), then optionally
// the post-force half-kick (velocity_verlet.cu:97-111) and the kinetic-energy
// reduction for the Berendsen factor (berendsen_thermostat.cu:109-163), with
// the factor computed in FP64 by the last workgroup.
// TYPES: 0 = single type (coefficients in registers), 1 = table in LDS,
// 2 = table in global memory.
template <bool KICK, bool OBS, int TYPES>
__global__ __launch_bounds__(kBlock) void force_kernel(ForceParams p) {
  __shared__ LjCoeff lj_lds[TYPES == 1 ? kMaxLdsTypes * kMaxLdsTypes : 1];
  __shared__ double red_lds[4];
  const int T = p.n_types;
  if (TYPES == 1) {
    for (int k = threadIdx.x; k < T * T; k += kBlock) {
      const float4 c = p.lj[k];
      lj_lds[k] = LjCoeff{c.x, c.y, c.z, c.w};
    }
    __syncthreads();
  }
  LjCoeff c0{};
  if (TYPES == 0) {
    const float4 c = p.lj[0];
    c0 = LjCoeff{c.x, c.y, c.z, c.w};
  }
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const size_t n = p.n;
  double ke = 0.0;
  double e_pair = 0.0, w_pair = 0.0;
  double e_bond = 0.0, w_bond = 0.0, e_angle = 0.0, w_angle = 0.0, e_dih = 0.0, w_dih = 0.0;
  if (i < p.n) {
    const float4 xi4 = p.pos[i];
    const int ti = __float_as_int(xi4.w);
    const float L0 = p.box.L[0], L1 = p.box.L[1], L2 = p.box.L[2];
    const float h0 = p.box.hL[0], h1 = p.box.hL[1], h2 = p.box.hL[2];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    double fxd = 0.0, fyd = 0.0, fzd = 0.0;
    const uint32_t cnt = p.nbr_count[i];
    const uint32_t* nb = p.nbr + i;
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t j = nb[static_cast<size_t>(k) * n];
      const float4 xj = p.pos[j];
      const float dx = mic_diff(xi4.x, xj.x, L0, h0);
      const float dy = mic_diff(xi4.y, xj.y, L1, h1);
      const float dz = mic_diff(xi4.z, xj.z, L2, h2);
      const float r2 = r2_strict(dx, dy, dz);
      if (r2 < p.rc2) {
        LjCoeff c;
        if (TYPES == 0) {
          c = c0;
        } else if (TYPES == 1) {
          c = lj_lds[ti * T + __float_as_int(xj.w)];
        } else {
          const float4 g = p.lj[ti * T + __float_as_int(xj.w)];
          c = LjCoeff{g.x, g.y, g.z, g.w};
        }
        // The pair value in FP64 (full-rate VALU on CDNA4): dx is exact in
        // binary32, so the only rounding left is FP64's; binary32 would
        // amplify the rounding of r^2 six-fold through (sigma^2/r^2)^6.
        const double ddx = dx, ddy = dy, ddz = dz;
        const double rr = ddx * ddx + ddy * ddy + ddz * ddz;
        double inv = __builtin_amdgcn_rcp(rr);
        inv = inv * (2.0 - rr * inv);
        inv = inv * (2.0 - rr * inv);
        const double sr2 = static_cast<double>(c.sigma2) * inv;
        const double sr6 = sr2 * sr2 * sr2;
        const double sr12 = sr6 * sr6;
        const double s = static_cast<double>(c.eps24) * inv * (2.0 * sr12 - sr6);
        fxd += s * ddx;
        fyd += s * ddy;
        fzd += s * ddz;
        if (OBS) {
          e_pair += static_cast<double>(c.eps4) * (sr12 - sr6) - static_cast<double>(c.shift);
          w_pair += s * rr;
        }
      }
    }
    fx = static_cast<float>(fxd);
    fy = static_cast<float>(fyd);
    fz = static_cast<float>(fzd);
    if (p.topo.has_listed) {
      const TopologyView& tp = p.topo;
      const uint32_t cw = tp.count[i];
      const uint32_t nbond = cw & 255u, nang = (cw >> 8) & 255u, ndih = (cw >> 16) & 255u;
      const float3 xi = f3(xi4);
      for (uint32_t k = 0; k < nbond; ++k) {
        const uint2 r = tp.bond[k * n + i];
        const float2 kc = tp.bond_coeff[r.y & 0x7fffffffu];
        const float3 d = mic3(xi, f3(p.pos[r.x]), p.box);
        const float r2 = dot3(d, d);
        const float rr = sqrtf(r2);
        const float stretch = rr - kc.y;
        const bool first = (r.y >> 31) != 0;
        if (OBS && first) e_bond += 0.5f * kc.x * stretch * stretch;
        if (rr != 0.f) {
          const float s = -kc.x * stretch / rr;
          fx += s * d.x;
          fy += s * d.y;
          fz += s * d.z;
          if (OBS && first) w_bond += s * r2;
        }
      }
      for (uint32_t k = 0; k < nang; ++k) {
        const uint4 r = tp.angle[k * n + i];
        const int role = static_cast<int>(r.w);
        const float3 a = f3(p.pos[r.x]);
        const float3 b = f3(p.pos[r.y]);
        const float3 pi = role == 0 ? xi : a;
        const float3 pj = role == 0 ? a : (role == 1 ? xi : b);
        const float3 pk = role == 2 ? xi : b;
        float e, w;
        const float3 f = angle_role_force(pi, pj, pk, tp.angle_coeff[r.z], role, p.box, e, w);
        fx += f.x;
        fy += f.y;
        fz += f.z;
        if (OBS && role == 0) {
          e_angle += e;
          w_angle += w;
        }
      }
      for (uint32_t k = 0; k < ndih; ++k) {
        const uint4 r = tp.dih[k * n + i];
        const int role = static_cast<int>(r.w >> 30);
        const float3 a = f3(p.pos[r.x]);
        const float3 b = f3(p.pos[r.y]);
        const float3 c = f3(p.pos[r.z]);
        float3 q0, q1, q2, q3;
        if (role == 0) { q0 = xi; q1 = a; q2 = b; q3 = c; }
        else if (role == 1) { q0 = a; q1 = xi; q2 = b; q3 = c; }
        else if (role == 2) { q0 = a; q1 = b; q2 = xi; q3 = c; }
        else { q0 = a; q1 = b; q2 = c; q3 = xi; }
        float e, w;
        const float3 f = dihedral_role_force(q0, q1, q2, q3, tp.dih_coeff[r.w & 0x3fffffffu],
                                             role, p.box, e, w);
        fx += f.x;
        fy += f.y;
        fz += f.z;
        if (OBS && role == 0) {
          e_dih += e;
          w_dih += w;
        }
      }
    }
    p.force[i] = make_float4(fx, fy, fz, 0.f);
    float4 v = p.vel[i];
    if (KICK) {
      const float h = 0.5f * p.dt / v.w;
      v.x += h * fx;
      v.y += h * fy;
      v.z += h * fz;
      p.vel[i] = v;
    }
    const double vx = v.x, vy = v.y, vz = v.z;
    ke = 0.5 * static_cast<double>(v.w) * (vx * vx + vy * vy + vz * vz);
  }
  const uint32_t nbk = gridDim.x;
  double* part = p.partials;
  double s = block_sum(ke, red_lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (OBS) {
    const double vals[8] = {0.5 * e_pair, e_bond, e_angle, e_dih, 0.5 * w_pair, w_bond, w_angle, w_dih};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      s = block_sum(vals[q], red_lds);
      if (threadIdx.x == 0) part[(q + 1) * static_cast<size_t>(nbk) + blockIdx.x] = s;
    }
  }
  if (!arrive_last(&p.ctrl->arrive_force)) {
    return;
  }
  const int nq = OBS ? 9 : 1;
  for (int q = 0; q < nq; ++q) {
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nbk; b += kBlock) {
      acc += load_coherent(part + q * static_cast<size_t>(nbk) + b);
    }
    acc = block_sum(acc, red_lds);
    if (threadIdx.x == 0) {
      Ctrl* ctrl = p.ctrl;
      if (q == 0) {
        ctrl->ke_pre = acc;
        double lambda = 1.0;
        if (KICK && p.berendsen && acc > 0.0 && p.dof > 0.0) {
          const double current = 2.0 * acc / p.dof;
          const double arg = 1.0 + p.dt_over_tau * (p.target_thermal / current - 1.0);
          lambda = sqrt(arg > 0.0 ? arg : 0.0);
        }
        ctrl->lambda = lambda;
        if (!KICK) ctrl->ke = acc;
      } else if (q <= 4) {
        ctrl->pe[q - 1] = acc;
      } else {
        ctrl->virial[q - 5] = acc;
      }
    }
  }
  if (threadIdx.x == 0) {
    p.ctrl->rebuild = 0;
    p.ctrl->reorder = 0;
    p.ctrl->arrive_force = 0;
  }
}

// Apply the pending Berendsen factor and reduce the kinetic energy of the
// stored velocities (output steps; berendsen_thermostat.cu:135-163 and
// thermo_output.cu:60-85).
template <bool THERMO>
__global__ __launch_bounds__(kBlock) void scale_kernel(ScaleParams p) {
  __shared__ double red_lds[4];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  double ke = 0.0;
  if (i < p.n) {
    float4 v = p.vel[i];
    if (THERMO) {
      const float lam = static_cast<float>(p.ctrl->lambda);
      v.x *= lam;
      v.y *= lam;
      v.z *= lam;
      p.vel[i] = v;
    }
    const double vx = v.x, vy = v.y, vz = v.z;
    ke = 0.5 * static_cast<double>(v.w) * (vx * vx + vy * vy + vz * vz);
  }
  const double s = block_sum(ke, red_lds);
  if (threadIdx.x == 0) p.partials[blockIdx.x] = s;
  if (!arrive_last(&p.ctrl->arrive_scale)) {
    return;
  }
  double acc = 0.0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock) {
    acc += load_coherent(p.partials + b);
  }
  acc = block_sum(acc, red_lds);
  if (threadIdx.x == 0) {
    p.ctrl->ke = acc;
    p.ctrl->lambda = 1.0;
    p.ctrl->arrive_scale = 0;
  }
}

// Pending Berendsen factor, first half-kick, drift and wrap
// (velocity_verlet.cu:59-95; box_geometry_view.cuh:92-102), then the
// displacement test against the last build (rebuild_tracker.cu:55-94).
template <bool THERMO, bool CHECK>
__global__ __launch_bounds__(kBlock) void integrate_kernel(IntegrateParams p) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) {
    return;
  }
  float4 v = p.vel[i];
  const float4 f = p.force[i];
  float4 x = p.pos[i];
  if (THERMO) {
    const float lam = static_cast<float>(p.ctrl->lambda);
    v.x *= lam;
    v.y *= lam;
    v.z *= lam;
  }
  const float h = 0.5f * p.dt / v.w;
  v.x += h * f.x;
  v.y += h * f.y;
  v.z += h * f.z;
  x.x += p.dt * v.x;
  x.y += p.dt * v.y;
  x.z += p.dt * v.z;
  float c[3];
  {
#pragma clang fp contract(off)
    c[0] = floorf((x.x - p.box.lo[0]) * p.box.invL[0]);
    c[1] = floorf((x.y - p.box.lo[1]) * p.box.invL[1]);
    c[2] = floorf((x.z - p.box.lo[2]) * p.box.invL[2]);
    x.x -= c[0] * p.box.L[0];
    x.y -= c[1] * p.box.L[1];
    x.z -= c[2] * p.box.L[2];
  }
  p.pos[i] = x;
  p.vel[i] = v;
  const bool wrapped = c[0] != 0.f || c[1] != 0.f || c[2] != 0.f;
  int4 im = make_int4(0, 0, 0, 0);
  if (wrapped || CHECK) {
    im = p.image[i];
  }
  if (wrapped) {
    im.x += static_cast<int>(c[0]);
    im.y += static_cast<int>(c[1]);
    im.z += static_cast<int>(c[2]);
    p.image[i] = im;
  }
  if (CHECK) {
#pragma clang fp contract(off)
    const float4 r = p.ref_pos[i];
    const int4 ri = p.ref_image[i];
    const float dx = (x.x - r.x) + static_cast<float>(im.x - ri.x) * p.box.L[0];
    const float dy = (x.y - r.y) + static_cast<float>(im.y - ri.y) * p.box.L[1];
    const float dz = (x.z - r.z) + static_cast<float>(im.z - ri.z) * p.box.L[2];
    if (dx * dx + dy * dy + dz * dz > p.threshold_sq) {
      __hip_atomic_store(&p.ctrl->rebuild, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace

void launch_control(Ctrl* ctrl, NeighborControl control, bool forced, hipStream_t stream) {
  control_kernel<<<1, 64, 0, stream>>>(ctrl, control, forced ? 1 : 0);
}

void launch_bin(const BinParams& p, hipStream_t stream) {
  bin_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
}

void launch_cell_scan(const CellScanParams& p, hipStream_t stream) {
  cell_scan_kernel<<<1, kScanBlock, 0, stream>>>(p);
}

void launch_cell_scatter(const CellScanParams& p, hipStream_t stream) {
  cell_scatter_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
}

void launch_cell_sort(const CellScanParams& p, hipStream_t stream) {
  cell_sort_kernel<<<grid_for(p.n_cells), kBlock, 0, stream>>>(p);
}

void launch_reorder(const ReorderParams& p, hipStream_t stream) {
  reorder_gather_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
  reorder_commit_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
}

void launch_build(const BuildParams& p, hipStream_t stream) {
  const unsigned g = grid_for(p.n);
  if (p.ecap == 0) {
    build_kernel<0><<<g, kBlock, 0, stream>>>(p);
  } else if (p.ecap <= 8) {
    build_kernel<1><<<g, kBlock, 0, stream>>>(p);
  } else {
    build_kernel<2><<<g, kBlock, 0, stream>>>(p);
  }
}

void launch_force(const ForceParams& p, bool kick, bool obs, hipStream_t stream) {
  const unsigned g = grid_for(p.n);
  const int types = p.n_types == 1 ? 0 : (p.n_types <= kMaxLdsTypes ? 1 : 2);
#define BEADS_FORCE_CASE(K, O, T)                                       \
  if (kick == K && obs == O && types == T) {                            \
    force_kernel<K, O, T><<<g, kBlock, 0, stream>>>(p);                 \
    return;                                                             \
  }
  BEADS_FORCE_CASE(true, false, 0)
  BEADS_FORCE_CASE(true, false, 1)
  BEADS_FORCE_CASE(true, false, 2)
  BEADS_FORCE_CASE(true, true, 0)
  BEADS_FORCE_CASE(true, true, 1)
  BEADS_FORCE_CASE(true, true, 2)
  BEADS_FORCE_CASE(false, false, 0)
  BEADS_FORCE_CASE(false, false, 1)
  BEADS_FORCE_CASE(false, false, 2)
  BEADS_FORCE_CASE(false, true, 0)
  BEADS_FORCE_CASE(false, true, 1)
  BEADS_FORCE_CASE(false, true, 2)
#undef BEADS_FORCE_CASE
}

void launch_integrate(const IntegrateParams& p, bool thermostat, bool check,
                      bool /*force_rebuild*/, hipStream_t stream) {
  const unsigned g = grid_for(p.n);
  if (thermostat) {
    if (check) integrate_kernel<true, true><<<g, kBlock, 0, stream>>>(p);
    else integrate_kernel<true, false><<<g, kBlock, 0, stream>>>(p);
  } else {
    if (check) integrate_kernel<false, true><<<g, kBlock, 0, stream>>>(p);
    else integrate_kernel<false, false><<<g, kBlock, 0, stream>>>(p);
  }
}

void launch_scale(const ScaleParams& p, bool thermostat, hipStream_t stream) {
  if (thermostat) scale_kernel<true><<<grid_for(p.n), kBlock, 0, stream>>>(p);
  else scale_kernel<false><<<grid_for(p.n), kBlock, 0, stream>>>(p);
}

}  // namespace beads_amd
