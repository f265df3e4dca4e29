// Device helpers shared by the per-step kernels (kernels.hip) and the batched
// replica kernel (replicas.hip): minimum image, decision arithmetic in the
// storage precision real_t, cell indexing, wavefront/workgroup reductions,
// listed-term algebra.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace beads_amd {
namespace dev {

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
// contiguous logical range [x * nb / 8, (x + 1) * nb / 8).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7u;
  const uint32_t xcd = b & 7u, idx = b >> 3;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// box_geometry_view.cuh:50-62 (minimum image of an already-formed difference).
__device__ __forceinline__ real_t mic(real_t d, real_t L, real_t hL) {
  return d > hL ? d - L : (d < -hL ? d + L : d);
}

// Minimum-image displacement xi - xj with exact intermediates: when the raw
// difference crosses half the box, xi - L (xi > L/2) or xj - L (xj > L/2) is
// exact by Sterbenz, so only the final small difference rounds. mic(xi - xj)
// instead rounds at the scale of L before folding back (about 1e-6 nm at
// L ~ 50 nm), which the r^-13 force amplifies past 1e-5.
__device__ __forceinline__ real_t mic_diff(real_t xi, real_t xj, real_t L, real_t hL) {
  const real_t d = xi - xj;
  return d > hL ? (xi - L) - xj : (d < -hL ? xi - (xj - L) : d);
}

// mic_diff with both folded forms evaluated and one selected (no branches):
// xi_m = xi - L is the caller's, hoisted out of a loop over partners.
__device__ __forceinline__ real_t mic_diff_sel(real_t xi, real_t xi_m, real_t xj, real_t L,
                                              real_t hL) {
  const real_t d = xi - xj;
  const real_t up = xi_m - xj;
  const real_t dn = xi - (xj - L);
  real_t r = __builtin_unpredictable(d > hL) ? up : d;
  r = __builtin_unpredictable(d < -hL) ? dn : r;
  return r;
}

// Squared distance exactly as written in the reference source (no FMA), so the
// r^2 <= r_s^2 and r^2 < r_c^2 decisions match a host evaluation in real_t.
__device__ __forceinline__ real_t r2_strict(real_t dx, real_t dy, real_t dz) {
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

// Sum of a value over an aligned team of TEAM adjacent lanes (xor butterfly:
// every lane ends with the same bits, since a + b == b + a exactly).
template <int TEAM>
__device__ __forceinline__ double team_sum(double v) {
#pragma unroll
  for (int m = TEAM / 2; m > 0; m >>= 1) {
    v += __shfl_xor(v, m, TEAM);
  }
  return v;
}

// counter[key] += 1 for every active lane of the wave with one atomic per
// distinct key: the lanes sharing a key add their count through the first of
// them. Returns the counter's value before this lane's increment (the lanes of
// a key take consecutive values in lane order).
__device__ __forceinline__ uint32_t wave_increment(uint32_t* counter, uint32_t key) {
  const uint32_t lane = threadIdx.x & 63u;
  uint64_t pending = __ballot(1);
  uint32_t result = 0;
  while (pending != 0) {
    const int first = __builtin_ctzll(pending);
    const uint32_t lead = static_cast<uint32_t>(__shfl(static_cast<int>(key), first));
    const uint64_t group = __ballot(key == lead) & pending;
    if ((group >> lane) & 1ull) {
      const uint32_t rank = static_cast<uint32_t>(__builtin_popcountll(group & ((1ull << lane) - 1ull)));
      uint32_t base = 0;
      if (rank == 0) {
        base = atomicAdd(&counter[lead], static_cast<uint32_t>(__builtin_popcountll(group)));
      }
      result = static_cast<uint32_t>(__shfl(static_cast<int>(base), first)) + rank;
    }
    pending &= ~group;
  }
  return result;
}

// 1/x in FP64: the hardware reciprocal (relative error <= 4.6e-8 measured on
// gfx950) refined by one Newton step (<= 2.3e-15).
__device__ __forceinline__ double rcp_f64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return r * (2.0 - x * r);
}

// 1/sqrt(x) in FP64 for x > 0: the hardware reciprocal square root refined by
// one Newton step (the hardware value's relative error squared, about 1e-15).
__device__ __forceinline__ double rsq_f64(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  return r * (1.5 - 0.5 * x * r * r);
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

// cell_list_build.cu:110-123: cell index in real_t, clamped.
__device__ __forceinline__ uint32_t axis_cell(real_t x, real_t lo, real_t inv, uint32_t count) {
  const real_t s = (x - lo) * inv;
  uint32_t i = s <= real_t(0) ? 0u : static_cast<uint32_t>(s);
  return i >= count ? count - 1 : i;
}

__device__ __forceinline__ uint32_t cell_of_position(real4 x, const BoxView& box,
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

// LJ pair value in FP64 from the displacement in real_t (exact in binary32
// too: FP64 VALU runs at the non-packed FP32 rate on CDNA4, and binary32 would
// amplify the rounding of r^2 six-fold through (sigma^2/r^2)^6). Returns the
// force scale s with f_i = s * d; energy gets 4 eps (sr12 - sr6) - shift.
__device__ __forceinline__ double lj_pair(real_t dx, real_t dy, real_t dz, const LjCoeff& c,
                                          double& rr, double& energy) {
  const double ddx = dx, ddy = dy, ddz = dz;
  rr = ddx * ddx + ddy * ddy + ddz * ddz;
  const double inv = rcp_f64(rr);
  const double sr2 = c.sigma2 * inv;
  const double sr6 = sr2 * sr2 * sr2;
  const double sr12 = sr6 * sr6;
  energy = c.eps4 * (sr12 - sr6) - c.shift;
  return c.eps24 * inv * (2.0 * sr12 - sr6);
}

// ------------------------------------------------------- listed terms ------

__device__ __forceinline__ real3 f3(real4 v) { return make_real3(v.x, v.y, v.z); }

// Minimum-image displacement p - q (wrapped coordinates), see mic_diff.
__device__ __forceinline__ real3 mic3(real3 p, real3 q, const BoxView& b) {
  return make_real3(mic_diff(p.x, q.x, b.L[0], b.hL[0]), mic_diff(p.y, q.y, b.L[1], b.hL[1]),
                    mic_diff(p.z, q.z, b.L[2], b.hL[2]));
}

// FP64 vectors for the listed terms. Their displacements come exact (or
// rounded once) from the positions through mic3; everything after that is
// FP64, because near-collinear triples make the angle and dihedral
// algebra ill-conditioned in binary32 (|v0 x v1| ~ 1e-5 |v0||v1| occurs in a
// million-bead melt and costs binary32 three significant digits there).
struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 dbl(real3 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ real3 flt(d3 v) {
  return make_real3(static_cast<real_t>(v.x), static_cast<real_t>(v.y), static_cast<real_t>(v.z));
}
__device__ __forceinline__ d3 dsub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 dadd(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 dscale(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// The listed terms take square roots and quotients as products with the
// FP64 reciprocal (square root) of rcp_f64 / rsq_f64 (one Newton step each,
// about 1e-15 relative) instead of IEEE division and square root, whose
// correctly rounded sequences cost several times the instructions.

// Harmonic bond U = k/2 (r - r0)^2 (bond_harmonic.cu:41-79) on the displacement
// d = x_self - x_partner; returns the force on self and the term's energy and
// virial d.f.
__device__ __forceinline__ real3 bond_force(real3 df, real2 kc, double& energy,
                                            double& virial) {
  const d3 d = dbl(df);
  const double k = kc.x, r0 = kc.y;
  const double r2 = ddot(d, d);
  if (r2 == 0.0) {
    energy = 0.5 * k * r0 * r0;
    virial = 0.0;
    return make_real3(0, 0, 0);
  }
  const double inv_r = rsq_f64(r2);
  const double stretch = r2 * inv_r - r0;
  energy = 0.5 * k * stretch * stretch;
  const double s = -k * stretch * inv_r;
  virial = s * r2;
  return flt(dscale(d, s));
}

// Harmonic angle U = k/2 (theta - theta0)^2 (angle_harmonic.cu:61-158) for one
// role. Same mathematics as the reference, in a stable form: theta =
// atan2(|a x b|, a.b) and sin(theta) = |a x b| / (|a||b|) instead of acos(c)
// and sqrt(1 - c^2), and the force directions as a x (a x b), b x (a x b)
// instead of differences of nearly parallel vectors. The reference's floor
// (sin theta replaced by kSinFloor at or below it, angle_harmonic.cu:25, 126;
// real_t{0.001}: the binary32 value in the binary32 build) is kept.
__device__ __forceinline__ real3 angle_role_force(real3 xi, real3 xj, real3 xk, real2 kc,
                                                  int role, const BoxView& box, double& energy,
                                                  double& virial) {
  const d3 a = dbl(mic3(xi, xj, box));
  const d3 b = dbl(mic3(xk, xj, box));
  const double k = kc.x, theta0 = kc.y;
  const double ra2 = ddot(a, a), rb2 = ddot(b, b);
  if (ra2 == 0.0 || rb2 == 0.0) {
    energy = 0.5 * k * theta0 * theta0;
    virial = 0.0;
    return make_real3(0, 0, 0);
  }
  const double ia = rsq_f64(ra2), ib = rsq_f64(rb2);
  const d3 nrm = dcross(a, b);
  const double nn2 = ddot(nrm, nrm);
  const double nn = nn2 > 0.0 ? nn2 * rsq_f64(nn2) : 0.0;
  // The FP64 arguments are well conditioned; in the binary32 build an atan2
  // of them in binary32 is accurate to a few 1e-8 relative (below the
  // force's own rounding) and keeps the pass's register budget.
  const double theta = kReal64 ? atan2(nn, ddot(a, b))
                               : atan2f(static_cast<float>(nn), static_cast<float>(ddot(a, b)));
  const double dth = theta - theta0;
  energy = 0.5 * k * dth * dth;
  const double sn_true = nn * ia * ib;
  constexpr double kSinFloor = static_cast<double>(static_cast<real_t>(0.001));
  const double sn = sn_true <= kSinFloor ? kSinFloor : sn_true;
  const double A = -k * dth * rcp_f64(sn);
  const double aib = A * ia * ib;
  const d3 fi = dscale(dcross(a, nrm), aib * ia * ia);
  const d3 fk = dscale(dcross(b, nrm), -aib * ib * ib);
  virial = ddot(a, fi) + ddot(b, fk);
  if (role == 0) return flt(fi);
  if (role == 2) return flt(fk);
  return flt(dscale(dadd(fi, fk), -1.0));
}

// Harmonic dihedral k[1 + d cos(n phi)] (dihedral_harmonic.cu:83-206), one role.
// phi = atan2(|v1| v0.c1, c0.c1) carries the reference's sign convention
// (negative when v0.c1 < 0) without the ill-conditioned acos near 0 and pi.
__device__ __forceinline__ real3 dihedral_role_force(real3 xi, real3 xj, real3 xk, real3 xl,
                                                     real4 kc, int role, const BoxView& box,
                                                     double& energy, double& virial) {
  const d3 v0 = dbl(mic3(xi, xj, box));
  const d3 v1 = dbl(mic3(xk, xj, box));
  const d3 v2 = dbl(mic3(xk, xl, box));
  const d3 c0 = dcross(v0, v1), c1 = dcross(v1, v2);
  const double n1 = ddot(v1, v1), nc0 = ddot(c0, c0), nc1 = ddot(c1, c1);
  const double k = kc.x, dsign = kc.y, nmul = kc.z;
  if (n1 == 0.0 || nc0 == 0.0 || nc1 == 0.0) {
    energy = nmul == 0.0 ? k * (1.0 + dsign) : k;
    virial = 0.0;
    return make_real3(0, 0, 0);
  }
  const double inb = rsq_f64(n1);
  const double nb = n1 * inb;
  // cos(phi) and sin(phi) from the cross products (sin carries the sign of
  // v0.c1), then cos(n phi), sin(n phi) by the angle-addition recurrence:
  // the reference's k[1 + d cos(n phi)] and its derivative without a
  // transcendental call.
  const double ic0 = rsq_f64(nc0), ic1 = rsq_f64(nc1);
  const double inv_c = ic0 * ic1;
  const double c1phi = ddot(c0, c1) * inv_c;
  const double s1phi = nb * ddot(v0, c1) * inv_c;
  double cs = 1.0, sn = 0.0;
  for (int m = 0; m < static_cast<int>(nmul); ++m) {
    const double c2 = cs * c1phi - sn * s1phi;
    sn = sn * c1phi + cs * s1phi;
    cs = c2;
  }
  energy = k * (1.0 + dsign * cs);
  const double dE = -k * dsign * nmul * sn;
  const d3 fi = dscale(c0, -dE * nb * ic0 * ic0);
  const d3 fl = dscale(c1, dE * nb * ic1 * ic1);
  const double inv_n1 = inb * inb;
  const d3 sv = dsub(dscale(fi, ddot(v0, v1) * inv_n1), dscale(fl, ddot(v2, v1) * inv_n1));
  const d3 fj = dsub(sv, fi);
  const d3 fk = dscale(dadd(sv, fl), -1.0);
  virial = ddot(v0, fi) + ddot(v1, fk) + ddot(dsub(v1, v2), fl);
  if (role == 0) return flt(fi);
  if (role == 1) return flt(fj);
  if (role == 2) return flt(fk);
  return flt(fl);
}

// Per-thread sums of one force pass.
struct StepSums {
  double ke = 0.0;
  double e_pair = 0.0, w_pair = 0.0;
  double e_bond = 0.0, w_bond = 0.0, e_angle = 0.0, w_angle = 0.0, e_dih = 0.0, w_dih = 0.0;
};

// Listed terms, atom-centric: atom i evaluates every term it belongs to and
// keeps its own role's force (no atomics, deterministic; the reference adds
// each term's forces with float atomics, bond_harmonic.cu, angle_harmonic.cu,
// dihedral_harmonic.cu). A term's energy and virial are counted once, by role
// 0 (bonds: the flag bit of the coefficient word). Records hold partner slots
// of the same position array; topo's per-slot arrays have stride n.
template <bool OBS>
__device__ __forceinline__ real3 bond_term(const real4* pos, const TopologyView& topo,
                                            const BoxView& box, real3 xi, uint2 r,
                                            StepSums& sums) {
  double e, w;
  const real3 f =
      bond_force(mic3(xi, f3(pos[r.x]), box), topo.bond_coeff[r.y & 0x7fffffffu], e, w);
  if (OBS && (r.y >> 31) != 0) {
    sums.e_bond += e;
    sums.w_bond += w;
  }
  return f;
}

template <bool OBS>
__device__ __forceinline__ real3 angle_term(const real4* pos, const TopologyView& topo,
                                             const BoxView& box, real3 xi, uint4 r,
                                             StepSums& sums) {
  const int role = static_cast<int>(r.w);
  const real3 a = f3(pos[r.x]);
  const real3 b = f3(pos[r.y]);
  const real3 pi = role == 0 ? xi : a;
  const real3 pj = role == 0 ? a : (role == 1 ? xi : b);
  const real3 pk = role == 2 ? xi : b;
  double e, w;
  const real3 f = angle_role_force(pi, pj, pk, topo.angle_coeff[r.z], role, box, e, w);
  if (OBS && role == 0) {
    sums.e_angle += e;
    sums.w_angle += w;
  }
  return f;
}

template <bool OBS>
__device__ __forceinline__ real3 dihedral_term(const real4* pos, const TopologyView& topo,
                                                const BoxView& box, real3 xi, uint4 r,
                                                StepSums& sums) {
  const int role = static_cast<int>(r.w >> 30);
  const real3 a = f3(pos[r.x]);
  const real3 b = f3(pos[r.y]);
  const real3 c = f3(pos[r.z]);
  real3 q0, q1, q2, q3;
  if (role == 0) { q0 = xi; q1 = a; q2 = b; q3 = c; }
  else if (role == 1) { q0 = a; q1 = xi; q2 = b; q3 = c; }
  else if (role == 2) { q0 = a; q1 = b; q2 = xi; q3 = c; }
  else { q0 = a; q1 = b; q2 = c; q3 = xi; }
  double e, w;
  const real3 f =
      dihedral_role_force(q0, q1, q2, q3, topo.dih_coeff[r.w & 0x3fffffffu], role, box, e, w);
  if (OBS && role == 0) {
    sums.e_dih += e;
    sums.w_dih += w;
  }
  return f;
}

// The listed terms of atom i shared over its team of TEAM lanes: lane t takes
// terms t, t + TEAM, ... of each kind and adds their forces to its FP64
// partial sums.
template <bool OBS, bool SPILL, int TEAM>
__device__ __forceinline__ void listed_team(const real4* pos, const TopologyView& topo,
                                            const BoxView& box, size_t n, uint32_t i, uint32_t t,
                                            real4 xi4, double& fxd, double& fyd, double& fzd,
                                            StepSums& sums) {
  const uint32_t cw = topo.count[i];
  const uint32_t nbond = cw & 255u, nang = (cw >> 8) & 255u, ndih = (cw >> 16) & 255u;
  const real3 xi = f3(xi4);
  auto add = [&](real3 f) {
    fxd += f.x;
    fyd += f.y;
    fzd += f.z;
  };
  for (uint32_t k = t; k < nbond; k += TEAM) {
    add(bond_term<OBS>(pos, topo, box, xi, topo.bond[k * n + i], sums));
  }
  for (uint32_t k = t; k < nang; k += TEAM) {
    add(angle_term<OBS>(pos, topo, box, xi, topo.angle[k * n + i], sums));
  }
  for (uint32_t k = t; k < ndih; k += TEAM) {
    add(dihedral_term<OBS>(pos, topo, box, xi, topo.dih[k * n + i], sums));
  }
  if (SPILL) {  // some particle spills: its records past the inline capacities
    const uint32_t si = topo.spill_index[i];
    if (si != kNoSpill) {  // uniform over the team
      const SpillEntry e = topo.spill[si];
      for (uint32_t k = t; k < e.cnt[0]; k += TEAM) {
        add(bond_term<OBS>(pos, topo, box, xi, topo.spill_bond[e.off[0] + k], sums));
      }
      for (uint32_t k = t; k < e.cnt[1]; k += TEAM) {
        add(angle_term<OBS>(pos, topo, box, xi, topo.spill_angle[e.off[1] + k], sums));
      }
      for (uint32_t k = t; k < e.cnt[2]; k += TEAM) {
        add(dihedral_term<OBS>(pos, topo, box, xi, topo.spill_dih[e.off[2] + k], sums));
      }
    }
  }
}

// Neighbour storage: each atom's list is cut into chunks of kNbrChunk entries;
// chunk c of atom i is kNbrChunk consecutive words at ((c * N) + i) * kNbrChunk.
// A build appends into an LDS staging row and writes whole 64-byte chunks; the
// force pass reads a chunk as four 16-byte loads.
__device__ __forceinline__ size_t nbr_chunk_base(uint32_t c, uint32_t i, uint32_t n) {
  return (static_cast<size_t>(c) * n + i) * kNbrChunk;
}

}  // namespace dev
}  // namespace beads_amd
