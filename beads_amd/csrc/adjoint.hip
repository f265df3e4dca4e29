// Adjoint of the force evaluation (no reference counterpart: the reference has
// no gradients). Given g = dL/dF for every atom and gE = dL/dE, one pass over
// the stored full neighbour list and the per-atom bond roles produces
//   * dL/dtheta for the LJ epsilon/sigma of every type pair and the harmonic
//     bond k/r0 of every bond type, accumulated in FP64, and
//   * dL/dx = J^T g - gE F, J = dF/dx (the Hessian-vector product), in FP64
//     per atom, rounded to real_t on store.
// The same pass serves force matching (g = dL/dF) and the reverse sweep of
// velocity-Verlet trajectory backprop (g = adjoint of the force).
//
// LJ pair, d = x_i - x_j, u = 1/r^2, sr6 = (sigma^2 u)^3, sr12 = sr6^2:
//   F_i = sum_j s d,                 s  = 24 eps u (2 sr12 - sr6)  (pair_lj.cu:56-66)
//   ds/d(r^2) = -24 eps u^2 (14 sr12 - 4 sr6)
//   ds/deps   = s / eps,  ds/dsigma = 24 eps u / sigma (24 sr12 - 6 sr6)
//   dL/dx_i  += s (g_i - g_j) + 2 ds/d(r^2) d (d . (g_i - g_j))   (symmetric tables)
//   dL/dtheta += (g_i . d) ds/dtheta        (each direction of the full list)
//   E = 1/2 sum_i sum_j [4 eps (sr12 - sr6) - shift],
//   shift = 4 eps (sc12 - sc6), sc6 = (sigma^2 / rc^2)^3 when the energy is shifted.
// Harmonic bond from the atom's side, d = x_self - x_partner, f = phi(r) d,
// phi = -k (r - r0) / r, E = k/2 (r - r0)^2 (bond_harmonic.cu:41-79):
//   dL/dx_self += phi (g_self - g_p) + (phi'/r) d (d . (g_self - g_p)), phi' = -k r0 / r^2
//   dL/dk  += g_self . (-(r - r0) / r) d,   dL/dr0 += g_self . (k / r) d
// Every bond is seen from both atoms, so each side adds half of dE/dk, dE/dr0.
// Harmonic angles and dihedrals: listed_backward_kernel below.
#include "adjoint.hpp"

#include <mutex>

#include "common.hpp"
#include "device_common.hpp"

namespace beads_amd {
namespace {

using namespace dev;

// Staging of dL/dF (real4 per atom) for frame_backward's allocation-free
// variant: one copy per GPU, held until the process ends. Host code must use
// the address from hipGetSymbolAddress (scratch_address below), never this
// name, which in host code is only a stand-in for the device symbol.
static __device__ __align__(256) real4 scratch[kNoAllocMaxAtoms];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
real4* g_scratch_ptr[kMaxDevices] = {};

template <bool CHECK>
__global__ __launch_bounds__(kBlock) void frame_load_kernel(FrameLoadParams p) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) return;
  real4 x = p.pos[i];
  x.x = p.xyz[3 * i];
  x.y = p.xyz[3 * i + 1];
  x.z = p.xyz[3 * i + 2];
  real_t c[3];
  {
#pragma clang fp contract(off)
    c[0] = floor((x.x - p.box.lo[0]) * p.box.invL[0]);
    c[1] = floor((x.y - p.box.lo[1]) * p.box.invL[1]);
    c[2] = floor((x.z - p.box.lo[2]) * p.box.invL[2]);
    x.x -= c[0] * p.box.L[0];
    x.y -= c[1] * p.box.L[1];
    x.z -= c[2] * p.box.L[2];
  }
  p.pos[i] = x;
  const int4 im = make_int4(static_cast<int>(c[0]), static_cast<int>(c[1]),
                            static_cast<int>(c[2]), 0);
  p.image[i] = im;
  if (CHECK) {
#pragma clang fp contract(off)
    const real4 r = p.ref_pos[i];
    const int4 ri = p.ref_image[i];
    const real_t dx = (x.x - r.x) + static_cast<real_t>(im.x - ri.x) * p.box.L[0];
    const real_t dy = (x.y - r.y) + static_cast<real_t>(im.y - ri.y) * p.box.L[1];
    const real_t dz = (x.z - r.z) + static_cast<real_t>(im.z - ri.z) * p.box.L[2];
    if (dx * dx + dy * dy + dz * dz > p.threshold_sq) {
      __hip_atomic_store(&p.ctrl->rebuild, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ __launch_bounds__(kBlock) void frame_store_kernel(uint32_t n, const real4* force,
                                                             real_t* xyz) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const real4 f = force[i];
  xyz[3 * i] = f.x;
  xyz[3 * i + 1] = f.y;
  xyz[3 * i + 2] = f.z;
}

__global__ __launch_bounds__(kBlock) void fill_f64_kernel(double* out, uint32_t count) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k < count) out[k] = 0.0;
}

__global__ __launch_bounds__(kBlock) void pack_xyz_kernel(uint32_t n, const real_t* xyz,
                                                          real4* out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  out[i] = make_real4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0);
}

// Lanes per atom and list slots per lane per round of the backward pass (the
// force pass's team layout: lane t takes slots base + u*kAdjTeam + t).
constexpr int kAdjTeam = 4;
constexpr int kAdjSlots = 4;
constexpr uint32_t kAdjAtoms = kBlock / kAdjTeam;

// Per type pair, the factors of the pair terms (d = x_i - x_j, u = 1/r^2,
// t1 = 2 sr12 - sr6):
//   s = e24 u t1,  ds/d(r^2) = -e24 u^2 (14 sr12 - 4 sr6),
//   ds/deps = 24 u t1,  ds/dsigma = cs u (4 sr12 - sr6)   (cs = 144 eps / sigma),
// and for dE/dtheta: 4 (sr12 - sr6) - sh_e and f4 (12 sr12 - 6 sr6) - sh_s
// (f4 = 4 eps / sigma; sh_e, sh_s the shift's derivatives, 0 when unshifted).
struct AdjLj {
  double sig2, e24, cs, f4, sh_e, sh_s;
};

__device__ __forceinline__ AdjLj adj_lj(const AdjointParams& p, int k) {
  const double eps = p.eps[k], sig = p.sigma[k];
  AdjLj c;
  c.sig2 = sig * sig;
  c.e24 = 24.0 * eps;
  c.cs = 144.0 * eps / sig;
  c.f4 = 4.0 * eps / sig;
  c.sh_e = 0.0;
  c.sh_s = 0.0;
  if (p.lj_shift) {
    const double sc2 = c.sig2 / p.rc2_f64;
    const double sc6 = sc2 * sc2 * sc2;
    const double sc12 = sc6 * sc6;
    c.sh_e = 4.0 * (sc12 - sc6);
    c.sh_s = c.f4 * (12.0 * sc12 - 6.0 * sc6);
  }
  return c;
}

// TREG > 0: the atom's row of the type table (at most TREG types) is
// accumulated in registers and flushed once per lane; TREG == 0: every
// interaction adds into LDS directly (any type count up to kMaxLdsTypes).
// One type: the pair factors live in registers; more: in an LDS table.
// GE: dL/dE is nonzero, so the pass also sums F for the -gE F term and the
// energy's parameter derivatives. Each atom's team sums its partial dL/dx in
// FP64 before the store.
template <int TREG, bool GE>
__global__ __launch_bounds__(kBlock) void fm_backward_kernel(AdjointParams p) {
  __shared__ double acc[2 * kMaxLdsTypes * kMaxLdsTypes + 2 * kMaxBondTypesAdjoint];
  __shared__ AdjLj coef[TREG == 1 ? 1 : kMaxLdsTypes * kMaxLdsTypes];
  const int T = p.n_types;
  const int n_pair = T * T;
  const int n_acc = 2 * n_pair + 2 * p.n_bond_types;
  for (int k = threadIdx.x; k < n_acc; k += kBlock) acc[k] = 0.0;
  if (TREG != 1) {
    for (int k = threadIdx.x; k < n_pair; k += kBlock) coef[k] = adj_lj(p, k);
  }
  __syncthreads();
  const AdjLj c0 = TREG == 1 ? adj_lj(p, 0) : AdjLj{};
  const uint32_t n = p.n;
  const uint32_t t = threadIdx.x & (kAdjTeam - 1);
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kAdjAtoms + threadIdx.x / kAdjTeam;
  const double gE = p.grad_energy;
  if (i < n) {  // uniform over a team
    const real4 xi = p.pos[i];
    const int ti = static_cast<int>(id_of(xi.w));
    const real4 gi = p.g[i];
    const double gix = gi.x, giy = gi.y, giz = gi.z;
    const real_t L0 = p.box.L[0], L1 = p.box.L[1], L2 = p.box.L[2];
    const real_t h0 = p.box.hL[0], h1 = p.box.hL[1], h2 = p.box.hL[2];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double re[TREG > 0 ? TREG : 1], rs[TREG > 0 ? TREG : 1];
#pragma unroll
    for (int q = 0; q < (TREG > 0 ? TREG : 1); ++q) re[q] = rs[q] = 0.0;
    const uint32_t cnt = p.nbr_count[i];
    for (uint32_t base = 0; base < cnt; base += kAdjTeam * kAdjSlots) {
      uint32_t jj[kAdjSlots];
#pragma unroll
      for (int u = 0; u < kAdjSlots; ++u) {
        const uint32_t s = base + u * kAdjTeam + t;
        jj[u] = s < cnt ? p.nbr[nbr_chunk_base(s / kNbrChunk, i, n) + (s & (kNbrChunk - 1))] : i;
      }
      real4 xj[kAdjSlots], gj[kAdjSlots];
#pragma unroll
      for (int u = 0; u < kAdjSlots; ++u) {
        xj[u] = p.pos[jj[u]];
        gj[u] = p.g[jj[u]];
      }
#pragma unroll
      for (int u = 0; u < kAdjSlots; ++u) {
        // mic_diff returns the raw difference unless it exceeds half the box;
        // the folded forms only when some lane of the wave needs them.
        real_t dx = xi.x - xj[u].x;
        real_t dy = xi.y - xj[u].y;
        real_t dz = xi.z - xj[u].z;
        if ((__builtin_amdgcn_ballot_w64(fabs(dx) > h0) |
             __builtin_amdgcn_ballot_w64(fabs(dy) > h1) |
             __builtin_amdgcn_ballot_w64(fabs(dz) > h2)) != 0) {
          dx = mic_diff(xi.x, xj[u].x, L0, h0);
          dy = mic_diff(xi.y, xj[u].y, L1, h1);
          dz = mic_diff(xi.z, xj[u].z, L2, h2);
        }
        if (base + u * kAdjTeam + t >= cnt || r2_strict(dx, dy, dz) >= p.rc2) continue;
        const int tj = static_cast<int>(id_of(xj[u].w));
        const AdjLj c = TREG == 1 ? c0 : coef[ti * T + tj];
        const double ddx = dx, ddy = dy, ddz = dz;
        const double u2 = rcp_f64(ddx * ddx + ddy * ddy + ddz * ddz);
        const double sr2 = c.sig2 * u2;
        const double sr6 = sr2 * sr2 * sr2;
        const double sr12 = sr6 * sr6;
        const double b = u2 * (2.0 * sr12 - sr6);
        const double s = c.e24 * b;
        const double sp2 = -2.0 * c.e24 * u2 * u2 * (14.0 * sr12 - 4.0 * sr6);
        const double ex = gix - gj[u].x, ey = giy - gj[u].y, ez = giz - gj[u].z;
        const double q2 = sp2 * (ddx * ex + ddy * ey + ddz * ez);
        gx += s * ex + q2 * ddx;
        gy += s * ey + q2 * ddy;
        gz += s * ez + q2 * ddz;
        if (GE) {
          fx += s * ddx;
          fy += s * ddy;
          fz += s * ddz;
        }
        const double w = gix * ddx + giy * ddy + giz * ddz;
        // dF/deps and dF/dsigma terms, plus half of dE/deps and dE/dsigma.
        double de = 24.0 * w * b;
        double ds = w * c.cs * u2 * (4.0 * sr12 - sr6);
        if (GE) {
          de += 0.5 * gE * (4.0 * (sr12 - sr6) - c.sh_e);
          ds += 0.5 * gE * (c.f4 * (12.0 * sr12 - 6.0 * sr6) - c.sh_s);
        }
        if constexpr (TREG > 0) {
#pragma unroll
          for (int q = 0; q < TREG; ++q) {
            if (tj == q) {
              re[q] += de;
              rs[q] += ds;
            }
          }
        } else {
          atomicAdd(&acc[ti * T + tj], de);
          atomicAdd(&acc[n_pair + ti * T + tj], ds);
        }
      }
    }
    if constexpr (TREG > 0) {
#pragma unroll
      for (int q = 0; q < TREG; ++q) {
        if (q < T) {
          if (re[q] != 0.0) atomicAdd(&acc[ti * T + q], re[q]);
          if (rs[q] != 0.0) atomicAdd(&acc[n_pair + ti * T + q], rs[q]);
        }
      }
    }
    if (p.bcap > 0) {
      const uint32_t nbond_inline = p.topo_count[i] & 255u;
      uint32_t nbond = nbond_inline, spill_off = 0;
      if (p.spill_index != nullptr) {
        const uint32_t si = p.spill_index[i];
        if (si != kNoSpill) {
          nbond += p.spill[si].cnt[0];
          spill_off = p.spill[si].off[0];
        }
      }
      const real3 x3 = f3(xi);
      for (uint32_t k = t; k < nbond; k += kAdjTeam) {
        const uint2 r = k < nbond_inline ? p.bond[static_cast<size_t>(k) * n + i]
                                         : p.spill_bond[spill_off + (k - nbond_inline)];
        const uint32_t bt = r.y & 0x7fffffffu;
        const double kk = p.bond_k[bt], r0 = p.bond_r0[bt];
        const real3 d = mic3(x3, f3(p.pos[r.x]), p.box);
        const double ddx = d.x, ddy = d.y, ddz = d.z;
        const double rr = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        if (rr == 0.0) continue;
        const double inv_r = rcp_f64(rr);
        const double phi = -kk * (rr - r0) * inv_r;
        const double dphi_r = -kk * r0 * inv_r * inv_r * inv_r;  // phi' / r
        const real4 gp = p.g[r.x];
        const double ex = gix - gp.x, ey = giy - gp.y, ez = giz - gp.z;
        const double dd = ddx * ex + ddy * ey + ddz * ez;
        gx += phi * ex + dphi_r * ddx * dd;
        gy += phi * ey + dphi_r * ddy * dd;
        gz += phi * ez + dphi_r * ddz * dd;
        if (GE) {
          fx += phi * ddx;
          fy += phi * ddy;
          fz += phi * ddz;
        }
        const double w = gix * ddx + giy * ddy + giz * ddz;
        const double stretch = rr - r0;
        atomicAdd(&acc[2 * n_pair + bt], w * (-stretch * inv_r) + gE * 0.25 * stretch * stretch);
        atomicAdd(&acc[2 * n_pair + p.n_bond_types + bt],
                  w * (kk * inv_r) - gE * 0.5 * kk * stretch);
      }
    }
    gx = team_sum<kAdjTeam>(gx);
    gy = team_sum<kAdjTeam>(gy);
    gz = team_sum<kAdjTeam>(gz);
    if (GE) {
      fx = team_sum<kAdjTeam>(fx);
      fy = team_sum<kAdjTeam>(fy);
      fz = team_sum<kAdjTeam>(fz);
    }
    if (t == 0) {
      p.grad_pos[3 * i] = static_cast<real_t>(gx - gE * fx);
      p.grad_pos[3 * i + 1] = static_cast<real_t>(gy - gE * fy);
      p.grad_pos[3 * i + 2] = static_cast<real_t>(gz - gE * fz);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n_acc; k += kBlock) {
    if (acc[k] != 0.0) atomicAdd(&p.grad_params[k], acc[k]);
  }
}

// ------------------------------------------------ angles and dihedrals ----
//
// The listed terms' part of dL/dx = J^T g - gE F and of the parameter
// gradients, in a pass of its own (after fm_backward_kernel, adding to its
// dL/dx). For a term with atoms a and the force F_b on atom b,
// (J^T g)_b = sum_a (dF_a/dx_b)^T g_a = sum_a (dF_b/dx_a) g_a, since the
// blocks dF_a/dx_b = -d2U/dx_a dx_b are each other's transposes: the
// directional derivative of F_b when every atom a of the term moves along
// g_a. Each role evaluates the term's force on its atom in forward-mode FP64
// arithmetic (value, derivative along g) following angle_role_force and
// dihedral_role_force (device_common.hpp), with theta and phi from FP64
// atan2. Parameters: F is linear in k, and for angles F = A G with
// A = -k (theta - theta0) / sin, so dF/dtheta0 = (k / sin) G; the role-0
// record of each term adds dE/dk and dE/dtheta0 once.

struct dual {
  double v, d;  // value and derivative along the atoms' g
};
__device__ __forceinline__ dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ dual operator*(dual a, double s) { return {a.v * s, a.d * s}; }
__device__ __forceinline__ dual dual_rsq(dual x) {
  const double r = rsq_f64(x.v);
  return {r, -0.5 * r * r * r * x.d};
}
__device__ __forceinline__ dual dual_rcp(dual x) {
  const double r = rcp_f64(x.v);
  return {r, -r * r * x.d};
}
__device__ __forceinline__ dual dual_atan2(dual y, dual x) {
  const double q = x.v * x.v + y.v * y.v;
  return {atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / q};
}
struct dv3 {
  dual x, y, z;
};
__device__ __forceinline__ dv3 operator+(dv3 a, dv3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ dv3 operator-(dv3 a, dv3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ dv3 operator*(dv3 a, dual s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ dv3 operator*(dv3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ dual ddot(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ dv3 dcross(dv3 a, dv3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Minimum-image displacement p - q (the forward's mic3) with derivative gp - gq.
__device__ __forceinline__ dv3 dual_disp(real3 p, real3 q, real4 gp, real4 gq, const BoxView& box) {
  const real3 d = mic3(p, q, box);
  return {{d.x, static_cast<double>(gp.x) - gq.x},
          {d.y, static_cast<double>(gp.y) - gq.y},
          {d.z, static_cast<double>(gp.z) - gq.z}};
}
__device__ __forceinline__ double gdot(real4 g, double x, double y, double z) {
  return static_cast<double>(g.x) * x + static_cast<double>(g.y) * y + static_cast<double>(g.z) * z;
}

// One listed role's contribution: force on the role's atom (value and
// derivative), and the parameter gradient terms it adds.
struct RoleAdjoint {
  dv3 f;
  double dk = 0.0, dtheta0 = 0.0;  // g_self . dF/dparam, plus the term's gE dE/dparam on role 0
};

// angle_role_force's mathematics; atoms (i, j, k) with j the vertex.
__device__ RoleAdjoint angle_adjoint(real3 pi, real3 pj, real3 pk, real4 gi, real4 gj, real4 gk,
                                     double k, double theta0, int role, real4 gself, double gE,
                                     const BoxView& box) {
  RoleAdjoint out;
  out.f = {{0, 0}, {0, 0}, {0, 0}};
  const dv3 a = dual_disp(pi, pj, gi, gj, box);
  const dv3 b = dual_disp(pk, pj, gk, gj, box);
  const dual ra2 = ddot(a, a), rb2 = ddot(b, b);
  if (ra2.v == 0.0 || rb2.v == 0.0) {
    if (role == 0) out.dk = gE * 0.5 * theta0 * theta0, out.dtheta0 = gE * k * theta0;
    return out;
  }
  const dual ia = dual_rsq(ra2), ib = dual_rsq(rb2);
  const dv3 nrm = dcross(a, b);
  const dual nn2 = ddot(nrm, nrm);
  const dual nn = nn2.v > 0.0 ? nn2 * dual_rsq(nn2) : dual{0.0, 0.0};
  const dual theta = dual_atan2(nn, ddot(a, b));
  const dual dth = theta - dual{theta0, 0.0};
  const dual sn_true = nn * ia * ib;
  constexpr double kSinFloor = static_cast<double>(static_cast<real_t>(0.001));
  const dual sn = sn_true.v <= kSinFloor ? dual{kSinFloor, 0.0} : sn_true;
  const dual A = dth * dual_rcp(sn) * (-k);
  const dual aib = A * ia * ib;
  const dv3 ai = dcross(a, nrm), bk = dcross(b, nrm);
  const dv3 fi = ai * (aib * ia * ia);
  const dv3 fk = bk * (aib * ib * ib * -1.0);
  // G: the force with A = 1 (value only).
  const double gi_s = ia.v * ib.v * ia.v * ia.v, gk_s = -ia.v * ib.v * ib.v * ib.v;
  double Gx, Gy, Gz;
  if (role == 0) {
    out.f = fi;
    Gx = ai.x.v * gi_s, Gy = ai.y.v * gi_s, Gz = ai.z.v * gi_s;
  } else if (role == 2) {
    out.f = fk;
    Gx = bk.x.v * gk_s, Gy = bk.y.v * gk_s, Gz = bk.z.v * gk_s;
  } else {
    out.f = (fi + fk) * -1.0;
    Gx = -(ai.x.v * gi_s + bk.x.v * gk_s);
    Gy = -(ai.y.v * gi_s + bk.y.v * gk_s);
    Gz = -(ai.z.v * gi_s + bk.z.v * gk_s);
  }
  const double isn = 1.0 / sn.v;
  const double wg = gdot(gself, Gx, Gy, Gz);
  out.dk = wg * (-dth.v * isn);
  out.dtheta0 = wg * (k * isn);
  if (role == 0) {
    out.dk += gE * 0.5 * dth.v * dth.v;
    out.dtheta0 += gE * (-k * dth.v);
  }
  return out;
}

// dihedral_role_force's mathematics; atoms (0, 1, 2, 3) along the chain.
__device__ RoleAdjoint dihedral_adjoint(real3 q0, real3 q1, real3 q2, real3 q3, real4 g0, real4 g1,
                                        real4 g2, real4 g3, double k, double dsign, double nmul,
                                        int role, real4 gself, double gE, const BoxView& box) {
  RoleAdjoint out;
  out.f = {{0, 0}, {0, 0}, {0, 0}};
  const dv3 v0 = dual_disp(q0, q1, g0, g1, box);
  const dv3 v1 = dual_disp(q2, q1, g2, g1, box);
  const dv3 v2 = dual_disp(q2, q3, g2, g3, box);
  const dv3 c0 = dcross(v0, v1), c1 = dcross(v1, v2);
  const dual n1 = ddot(v1, v1), nc0 = ddot(c0, c0), nc1 = ddot(c1, c1);
  if (n1.v == 0.0 || nc0.v == 0.0 || nc1.v == 0.0) {
    if (role == 0) out.dk = gE * (nmul == 0.0 ? 1.0 + dsign : 1.0);
    return out;
  }
  const dual inb = dual_rsq(n1);
  const dual nb = n1 * inb;
  const dual ic0 = dual_rsq(nc0), ic1 = dual_rsq(nc1);
  const dual inv_c = ic0 * ic1;
  const dual c1phi = ddot(c0, c1) * inv_c;
  const dual s1phi = nb * ddot(v0, c1) * inv_c;
  dual cs{1.0, 0.0}, sn{0.0, 0.0};
  for (int m = 0; m < static_cast<int>(nmul); ++m) {
    const dual c2 = cs * c1phi - sn * s1phi;
    sn = sn * c1phi + cs * s1phi;
    cs = c2;
  }
  // F is linear in k: evaluated at k = 1 (dF/dk), then scaled.
  const dual dE1 = sn * (-dsign * nmul);
  const dv3 fi = c0 * (dE1 * nb * ic0 * ic0 * -1.0);
  const dv3 fl = c1 * (dE1 * nb * ic1 * ic1);
  const dual inv_n1 = inb * inb;
  const dv3 sv = fi * (ddot(v0, v1) * inv_n1) - fl * (ddot(v2, v1) * inv_n1);
  dv3 f1;
  if (role == 0) {
    f1 = fi;
  } else if (role == 1) {
    f1 = sv - fi;
  } else if (role == 2) {
    f1 = (sv + fl) * -1.0;
  } else {
    f1 = fl;
  }
  out.f = f1 * k;
  out.dk = gdot(gself, f1.x.v, f1.y.v, f1.z.v);
  if (role == 0) out.dk += gE * (1.0 + dsign * cs.v);
  return out;
}

constexpr int kListTeam = 4;
constexpr uint32_t kListAtoms = kBlock / kListTeam;

__global__ __launch_bounds__(kBlock) void listed_backward_kernel(AdjointParams p) {
  __shared__ double acc[2 * kMaxAngleTypesAdjoint + kMaxDihedralTypesAdjoint];
  const int na = p.n_angle_types, nd = p.n_dihedral_types;
  const int n_acc = 2 * na + nd;
  for (int k = threadIdx.x; k < n_acc; k += kBlock) acc[k] = 0.0;
  __syncthreads();
  const uint32_t n = p.n;
  const uint32_t t = threadIdx.x & (kListTeam - 1);
  const uint32_t i = blockIdx.x * kListAtoms + threadIdx.x / kListTeam;
  const double gE = p.grad_energy;
  if (i < n) {  // uniform over a team
    const real3 xi = f3(p.pos[i]);
    const real4 gi = p.g[i];
    const uint32_t cw = p.topo_count[i];
    const uint32_t ang_inline = p.acap > 0 ? (cw >> 8) & 255u : 0u;
    const uint32_t dih_inline = p.dcap > 0 ? (cw >> 16) & 255u : 0u;
    uint32_t nang = ang_inline, ndih = dih_inline, off_a = 0, off_d = 0;
    if (p.spill_index != nullptr) {
      const uint32_t si = p.spill_index[i];
      if (si != kNoSpill) {
        nang += p.spill[si].cnt[1];
        off_a = p.spill[si].off[1];
        ndih += p.spill[si].cnt[2];
        off_d = p.spill[si].off[2];
      }
    }
    double gx = 0.0, gy = 0.0, gz = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    auto add = [&](const RoleAdjoint& r) {
      gx += r.f.x.d;
      gy += r.f.y.d;
      gz += r.f.z.d;
      fx += r.f.x.v;
      fy += r.f.y.v;
      fz += r.f.z.v;
    };
    for (uint32_t k = t; k < nang; k += kListTeam) {
      const uint4 r = k < ang_inline ? p.angle[static_cast<size_t>(k) * n + i]
                                     : p.spill_angle[off_a + (k - ang_inline)];
      const int role = static_cast<int>(r.w);
      const real3 a = f3(p.pos[r.x]), b = f3(p.pos[r.y]);
      const real4 ga = p.g[r.x], gb = p.g[r.y];
      const real3 pi = role == 0 ? xi : a;
      const real3 pj = role == 0 ? a : (role == 1 ? xi : b);
      const real3 pk = role == 2 ? xi : b;
      const real4 qi = role == 0 ? gi : ga;
      const real4 qj = role == 0 ? ga : (role == 1 ? gi : gb);
      const real4 qk = role == 2 ? gi : gb;
      const RoleAdjoint ra = angle_adjoint(pi, pj, pk, qi, qj, qk, p.angle_k[r.z],
                                           p.angle_theta0[r.z], role, gi, gE, p.box);
      add(ra);
      atomicAdd(&acc[r.z], ra.dk);
      atomicAdd(&acc[na + r.z], ra.dtheta0);
    }
    for (uint32_t k = t; k < ndih; k += kListTeam) {
      const uint4 r = k < dih_inline ? p.dih[static_cast<size_t>(k) * n + i]
                                     : p.spill_dih[off_d + (k - dih_inline)];
      const int role = static_cast<int>(r.w >> 30);
      const uint32_t ty = r.w & 0x3fffffffu;
      const real3 a = f3(p.pos[r.x]), b = f3(p.pos[r.y]), c = f3(p.pos[r.z]);
      const real4 ga = p.g[r.x], gb = p.g[r.y], gc = p.g[r.z];
      real3 q0, q1, q2, q3;
      real4 h0, h1, h2, h3;
      if (role == 0) { q0 = xi; q1 = a; q2 = b; q3 = c; h0 = gi; h1 = ga; h2 = gb; h3 = gc; }
      else if (role == 1) { q0 = a; q1 = xi; q2 = b; q3 = c; h0 = ga; h1 = gi; h2 = gb; h3 = gc; }
      else if (role == 2) { q0 = a; q1 = b; q2 = xi; q3 = c; h0 = ga; h1 = gb; h2 = gi; h3 = gc; }
      else { q0 = a; q1 = b; q2 = c; q3 = xi; h0 = ga; h1 = gb; h2 = gc; h3 = gi; }
      const RoleAdjoint rd = dihedral_adjoint(q0, q1, q2, q3, h0, h1, h2, h3, p.dih_k[ty],
                                              p.dih_d[ty], p.dih_n[ty], role, gi, gE, p.box);
      add(rd);
      atomicAdd(&acc[2 * na + ty], rd.dk);
    }
    gx = team_sum<kListTeam>(gx);
    gy = team_sum<kListTeam>(gy);
    gz = team_sum<kListTeam>(gz);
    fx = team_sum<kListTeam>(fx);
    fy = team_sum<kListTeam>(fy);
    fz = team_sum<kListTeam>(fz);
    if (t == 0 && (nang | ndih) != 0) {
      // fm_backward_kernel stored the pair and bond part of dL/dx.
      real_t* o = p.grad_pos + 3 * static_cast<size_t>(i);
      o[0] = static_cast<real_t>(static_cast<double>(o[0]) + (gx - gE * fx));
      o[1] = static_cast<real_t>(static_cast<double>(o[1]) + (gy - gE * fy));
      o[2] = static_cast<real_t>(static_cast<double>(o[2]) + (gz - gE * fz));
    }
  }
  __syncthreads();
  const size_t base = 2 * static_cast<size_t>(p.n_types) * p.n_types + 2 * p.n_bond_types;
  for (int k = threadIdx.x; k < n_acc; k += kBlock) {
    if (acc[k] != 0.0) atomicAdd(&p.grad_params[base + k], acc[k]);
  }
}

}  // namespace

void launch_frame_load(const FrameLoadParams& p, bool check, hipStream_t stream) {
  if (check) {
    frame_load_kernel<true><<<grid_for(p.n), kBlock, 0, stream>>>(p);
  } else {
    frame_load_kernel<false><<<grid_for(p.n), kBlock, 0, stream>>>(p);
  }
}

void launch_frame_store(uint32_t n, const real4* force, real_t* xyz, hipStream_t stream) {
  frame_store_kernel<<<grid_for(n), kBlock, 0, stream>>>(n, force, xyz);
}

void launch_fill_f64(double* out, uint32_t count, hipStream_t stream) {
  if (count == 0) return;
  fill_f64_kernel<<<grid_for(count), kBlock, 0, stream>>>(out, count);
}

// Device address of this GPU's scratch array, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
real4* scratch_address() {
  int device = 0;
  BEADS_HIP(hipGetDevice(&device));
  require(device >= 0 && device < kMaxDevices, "device index out of range for the scratch table",
          BEADS_E_RUNTIME);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    BEADS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array",
            BEADS_E_DEVICE);
    g_scratch_ptr[device] = static_cast<real4*>(ptr);
  }
  return g_scratch_ptr[device];
}

void launch_pack_xyz(uint32_t n, const real_t* xyz, real4* out, hipStream_t stream) {
  pack_xyz_kernel<<<grid_for(n), kBlock, 0, stream>>>(n, xyz, out);
}

void launch_fm_backward(const AdjointParams& p, hipStream_t stream) {
  const unsigned g = grid_for(p.n, kAdjAtoms);
  const bool ge = p.grad_energy != 0.0;
  if (p.n_types <= 1) {
    if (ge) fm_backward_kernel<1, true><<<g, kBlock, 0, stream>>>(p);
    else fm_backward_kernel<1, false><<<g, kBlock, 0, stream>>>(p);
  } else if (p.n_types <= 4) {
    if (ge) fm_backward_kernel<4, true><<<g, kBlock, 0, stream>>>(p);
    else fm_backward_kernel<4, false><<<g, kBlock, 0, stream>>>(p);
  } else {
    if (ge) fm_backward_kernel<0, true><<<g, kBlock, 0, stream>>>(p);
    else fm_backward_kernel<0, false><<<g, kBlock, 0, stream>>>(p);
  }
  if (p.acap > 0 || p.dcap > 0) {
    listed_backward_kernel<<<grid_for(p.n, kListAtoms), kBlock, 0, stream>>>(p);
  }
}

}  // namespace beads_amd
