// Batched replicas: many independent LJ systems advanced together, in one of
// two modes.
//
// Persistent mode: ONE launch for all steps. Each workgroup owns one replica
// and runs its whole step loop --
// integrate + displacement check, the cell-list rebuild with stable sort and
// reorder, the full-list LJ force with the second half-kick and the Berendsen
// factor -- with workgroup barriers between phases and no grid-wide sync.
// Replicas are independent, so the launch needs no inter-workgroup
// communication; uneven sizes are balanced by dispatching the largest replicas
// first (the host orders the descriptors) and by the autotuned workgroup size.
//
// Phased mode: each step is a sequence of batched kernels whose boundaries
// stand for the persistent kernel's workgroup barriers. Per-atom work goes by
// chunks of kReplicaChunk atoms of one replica (one wave each), so a replica
// spreads over as many waves as its size needs and no replica sets a critical
// path; the rebuild kernels walk a device-built list of the chunks of the
// replicas that rebuild in this step, so a step without rebuilds costs them a
// launch each and no more. Both modes share the per-atom code below (integrate
// and check, list build, team force), so they make the same decisions with
// the same arithmetic; the phased sort ranks a cell's members by slot instead
// of insertion-sorting them (the same order).
//
// Per replica the arithmetic is the single-system path's (same decisions, same
// FP64 pair values and reductions): velocity_verlet.cu:59-111,
// berendsen_thermostat.cu:104-163, cell_list_build.cu:110-219,
// full_neighbor_path_common.cuh:49-94, regular/small_dimension/direct builders,
// pair_lj.cu:21-167, neighbor_system.cpp:79-146,198-216 (paths under the
// reference's native/src).
#include "replicas.hpp"

#include <algorithm>

#include "device_common.hpp"

namespace beads_amd {
namespace {

using namespace dev;

// Force-pass team layout inside a replica's workgroup.
constexpr int kReplicaTeam = 4;
constexpr int kReplicaSlots = 4;

template <int BLOCK>
struct ReplicaShared {
  double red[BLOCK / 64];
  uint32_t scan[BLOCK];
  LjCoeff lj[kMaxLdsTypes * kMaxLdsTypes];
  double lambda;
  uint32_t flag;
};

// Exclusive scan of count[0..nc) into start[0..nc] with one workgroup.
template <int BLOCK>
__device__ void block_scan(const uint32_t* count, uint32_t* start, uint32_t* cursor, uint32_t nc,
                           uint32_t* sums) {
  const uint32_t t = threadIdx.x;
  const uint32_t chunk = (nc + BLOCK - 1) / BLOCK;
  const uint32_t begin = min(nc, t * chunk), end = min(nc, begin + chunk);
  uint32_t local = 0;
  for (uint32_t c = begin; c < end; ++c) local += count[c];
  sums[t] = local;
  __syncthreads();
  for (uint32_t off = 1; off < BLOCK; off <<= 1) {
    const uint32_t v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  uint32_t run = sums[t] - local;
  for (uint32_t c = begin; c < end; ++c) {
    const uint32_t cnt = count[c];
    start[c] = run;
    cursor[c] = run;
    run += cnt;
  }
  if (t == BLOCK - 1) start[nc] = sums[t];
  __syncthreads();
}

// A replica's lists are stored in chunks of kNbrChunk entries (the plain
// path's layout): entry s of atom i at nbr_chunk_base(s / kNbrChunk, i, n) +
// s % kNbrChunk, so a team's kNbrChunk consecutive slots are one 64-byte run.
__device__ __forceinline__ size_t replica_nbr_index(uint32_t s, uint32_t i, uint32_t n) {
  return nbr_chunk_base(s / kNbrChunk, i, n) + (s & (kNbrChunk - 1));
}

// A replica's listed-term records and exclusions as the single-system view
// (per-slot arrays of stride n, partner slots local to the replica, no
// spill).
__device__ __forceinline__ TopologyView replica_topology(const ReplicaParams& P,
                                                         const ReplicaDesc& d) {
  const ReplicaArrays& A = P.arrays;
  TopologyView v{};
  v.count = A.topo_count + d.atom0;
  v.bond = A.bond + static_cast<size_t>(P.bcap) * d.atom0;
  v.angle = A.angle + static_cast<size_t>(P.acap) * d.atom0;
  v.dih = A.dih + static_cast<size_t>(P.dcap) * d.atom0;
  v.excl = A.excl + static_cast<size_t>(P.ecap) * d.atom0;
  v.bcap = P.bcap;
  v.acap = P.acap;
  v.dcap = P.dcap;
  v.ecap = P.ecap;
  v.bond_coeff = P.bond_coeff + d.bond_t0;
  v.angle_coeff = P.angle_coeff + d.angle_t0;
  v.dih_coeff = P.dih_coeff + d.dih_t0;
  v.has_listed = d.has_listed;
  return v;
}

// Whether local slot j is one of the ne excluded partners of local slot i.
__device__ __forceinline__ bool excluded_partner(const uint32_t* excl, uint32_t n, uint32_t i,
                                                 uint32_t ne, uint32_t j) {
  bool hit = false;
  for (uint32_t k = 0; k < ne; ++k) hit |= excl[static_cast<size_t>(k) * n + i] == j;
  return hit;
}

// Reorder of a replica's records: new slot s takes old slot o's records
// (gather) and, once every gather of the replica is done, its partner slots
// are translated through new_of_old (commit); cell_list_build.cu:183-219 for
// the particle arrays, the single-system reorder kernels for the records.
__device__ __forceinline__ void topology_gather(const ReplicaParams& P, const ReplicaDesc& d,
                                                uint32_t s, uint32_t o) {
  const ReplicaArrays& A = P.arrays;
  const size_t n = d.n, a0 = d.atom0;
  A.new_of_old[a0 + o] = s;
  A.s_topo_count[a0 + s] = A.topo_count[a0 + o];
  const uint2* bond = A.bond + P.bcap * a0;
  const uint4* angle = A.angle + P.acap * a0;
  const uint4* dih = A.dih + P.dcap * a0;
  const uint32_t* excl = A.excl + P.ecap * a0;
  for (uint32_t k = 0; k < P.bcap; ++k) A.s_bond[P.bcap * a0 + k * n + s] = bond[k * n + o];
  for (uint32_t k = 0; k < P.acap; ++k) A.s_angle[P.acap * a0 + k * n + s] = angle[k * n + o];
  for (uint32_t k = 0; k < P.dcap; ++k) A.s_dih[P.dcap * a0 + k * n + s] = dih[k * n + o];
  for (uint32_t k = 0; k < P.ecap; ++k) A.s_excl[P.ecap * a0 + k * n + s] = excl[k * n + o];
}

__device__ __forceinline__ void topology_commit(const ReplicaParams& P, const ReplicaDesc& d,
                                                uint32_t s) {
  const ReplicaArrays& A = P.arrays;
  const size_t n = d.n, a0 = d.atom0;
  const uint32_t* m = A.new_of_old + a0;
  const uint32_t cw = A.s_topo_count[a0 + s];
  A.topo_count[a0 + s] = cw;
  const uint32_t nb = cw & 255u, na = (cw >> 8) & 255u, nd = (cw >> 16) & 255u, ne = cw >> 24;
  for (uint32_t k = 0; k < nb; ++k) {
    uint2 r = A.s_bond[P.bcap * a0 + k * n + s];
    r.x = m[r.x];
    A.bond[P.bcap * a0 + k * n + s] = r;
  }
  for (uint32_t k = 0; k < na; ++k) {
    uint4 r = A.s_angle[P.acap * a0 + k * n + s];
    r.x = m[r.x];
    r.y = m[r.y];
    A.angle[P.acap * a0 + k * n + s] = r;
  }
  for (uint32_t k = 0; k < nd; ++k) {
    uint4 r = A.s_dih[P.dcap * a0 + k * n + s];
    r.x = m[r.x];
    r.y = m[r.y];
    r.z = m[r.z];
    A.dih[P.dcap * a0 + k * n + s] = r;
  }
  for (uint32_t k = 0; k < ne; ++k) {
    A.excl[P.ecap * a0 + k * n + s] = m[A.s_excl[P.ecap * a0 + k * n + s]];
  }
}

// Block totals of a force pass's sums (kinetic energy; with obs also the
// four energies and virials) into out[kReplicaSums].
template <int BLOCK>
__device__ __forceinline__ void step_totals(const StepSums& sums, bool obs, double* out,
                                            double* red) {
  out[0] = block_sum<BLOCK>(sums.ke, red);
  if (obs) {
    const double v[8] = {sums.e_pair, sums.e_bond, sums.e_angle, sums.e_dih,
                         sums.w_pair, sums.w_bond, sums.w_angle, sums.w_dih};
#pragma unroll
    for (int q = 0; q < 8; ++q) out[q + 1] = block_sum<BLOCK>(v[q], red);
  }
}

// Velocity-Verlet first half-kick (with the pending Berendsen factor), drift,
// wrap and image update of atom a of a replica (arrays offset to the replica),
// and, when check, the displacement test against the reference positions.
// Returns 1 when the atom moved farther than the half skin.
__device__ __forceinline__ int integrate_atom(const ReplicaDesc& d, real4* pos, real4* vel,
                                              const real4* force, int4* image,
                                              const real4* ref_pos, const int4* ref_image,
                                              uint32_t a, real_t lam, bool check) {
  real4 v = vel[a];
  const real4 f = force[a];
  real4 x = pos[a];
  if (d.berendsen) {
    v.x *= lam;
    v.y *= lam;
    v.z *= lam;
  }
  const real_t h = real_t(0.5) * d.dt / v.w;
  v.x += h * f.x;
  v.y += h * f.y;
  v.z += h * f.z;
  x.x += d.dt * v.x;
  x.y += d.dt * v.y;
  x.z += d.dt * v.z;
  real_t c[3];
  {
#pragma clang fp contract(off)
    c[0] = floor((x.x - d.box.lo[0]) * d.box.invL[0]);
    c[1] = floor((x.y - d.box.lo[1]) * d.box.invL[1]);
    c[2] = floor((x.z - d.box.lo[2]) * d.box.invL[2]);
    x.x -= c[0] * d.box.L[0];
    x.y -= c[1] * d.box.L[1];
    x.z -= c[2] * d.box.L[2];
  }
  pos[a] = x;
  vel[a] = v;
  int4 im = image[a];
  if (c[0] != real_t(0) || c[1] != real_t(0) || c[2] != real_t(0)) {
    im.x += static_cast<int>(c[0]);
    im.y += static_cast<int>(c[1]);
    im.z += static_cast<int>(c[2]);
    image[a] = im;
  }
  int exceeded = 0;
  if (check) {
#pragma clang fp contract(off)
    const real4 r = ref_pos[a];
    const int4 ri = ref_image[a];
    const real_t dx = (x.x - r.x) + static_cast<real_t>(im.x - ri.x) * d.box.L[0];
    const real_t dy = (x.y - r.y) + static_cast<real_t>(im.y - ri.y) * d.box.L[1];
    const real_t dz = (x.z - r.z) + static_cast<real_t>(im.z - ri.z) * d.box.L[2];
    if (dx * dx + dy * dy + dz * dz > d.thr2) exceeded = 1;
  }
  return exceeded;
}

// Full neighbour list of atom i of a replica (arrays offset to the replica):
// candidates in the 27 (or per-axis unique) neighbour cells in neighbourhood
// order, members of a cell in slot order, or every atom on the direct path,
// less i's ne excluded partners (excl: the replica's exclusion records).
// Captures the reference position. Returns 1 on overflow.
__device__ __forceinline__ uint32_t build_atom(const ReplicaDesc& d, const real4* pos,
                                               const int4* image, const uint32_t* cell_of,
                                               const uint32_t* cell_atoms, const uint32_t* start,
                                               uint32_t* nbr, uint32_t* nbr_count,
                                               real4* ref_pos, int4* ref_image,
                                               const uint32_t* excl, uint32_t ne, uint32_t i) {
  const uint32_t n = d.n;
  const real_t L0 = d.box.L[0], L1 = d.box.L[1], L2 = d.box.L[2];
  const real_t h0 = d.box.hL[0], h1 = d.box.hL[1], h2 = d.box.hL[2];
  uint32_t overflow = 0;
  const real4 xi = pos[i];
  uint32_t cnt = 0;
  auto visit = [&](uint32_t j) {
    const real4 xj = pos[j];
    const real_t dx = mic(xi.x - xj.x, L0, h0);
    const real_t dy = mic(xi.y - xj.y, L1, h1);
    const real_t dz = mic(xi.z - xj.z, L2, h2);
    if (r2_strict(dx, dy, dz) > d.rs2 || j == i) return;
    if (ne != 0 && excluded_partner(excl, n, i, ne, j)) return;
    if (cnt < d.max_neighbors) {
      nbr[replica_nbr_index(cnt, i, n)] = j;
      ++cnt;
    } else {
      overflow = 1;
    }
  };
  if (d.grid.path == 0) {
    for (uint32_t j = 0; j < n; ++j) visit(j);
  } else {
    const CellNeighborhood h = neighborhood(cell_of[i], d.grid);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a >= h.nz) break;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        if (b >= h.ny) break;
        const uint32_t row = d.grid.nx * (h.ys[b] + d.grid.ny * h.zs[a]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (c >= h.nx) break;
          const uint32_t cell = h.xs[c] + row;
          for (uint32_t k = start[cell]; k < start[cell + 1]; ++k) visit(cell_atoms[k]);
        }
      }
    }
  }
  nbr_count[i] = cnt;
  ref_pos[i] = xi;
  ref_image[i] = image[i];
  return overflow;
}

// Full-list LJ force plus the listed terms of atom i of a replica (arrays
// offset to the replica) by a team of TEAM adjacent lanes (lane t of the
// team): lane t takes list slots base + u*TEAM + t, SLOTS of them in flight
// per round, then listed terms t, t + TEAM, ...; the team's FP64 partial
// forces meet in a butterfly and lane 0 stores the force and, when kick, the
// second half-kick. Adds this lane's share of the kinetic energy (lane 0)
// and, when OBS, of the energies and virials.
template <int TEAM, int SLOTS, bool OBS, bool LISTED>
__device__ __forceinline__ void force_atom(const ReplicaDesc& d, const real4* pos, real4* vel,
                                           real4* force, const uint32_t* nbr_count,
                                           const uint32_t* nbr, const LjCoeff* lj,
                                           const TopologyView& topo, bool kick, bool single,
                                           uint32_t i, uint32_t t, StepSums& sums) {
  const uint32_t n = d.n;
  const real_t L0 = d.box.L[0], L1 = d.box.L[1], L2 = d.box.L[2];
  const real_t h0 = d.box.hL[0], h1 = d.box.hL[1], h2 = d.box.hL[2];
  const int T = d.n_types;
  const LjCoeff c0 = lj[0];
  const real4 xi = pos[i];
  const int ti = static_cast<int>(id_of(xi.w));
  double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, w = 0.0;
  const uint32_t cnt = nbr_count[i];
  for (uint32_t base = 0; base < cnt; base += TEAM * SLOTS) {
    uint32_t jj[SLOTS];
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
      const uint32_t s = base + u * TEAM + t;
      jj[u] = s < cnt ? nbr[replica_nbr_index(s, i, n)] : i;
    }
    real4 xj[SLOTS];
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
      xj[u] = pos[jj[u]];
    }
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
      // mic_diff returns the raw difference unless it exceeds half the box;
      // the folded forms only when some lane of the wave needs them.
      real_t dx = xi.x - xj[u].x;
      real_t dy = xi.y - xj[u].y;
      real_t dz = xi.z - xj[u].z;
      if ((__builtin_amdgcn_ballot_w64(fabs(dx) > h0) | __builtin_amdgcn_ballot_w64(fabs(dy) > h1) |
           __builtin_amdgcn_ballot_w64(fabs(dz) > h2)) != 0) {
        dx = mic_diff(xi.x, xj[u].x, L0, h0);
        dy = mic_diff(xi.y, xj[u].y, L1, h1);
        dz = mic_diff(xi.z, xj[u].z, L2, h2);
      }
      if (base + u * TEAM + t < cnt && r2_strict(dx, dy, dz) < d.rc2) {
        const LjCoeff c = single ? c0 : lj[ti * T + static_cast<int>(id_of(xj[u].w))];
        double rr, uu;
        const double s = lj_pair(dx, dy, dz, c, rr, uu);
        fx += s * dx;
        fy += s * dy;
        fz += s * dz;
        if (OBS) {
          e += uu;
          w += s * rr;
        }
      }
    }
  }
  if (LISTED && d.has_listed) {
    listed_team<OBS, false, TEAM>(pos, topo, d.box, n, i, t, xi, fx, fy, fz, sums);
  }
  fx = team_sum<TEAM>(fx);
  fy = team_sum<TEAM>(fy);
  fz = team_sum<TEAM>(fz);
  if (OBS) {
    sums.e_pair += 0.5 * e;
    sums.w_pair += 0.5 * w;
  }
  if (t != 0) {
    return;
  }
  const real_t ffx = static_cast<real_t>(fx), ffy = static_cast<real_t>(fy),
              ffz = static_cast<real_t>(fz);
  force[i] = make_real4(ffx, ffy, ffz, 0);
  real4 v = vel[i];
  if (kick) {
    const real_t h = real_t(0.5) * d.dt / v.w;
    v.x += h * ffx;
    v.y += h * ffy;
    v.z += h * ffz;
    vel[i] = v;
  }
  const double vx = v.x, vy = v.y, vz = v.z;
  sums.ke += 0.5 * static_cast<double>(v.w) * (vx * vx + vy * vy + vz * vz);
}

// Rebuild decisions of one replica at step now (NeighborSystem::rebuild_now's
// counters): returns whether the rebuild also reorders.
__device__ __forceinline__ bool begin_rebuild(const ReplicaDesc& d, ReplicaState* st,
                                              uint64_t now) {
  const uint64_t builds = st->build_count + 1;
  const bool due = d.sort_every_steps != 0 ? now - st->last_sort_step >= d.sort_every_steps
                                           : builds % d.sort_every == 0;
  const bool reorder = d.grid.path != 0 && due;
  st->build_count = builds;
  if (reorder) {
    st->reorder_count += 1;
    st->last_sort_step = now;
  }
  st->checked = 0;
  return reorder;
}

// Neighbour rebuild of one replica (NeighborSystem::rebuild_now restated for a
// workgroup): bin, counting sort, per-cell slot order, optional reorder, build.
template <int BLOCK, bool TOPO>
__device__ void rebuild_replica(const ReplicaParams& P, const ReplicaDesc& d, ReplicaState* st,
                                ReplicaShared<BLOCK>& sh, uint64_t now) {
  const ReplicaArrays& A = P.arrays;
  const bool topo = TOPO && (d.has_listed || d.has_excl);
  const uint32_t n = d.n, tid = threadIdx.x;
  real4* pos = A.pos + d.atom0;
  real4* vel = A.vel + d.atom0;
  int4* image = A.image + d.atom0;
  uint32_t* tag = A.tag + d.atom0;
  uint32_t* cell_of = A.cell_of + d.atom0;
  uint32_t* cell_atoms = A.cell_atoms + d.atom0;
  uint32_t* nbr_count = A.nbr_count + d.atom0;
  uint32_t* nbr = A.nbr + d.nbr0;
  real4* ref_pos = A.ref_pos + d.atom0;
  int4* ref_image = A.ref_image + d.atom0;
  const uint32_t nc = d.grid.n_cells;
  uint32_t* count = A.cell_count + d.cell0;
  uint32_t* start = A.cell_start + d.cell0;
  uint32_t* cursor = A.cell_cursor + d.cell0;

  const uint64_t builds = st->build_count + 1;
  const bool due = d.sort_every_steps != 0 ? now - st->last_sort_step >= d.sort_every_steps
                                           : builds % d.sort_every == 0;
  const bool reorder = d.grid.path != 0 && due;
  __syncthreads();
  if (tid == 0) {
    (void)begin_rebuild(d, st, now);
  }
  if (d.grid.path != 0) {
    for (uint32_t c = tid; c < nc; c += BLOCK) count[c] = 0;
    __syncthreads();
    for (uint32_t a = tid; a < n; a += BLOCK) {
      const uint32_t c = cell_of_position(pos[a], d.box, d.grid);
      cell_of[a] = c;
      atomicAdd(&count[c], 1u);
    }
    __syncthreads();
    block_scan<BLOCK>(count, start, cursor, nc, sh.scan);
    // Counts left at zero, as the phased mode's scan leaves them.
    for (uint32_t c = tid; c < nc; c += BLOCK) count[c] = 0;
    for (uint32_t a = tid; a < n; a += BLOCK) {
      cell_atoms[atomicAdd(&cursor[cell_of[a]], 1u)] = a;
    }
    __syncthreads();
    for (uint32_t c = tid; c < nc; c += BLOCK) {
      const uint32_t b = start[c], e = start[c + 1];
      for (uint32_t k = b + 1; k < e; ++k) {
        const uint32_t v = cell_atoms[k];
        uint32_t m = k;
        while (m > b && cell_atoms[m - 1] > v) {
          cell_atoms[m] = cell_atoms[m - 1];
          --m;
        }
        cell_atoms[m] = v;
      }
    }
    __syncthreads();
    if (reorder) {
      real4* s_pos = A.s_pos + d.atom0;
      real4* s_vel = A.s_vel + d.atom0;
      int4* s_img = A.s_image + d.atom0;
      uint32_t* s_tag = A.s_tag + d.atom0;
      uint32_t* s_cell = A.s_cell_of + d.atom0;
      for (uint32_t s = tid; s < n; s += BLOCK) {
        const uint32_t o = cell_atoms[s];
        s_pos[s] = pos[o];
        s_vel[s] = vel[o];
        s_img[s] = image[o];
        s_tag[s] = tag[o];
        s_cell[s] = cell_of[o];
        if (topo) topology_gather(P, d, s, o);
      }
      __syncthreads();
      for (uint32_t s = tid; s < n; s += BLOCK) {
        pos[s] = s_pos[s];
        vel[s] = s_vel[s];
        image[s] = s_img[s];
        tag[s] = s_tag[s];
        cell_of[s] = s_cell[s];
        cell_atoms[s] = s;
        if (topo) topology_commit(P, d, s);
      }
      __syncthreads();
    }
  }
  const uint32_t* topo_count = A.topo_count + d.atom0;
  const uint32_t* excl = A.excl + static_cast<size_t>(P.ecap) * d.atom0;
  uint32_t overflow = 0;
  for (uint32_t i = tid; i < n; i += BLOCK) {
    const uint32_t ne = TOPO && d.has_excl ? topo_count[i] >> 24 : 0u;
    overflow |= build_atom(d, pos, image, cell_of, cell_atoms, start, nbr, nbr_count, ref_pos,
                           ref_image, excl, ne, i);
  }
  if (overflow) atomicOr(&st->overflow, 1u);
  __syncthreads();
}

// Full-list LJ + listed terms + optional second half-kick; adds this thread's
// share of the kinetic energy, energies and virials.
template <int BLOCK, bool TOPO>
__device__ void force_replica(const ReplicaParams& P, const ReplicaDesc& d, bool kick,
                              bool single, const LjCoeff* lj, StepSums& sums) {
  const ReplicaArrays& A = P.arrays;
  const TopologyView topo = replica_topology(P, d);
  const uint32_t t = threadIdx.x & (kReplicaTeam - 1);
  for (uint32_t i = threadIdx.x / kReplicaTeam; i < d.n; i += BLOCK / kReplicaTeam) {
    force_atom<kReplicaTeam, kReplicaSlots, true, TOPO>(d, A.pos + d.atom0, A.vel + d.atom0, A.force + d.atom0,
                                            A.nbr_count + d.atom0, A.nbr + d.nbr0, lj, topo, kick,
                                            single, i, t, sums);
  }
}

__device__ __forceinline__ void store_observables(ReplicaState* st, const double* tot) {
  for (int q = 0; q < 4; ++q) {
    st->pe[q] = tot[1 + q];
    st->virial[q] = tot[5 + q];
  }
}

// TOPO: the batch has listed terms or exclusions (a pair-only batch runs the
// kernel without their code).
template <int BLOCK, bool TOPO>
__global__ __launch_bounds__(BLOCK) void replica_kernel(ReplicaParams P, uint64_t n_steps,
                                                        int observables_on_last) {
  __shared__ ReplicaShared<BLOCK> sh;
  const ReplicaDesc d = P.desc[blockIdx.x];
  const ReplicaArrays& A = P.arrays;
  ReplicaState* st = P.state + blockIdx.x;
  const uint32_t tid = threadIdx.x, n = d.n;
  const bool single = d.n_types == 1;
  for (int k = tid; k < d.n_types * d.n_types; k += BLOCK) sh.lj[k] = P.lj[d.lj0 + k];
  if (tid == 0) {
    sh.lambda = st->lambda;
  }
  __syncthreads();

  double tot[kReplicaSums];
  if (st->initialized == 0) {
    // Initial build and step-0 forces (simulation_run.cpp:113-119).
    rebuild_replica<BLOCK, TOPO>(P, d, st, sh, 0);
    StepSums sums;
    force_replica<BLOCK, TOPO>(P, d, false, single, sh.lj, sums);
    step_totals<BLOCK>(sums, true, tot, sh.red);
    if (tid == 0) {
      st->initialized = 1;
      st->ke = tot[0];
      store_observables(st, tot);
      sh.lambda = 1.0;
    }
    __syncthreads();
  }

  real4* pos = A.pos + d.atom0;
  real4* vel = A.vel + d.atom0;
  const real4* force = A.force + d.atom0;
  int4* image = A.image + d.atom0;
  const real4* ref_pos = A.ref_pos + d.atom0;
  const int4* ref_image = A.ref_image + d.atom0;
  uint64_t step = st->step;
  for (uint64_t s = 0; s < n_steps; ++s) {
    const uint64_t physical = step + 1;
    const bool every = d.cutoff_buffer == real_t(0);
    const bool check = !every && physical % d.check_every == 0;
    const real_t lam = static_cast<real_t>(sh.lambda);
    int exceeded = 0;
    for (uint32_t a = tid; a < n; a += BLOCK) {
      exceeded |= integrate_atom(d, pos, vel, force, image, ref_pos, ref_image, a, lam, check);
    }
    const bool any = __syncthreads_or(exceeded) != 0;
    bool rebuild = every;
    if (check) {
      const bool first = st->checked == 0;
      rebuild = any;
      __syncthreads();
      if (tid == 0) {
        st->checked = 1;
        if (any && d.check_every > 1 && first) st->dangerous += 1;
      }
    }
    if (rebuild) {
      rebuild_replica<BLOCK, TOPO>(P, d, st, sh, physical);
    }
    const bool obs = observables_on_last && s + 1 == n_steps;
    StepSums sums;
    force_replica<BLOCK, TOPO>(P, d, true, single, sh.lj, sums);
    step_totals<BLOCK>(sums, obs, tot, sh.red);
    const double ke = tot[0];
    if (tid == 0) {
      double lambda = 1.0;
      if (d.berendsen && ke > 0.0 && d.dof > 0.0) {
        const double current = 2.0 * ke / d.dof;
        const double arg = 1.0 + d.dt_over_tau * (d.target_thermal / current - 1.0);
        lambda = sqrt(arg > 0.0 ? arg : 0.0);
      }
      sh.lambda = lambda;
      st->ke = ke;
      if (obs) store_observables(st, tot);
    }
    __syncthreads();
    step = physical;
  }
  if (tid == 0) {
    st->step = step;
    st->lambda = sh.lambda;
  }
}

// ---------------------------------------------------------------- phased ---

constexpr int kPhasedBlock = 256;
constexpr uint32_t kPhasedWaves = kPhasedBlock / 64;  // chunks per workgroup
constexpr uint32_t kWorkBlocks = 1024;                // grid of the work-list kernels

__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Local index range of chunk c within its replica: atom a0 + lane.
struct ChunkAtoms {
  uint32_t r;   // replica
  uint32_t a0;  // first atom of the chunk, local to the replica
};
__device__ __forceinline__ ChunkAtoms chunk_atoms(const ReplicaParams& P, uint32_t c) {
  const uint32_t r = wave_uniform(P.chunk_replica[c]);
  return {r, (c - P.desc[r].chunk0) * kReplicaChunk};
}

// Step kernel 1: first half-kick, drift, wrap and (on check steps) the
// displacement test of every atom; a replica's test result is the OR over its
// chunks.
__global__ __launch_bounds__(kPhasedBlock) void phased_integrate_kernel(ReplicaParams P) {
  const uint32_t c = wave_uniform(blockIdx.x * kPhasedWaves + threadIdx.x / 64);
  if (c >= P.total_chunks) return;
  const ChunkAtoms ca = chunk_atoms(P, c);
  const ReplicaDesc& d = P.desc[ca.r];
  ReplicaState* st = P.state + ca.r;
  const ReplicaArrays& A = P.arrays;
  const uint64_t physical = st->step + 1;
  const bool every = d.cutoff_buffer == real_t(0);
  const bool check = !every && physical % d.check_every == 0;
  const real_t lam = static_cast<real_t>(st->lambda);
  const uint32_t a = ca.a0 + (threadIdx.x & 63u);
  int exceeded = 0;
  if (a < d.n) {
    exceeded = integrate_atom(d, A.pos + d.atom0, A.vel + d.atom0, A.force + d.atom0,
                              A.image + d.atom0, A.ref_pos + d.atom0, A.ref_image + d.atom0, a,
                              lam, check);
  }
  if (__builtin_amdgcn_ballot_w64(exceeded != 0) != 0 && (threadIdx.x & 63u) == 0) {
    atomicOr(&st->exceeded, 1u);
  }
}

// Step kernel 2, one wave per replica: the check and rebuild decisions of the
// persistent step loop (initial: the first build at step 0), and the
// replica's entries of the rebuild kernels' work lists.
__global__ __launch_bounds__(64) void phased_decide_kernel(ReplicaParams P, int initial) {
  __shared__ uint32_t base[2];
  const uint32_t r = blockIdx.x;
  const ReplicaDesc& d = P.desc[r];
  ReplicaState* st = P.state + r;
  if (threadIdx.x == 0) {
    bool rebuild = true, reorder = false;
    uint64_t now = 0;
    if (!initial) {
      now = st->step + 1;
      const bool every = d.cutoff_buffer == real_t(0);
      const bool check = !every && now % d.check_every == 0;
      rebuild = every;
      if (check) {
        const bool any = st->exceeded != 0;
        const bool first = st->checked == 0;
        rebuild = any;
        st->checked = 1;
        if (any && d.check_every > 1 && first) st->dangerous += 1;
      }
    }
    st->exceeded = 0;
    if (rebuild) reorder = begin_rebuild(d, st, now);
    st->rebuild = rebuild ? 1u : 0u;
    st->reorder = reorder ? 1u : 0u;
    base[0] = rebuild ? atomicAdd(&P.work_count[0], d.n_chunks) : 0xffffffffu;
    base[1] = reorder ? atomicAdd(&P.work_count[1], d.n_chunks) : 0xffffffffu;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < d.n_chunks; k += 64) {
    if (base[0] != 0xffffffffu) P.work[base[0] + k] = d.chunk0 + k;
    if (base[1] != 0xffffffffu) P.work_reorder[base[1] + k] = d.chunk0 + k;
  }
}

// Runs body(chunk, replica, local atom) for every atom of the chunks on a
// work list, one wave per chunk.
template <typename Body>
__device__ __forceinline__ void for_work(const ReplicaParams& P, const uint32_t* list,
                                         uint32_t entries, Body body) {
  for (uint32_t w = wave_uniform(blockIdx.x * kPhasedWaves + threadIdx.x / 64); w < entries;
       w += gridDim.x * kPhasedWaves) {
    const ChunkAtoms ca = chunk_atoms(P, wave_uniform(list[w]));
    body(ca.r, ca.a0 + (threadIdx.x & 63u));
  }
}

// Rebuild kernel: cell of every atom and per-cell counts (counts start at
// zero: the scan clears them after use).
__global__ __launch_bounds__(kPhasedBlock) void phased_bin_kernel(ReplicaParams P) {
  const ReplicaArrays& A = P.arrays;
  for_work(P, P.work, P.work_count[0], [&](uint32_t r, uint32_t a) {
    const ReplicaDesc& d = P.desc[r];
    if (d.grid.path == 0 || a >= d.n) return;
    const uint32_t c = cell_of_position(A.pos[d.atom0 + a], d.box, d.grid);
    A.cell_of[d.atom0 + a] = c;
    (void)wave_increment(A.cell_count + d.cell0, c);  // one atomic per distinct cell
  });
}

// Rebuild kernel, one workgroup per replica: exclusive scan of the counts into
// cell starts and scatter cursors; clears the counts.
__global__ __launch_bounds__(kPhasedBlock) void phased_scan_kernel(ReplicaParams P) {
  __shared__ uint32_t sums[kPhasedBlock];
  const ReplicaDesc& d = P.desc[blockIdx.x];
  if (P.state[blockIdx.x].rebuild == 0 || d.grid.path == 0) return;
  const ReplicaArrays& A = P.arrays;
  uint32_t* count = A.cell_count + d.cell0;
  block_scan<kPhasedBlock>(count, A.cell_start + d.cell0, A.cell_cursor + d.cell0,
                           d.grid.n_cells, sums);
  for (uint32_t c = threadIdx.x; c < d.grid.n_cells; c += kPhasedBlock) count[c] = 0;
}

// Rebuild kernel: counting-sort scatter of the atoms into their cells.
__global__ __launch_bounds__(kPhasedBlock) void phased_scatter_kernel(ReplicaParams P) {
  const ReplicaArrays& A = P.arrays;
  for_work(P, P.work, P.work_count[0], [&](uint32_t r, uint32_t a) {
    const ReplicaDesc& d = P.desc[r];
    if (d.grid.path == 0 || a >= d.n) return;
    const uint32_t c = A.cell_of[d.atom0 + a];
    A.cell_tmp[d.atom0 + wave_increment(A.cell_cursor + d.cell0, c)] = a;
  });
}

// Rebuild kernel: each atom's place in its cell's slot order (the number of
// members with a lower slot), giving the sorted member lists.
__global__ __launch_bounds__(kPhasedBlock) void phased_rank_kernel(ReplicaParams P) {
  const ReplicaArrays& A = P.arrays;
  for_work(P, P.work, P.work_count[0], [&](uint32_t r, uint32_t a) {
    const ReplicaDesc& d = P.desc[r];
    if (d.grid.path == 0 || a >= d.n) return;
    const uint32_t c = A.cell_of[d.atom0 + a];
    const uint32_t b = A.cell_start[d.cell0 + c], e = A.cell_start[d.cell0 + c + 1];
    const uint32_t* members = A.cell_tmp + d.atom0;
    uint32_t rank = 0;
    for (uint32_t k = b; k < e; ++k) rank += members[k] < a ? 1u : 0u;
    A.cell_atoms[d.atom0 + b + rank] = a;
    const real4 x = A.pos[d.atom0 + a];
    A.cell_pos[d.atom0 + b + rank] = make_real4(x.x, x.y, x.z, real_of_id(a));
  });
}

// Reorder kernels: gather the per-atom state into cell order, then commit it.
__global__ __launch_bounds__(kPhasedBlock) void phased_reorder_gather_kernel(ReplicaParams P) {
  const ReplicaArrays& A = P.arrays;
  for_work(P, P.work_reorder, P.work_count[1], [&](uint32_t r, uint32_t s) {
    const ReplicaDesc& d = P.desc[r];
    if (s >= d.n) return;
    const uint32_t o = d.atom0 + A.cell_atoms[d.atom0 + s];
    const uint32_t g = d.atom0 + s;
    A.s_pos[g] = A.pos[o];
    A.s_vel[g] = A.vel[o];
    A.s_image[g] = A.image[o];
    A.s_tag[g] = A.tag[o];
    A.s_cell_of[g] = A.cell_of[o];
    if (d.has_listed || d.has_excl) topology_gather(P, d, s, o - d.atom0);
  });
}

__global__ __launch_bounds__(kPhasedBlock) void phased_reorder_commit_kernel(ReplicaParams P) {
  const ReplicaArrays& A = P.arrays;
  for_work(P, P.work_reorder, P.work_count[1], [&](uint32_t r, uint32_t s) {
    const ReplicaDesc& d = P.desc[r];
    if (s >= d.n) return;
    const uint32_t g = d.atom0 + s;
    A.pos[g] = A.s_pos[g];
    A.vel[g] = A.s_vel[g];
    A.image[g] = A.s_image[g];
    A.tag[g] = A.s_tag[g];
    A.cell_of[g] = A.s_cell_of[g];
    A.cell_atoms[g] = s;
    const real4 x = A.s_pos[g];
    A.cell_pos[g] = make_real4(x.x, x.y, x.z, real_of_id(s));
    if (d.has_listed || d.has_excl) topology_commit(P, d, s);
  });
}

// Rebuild kernel, one chunk per workgroup: the neighbour lists, with a team
// of kPhasedBuildTeam adjacent lanes per atom. The candidates are build_atom's
// in its order (neighbour cells in neighbourhood order, a cell's members in
// slot order, read from the cell-ordered positions; every atom on the direct
// path): lane t tests candidates k + u*kPhasedBuildTeam + t of each round and
// the team appends the accepted ones in candidate order through a ballot, so
// the lists are build_atom's. A replica's overflow is the OR over its chunks.
constexpr int kPhasedBuildTeam = 4;
constexpr int kPhasedBuildSlots = 4;
static_assert(kPhasedBlock == kReplicaChunk * kPhasedBuildTeam, "one chunk per workgroup");

// EXCL: 0 no exclusions in the batch; 4 or 8: an atom's excluded partners
// (at most EXCL) held in registers, unused entries never a slot; -1: read
// from memory per candidate.
template <int EXCL>
__global__ __launch_bounds__(kPhasedBlock) void phased_build_kernel(ReplicaParams P) {
  const ReplicaArrays& A = P.arrays;
  const uint32_t t = threadIdx.x & (kPhasedBuildTeam - 1);
  const uint32_t team_lane0 = (threadIdx.x & 63u) & ~static_cast<uint32_t>(kPhasedBuildTeam - 1);
  const uint32_t below = (1u << t) - 1u;
  const uint32_t entries = P.work_count[0];
  for (uint32_t w = blockIdx.x; w < entries; w += gridDim.x) {
    const ChunkAtoms ca = chunk_atoms(P, wave_uniform(P.work[w]));
    const ReplicaDesc& d = P.desc[ca.r];
    const uint32_t n = d.n;
    const uint32_t i = ca.a0 + threadIdx.x / kPhasedBuildTeam;
    uint32_t count = 0;
    if (i < n) {  // uniform over a team
      const real4* pos = A.pos + d.atom0;
      const real4* cpos = A.cell_pos + d.atom0;
      uint32_t* nbr = A.nbr + d.nbr0;
      const uint32_t maxn = d.max_neighbors;
      const real_t L0 = d.box.L[0], L1 = d.box.L[1], L2 = d.box.L[2];
      const real_t h0 = d.box.hL[0], h1 = d.box.hL[1], h2 = d.box.hL[2];
      const real4 xi = pos[i];
      const uint32_t* excl = A.excl + static_cast<size_t>(P.ecap) * d.atom0;
      const uint32_t ne = EXCL != 0 && d.has_excl ? A.topo_count[d.atom0 + i] >> 24 : 0u;
      uint32_t ex[EXCL > 0 ? EXCL : 1];
#pragma unroll
      for (int k = 0; k < (EXCL > 0 ? EXCL : 1); ++k) {
        ex[k] = EXCL > 0 && static_cast<uint32_t>(k) < ne
                    ? excl[static_cast<size_t>(k) * n + i]
                    : 0xffffffffu;
      }
      const auto excluded = [&](uint32_t j) {
        bool hit = false;
        if (EXCL > 0) {
#pragma unroll
          for (int k = 0; k < (EXCL > 0 ? EXCL : 1); ++k) hit |= ex[k] == j;
        } else if (EXCL < 0) {
          hit = ne != 0 && excluded_partner(excl, n, i, ne, j);
        }
        return hit;
      };
      // Candidates [b, e) of a run (b < e); entry(k) = {x, y, z, slot bits}.
      auto scan = [&](uint32_t b, uint32_t e, auto entry) {
        for (uint32_t k = b; k < e; k += kPhasedBuildTeam * kPhasedBuildSlots) {
          real4 x[kPhasedBuildSlots];
#pragma unroll
          for (int u = 0; u < kPhasedBuildSlots; ++u) {
            const uint32_t q = k + u * kPhasedBuildTeam + t;
            x[u] = entry(q < e ? q : e - 1);
          }
#pragma unroll
          for (int u = 0; u < kPhasedBuildSlots; ++u) {
            const uint32_t j = id_of(x[u].w);
            const real_t dx = mic(xi.x - x[u].x, L0, h0);
            const real_t dy = mic(xi.y - x[u].y, L1, h1);
            const real_t dz = mic(xi.z - x[u].z, L2, h2);
            const uint64_t m =
                __builtin_amdgcn_ballot_w64(k + u * kPhasedBuildTeam + t < e) &
                __builtin_amdgcn_ballot_w64(!(r2_strict(dx, dy, dz) > d.rs2)) &
                __builtin_amdgcn_ballot_w64(j != i && !excluded(j));
            const uint32_t bits =
                static_cast<uint32_t>(m >> team_lane0) & ((1u << kPhasedBuildTeam) - 1u);
            if (__builtin_amdgcn_inverse_ballot_w64(m)) {
              const uint32_t slot = count + __builtin_popcount(bits & below);
              if (slot < maxn) nbr[replica_nbr_index(slot, i, n)] = j;
            }
            count += __builtin_popcount(bits);
          }
        }
      };
      if (d.grid.path == 0) {
        scan(0u, n, [pos](uint32_t k) {
          const real4 q = pos[k];
          return make_real4(q.x, q.y, q.z, real_of_id(k));
        });
      } else {
        const uint32_t* start = A.cell_start + d.cell0;
        const auto entry = [cpos](uint32_t k) { return cpos[k]; };
        const CellNeighborhood h = neighborhood(A.cell_of[d.atom0 + i], d.grid);
        for (int a = 0; a < h.nz; ++a) {
          for (int b = 0; b < h.ny; ++b) {
            const uint32_t row = d.grid.nx * (h.ys[b] + d.grid.ny * h.zs[a]);
            for (int c = 0; c < h.nx; ++c) {
              const uint32_t cell = h.xs[c] + row;
              const uint32_t b0 = start[cell], e0 = start[cell + 1];
              if (b0 < e0) scan(b0, e0, entry);
            }
          }
        }
      }
      if (t == 0) {
        A.nbr_count[d.atom0 + i] = min(count, maxn);
        A.ref_pos[d.atom0 + i] = xi;
        A.ref_image[d.atom0 + i] = A.image[d.atom0 + i];
      }
    }
    if (__builtin_amdgcn_ballot_w64(count > d.max_neighbors && t == 0) != 0 &&
        (threadIdx.x & 63u) == 0) {
      atomicOr(&P.state[ca.r].overflow, 1u);
    }
  }
}

// Step kernel: full-list LJ force, listed terms and the second half-kick (not
// on the initial step) of one chunk per workgroup, TEAM lanes per atom; the
// chunk's kinetic energy and, when OBS, energies and virials go to its
// partials.
template <int TEAM, int SLOTS, bool OBS, bool LISTED>
__global__ __launch_bounds__(kReplicaChunk * TEAM) void phased_force_kernel(ReplicaParams P,
                                                                            int kick) {
  constexpr int BLOCK = kReplicaChunk * TEAM;
  __shared__ LjCoeff lj[kMaxLdsTypes * kMaxLdsTypes];
  __shared__ double red[BLOCK / 64];
  const uint32_t c = blockIdx.x;
  const uint32_t r = wave_uniform(P.chunk_replica[c]);
  const ReplicaDesc& d = P.desc[r];
  for (int k = threadIdx.x; k < d.n_types * d.n_types; k += BLOCK) lj[k] = P.lj[d.lj0 + k];
  __syncthreads();
  const ReplicaArrays& A = P.arrays;
  const uint32_t i = (c - d.chunk0) * kReplicaChunk + threadIdx.x / TEAM;
  const uint32_t t = threadIdx.x & (TEAM - 1);
  StepSums sums;
  if (i < d.n) {
    force_atom<TEAM, SLOTS, OBS, LISTED>(d, A.pos + d.atom0, A.vel + d.atom0, A.force + d.atom0,
                                 A.nbr_count + d.atom0, A.nbr + d.nbr0, lj,
                                 replica_topology(P, d), kick != 0, d.n_types == 1, i, t, sums);
  }
  double tot[kReplicaSums];
  step_totals<BLOCK>(sums, OBS, tot, red);
  if (threadIdx.x == 0) {
    const int nq = OBS ? kReplicaSums : 1;
    for (int q = 0; q < nq; ++q) P.partials[static_cast<size_t>(q) * P.total_chunks + c] = tot[q];
  }
}

// Step kernel, one workgroup per replica: totals of the chunk partials in
// chunk order, the Berendsen factor for the next step, the step counter
// (initial: the step-0 observables). Workgroup 0 also empties the work lists.
__global__ __launch_bounds__(kPhasedBlock) void phased_reduce_kernel(ReplicaParams P, int initial,
                                                                     int observables) {
  __shared__ double red[kPhasedBlock / 64];
  const uint32_t r = blockIdx.x;
  const ReplicaDesc& d = P.desc[r];
  ReplicaState* st = P.state + r;
  const int nq = observables || initial ? kReplicaSums : 1;
  double tot[kReplicaSums] = {};
  for (int q = 0; q < nq; ++q) {
    double v = 0.0;
    for (uint32_t k = threadIdx.x; k < d.n_chunks; k += kPhasedBlock) {
      v += P.partials[static_cast<size_t>(q) * P.total_chunks + d.chunk0 + k];
    }
    tot[q] = block_sum<kPhasedBlock>(v, red);
  }
  if (threadIdx.x == 0) {
    const double ke = tot[0];
    double lambda = 1.0;
    if (!initial && d.berendsen && ke > 0.0 && d.dof > 0.0) {
      const double current = 2.0 * ke / d.dof;
      const double arg = 1.0 + d.dt_over_tau * (d.target_thermal / current - 1.0);
      lambda = sqrt(arg > 0.0 ? arg : 0.0);
    }
    st->lambda = lambda;
    st->ke = ke;
    if (observables || initial) store_observables(st, tot);
    if (initial) {
      st->initialized = 1;
    } else {
      st->step += 1;
    }
    if (r == 0) {
      P.work_count[0] = 0;
      P.work_count[1] = 0;
    }
  }
}

template <bool TOPO>
void launch_persistent(const ReplicaParams& p, uint32_t count, int block, uint64_t n_steps,
                       int obs, hipStream_t stream) {
  switch (block) {
    case 256:
      replica_kernel<256, TOPO><<<count, 256, 0, stream>>>(p, n_steps, obs);
      break;
    case 512:
      replica_kernel<512, TOPO><<<count, 512, 0, stream>>>(p, n_steps, obs);
      break;
    default:
      replica_kernel<1024, TOPO><<<count, 1024, 0, stream>>>(p, n_steps, obs);
      break;
  }
}

template <int TEAM, int SLOTS>
void launch_phased_force(const ReplicaParams& p, bool obs, bool listed, int kick,
                         hipStream_t stream) {
  const dim3 grid(p.total_chunks), block(kReplicaChunk * TEAM);
  if (listed) {
    if (obs) {
      phased_force_kernel<TEAM, SLOTS, true, true><<<grid, block, 0, stream>>>(p, kick);
    } else {
      phased_force_kernel<TEAM, SLOTS, false, true><<<grid, block, 0, stream>>>(p, kick);
    }
  } else if (obs) {
    phased_force_kernel<TEAM, SLOTS, true, false><<<grid, block, 0, stream>>>(p, kick);
  } else {
    phased_force_kernel<TEAM, SLOTS, false, false><<<grid, block, 0, stream>>>(p, kick);
  }
}

// Whether any replica of the batch has listed terms (records of any kind).
bool batch_has_listed(const ReplicaParams& p) { return p.bcap + p.acap + p.dcap > 0; }

}  // namespace

void launch_replicas(const ReplicaParams& p, uint32_t count, int block, uint64_t n_steps,
                     bool observables_on_last, hipStream_t stream) {
  const int obs = observables_on_last ? 1 : 0;
  if (batch_has_listed(p) || p.ecap > 0) {
    launch_persistent<true>(p, count, block, n_steps, obs, stream);
  } else {
    launch_persistent<false>(p, count, block, n_steps, obs, stream);
  }
}

bool phased_shape_supported(PhasedShape shape) {
  const int code = shape.team * 100 + shape.slots;
  return code == 204 || code == 208 || code == 404 || code == 408 || code == 802 || code == 804;
}

void launch_phased_step(const ReplicaParams& p, PhasedShape shape, bool initial, bool observables,
                        hipStream_t stream) {
  const unsigned per_atom = (p.total_chunks + kPhasedWaves - 1) / kPhasedWaves;
  if (!initial) phased_integrate_kernel<<<per_atom, kPhasedBlock, 0, stream>>>(p);
  phased_decide_kernel<<<p.count, 64, 0, stream>>>(p, initial ? 1 : 0);
  // Work-list kernels loop over the listed chunks; their grid is fixed at
  // capture time, so a step without rebuilds costs each one empty launch.
  const unsigned work = std::min<unsigned>(kWorkBlocks, per_atom);
  phased_bin_kernel<<<work, kPhasedBlock, 0, stream>>>(p);
  phased_scan_kernel<<<p.count, kPhasedBlock, 0, stream>>>(p);
  phased_scatter_kernel<<<work, kPhasedBlock, 0, stream>>>(p);
  phased_rank_kernel<<<work, kPhasedBlock, 0, stream>>>(p);
  phased_reorder_gather_kernel<<<work, kPhasedBlock, 0, stream>>>(p);
  phased_reorder_commit_kernel<<<work, kPhasedBlock, 0, stream>>>(p);
  if (p.ecap == 0) {
    phased_build_kernel<0><<<work, kPhasedBlock, 0, stream>>>(p);
  } else if (p.ecap <= 4) {
    phased_build_kernel<4><<<work, kPhasedBlock, 0, stream>>>(p);
  } else if (p.ecap <= 8) {
    phased_build_kernel<8><<<work, kPhasedBlock, 0, stream>>>(p);
  } else {
    phased_build_kernel<-1><<<work, kPhasedBlock, 0, stream>>>(p);
  }
  const int kick = initial ? 0 : 1;
  const bool obs = observables || initial;
  const bool listed = batch_has_listed(p);
  switch (shape.team * 100 + shape.slots) {
    case 204:
      launch_phased_force<2, 4>(p, obs, listed, kick, stream);
      break;
    case 208:
      launch_phased_force<2, 8>(p, obs, listed, kick, stream);
      break;
    case 408:
      launch_phased_force<4, 8>(p, obs, listed, kick, stream);
      break;
    case 802:
      launch_phased_force<8, 2>(p, obs, listed, kick, stream);
      break;
    case 804:
      launch_phased_force<8, 4>(p, obs, listed, kick, stream);
      break;
    default:
      launch_phased_force<4, 4>(p, obs, listed, kick, stream);
      break;
  }
  phased_reduce_kernel<<<p.count, kPhasedBlock, 0, stream>>>(p, initial ? 1 : 0,
                                                             observables ? 1 : 0);
}

}  // namespace beads_amd
