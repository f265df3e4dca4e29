// Batched replicas: many independent LJ systems advanced by ONE persistent
// launch. Each workgroup owns one replica and runs its whole step loop --
// integrate + displacement check, the cell-list rebuild with stable sort and
// reorder, the full-list LJ force with the second half-kick and the Berendsen
// factor -- with workgroup barriers between phases and no grid-wide sync.
// Replicas are independent, so the launch needs no inter-workgroup
// communication; uneven sizes are balanced by dispatching the largest replicas
// first (the host orders the descriptors) and by the autotuned workgroup size.
//
// Per replica the arithmetic is the single-system path's (same decisions, same
// FP64 pair values and reductions): velocity_verlet.cu:59-111,
// berendsen_thermostat.cu:104-163, cell_list_build.cu:110-219,
// full_neighbor_path_common.cuh:49-94, regular/small_dimension/direct builders,
// pair_lj.cu:21-167, neighbor_system.cpp:79-146,198-216 (paths under the
// reference's native/src).
#include "replicas.hpp"

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

// Neighbour rebuild of one replica (NeighborSystem::rebuild_now restated for a
// workgroup): bin, counting sort, per-cell slot order, optional reorder, build.
template <int BLOCK>
__device__ void rebuild_replica(const ReplicaDesc& d, const ReplicaArrays& A, ReplicaState* st,
                                ReplicaShared<BLOCK>& sh, uint64_t now) {
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

  uint64_t builds = st->build_count + 1;
  const bool due = d.sort_every_steps != 0 ? now - st->last_sort_step >= d.sort_every_steps
                                           : builds % d.sort_every == 0;
  const bool reorder = d.grid.path != 0 && due;
  __syncthreads();
  if (tid == 0) {
    st->build_count = builds;
    if (reorder) {
      st->reorder_count += 1;
      st->last_sort_step = now;
    }
    st->checked = 0;
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
      }
      __syncthreads();
      for (uint32_t s = tid; s < n; s += BLOCK) {
        pos[s] = s_pos[s];
        vel[s] = s_vel[s];
        image[s] = s_img[s];
        tag[s] = s_tag[s];
        cell_of[s] = s_cell[s];
        cell_atoms[s] = s;
      }
      __syncthreads();
    }
  }
  const real_t L0 = d.box.L[0], L1 = d.box.L[1], L2 = d.box.L[2];
  const real_t h0 = d.box.hL[0], h1 = d.box.hL[1], h2 = d.box.hL[2];
  uint32_t overflow = 0;
  for (uint32_t i = tid; i < n; i += BLOCK) {
    const real4 xi = pos[i];
    uint32_t cnt = 0;
    auto visit = [&](uint32_t j) {
      const real4 xj = pos[j];
      const real_t dx = mic(xi.x - xj.x, L0, h0);
      const real_t dy = mic(xi.y - xj.y, L1, h1);
      const real_t dz = mic(xi.z - xj.z, L2, h2);
      if (r2_strict(dx, dy, dz) > d.rs2 || j == i) return;
      if (cnt < d.max_neighbors) {
        nbr[static_cast<size_t>(cnt) * n + i] = j;
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
  }
  if (overflow) atomicOr(&st->overflow, 1u);
  __syncthreads();
}

// Full-list LJ + optional second half-kick; returns this thread's share of the
// kinetic energy, pair energy and virial.
template <int BLOCK>
__device__ void force_replica(const ReplicaDesc& d, const ReplicaArrays& A, bool kick, bool single,
                              const LjCoeff* lj, double& ke, double& pe, double& vir) {
  const uint32_t n = d.n;
  const real4* pos = A.pos + d.atom0;
  real4* vel = A.vel + d.atom0;
  real4* force = A.force + d.atom0;
  const uint32_t* nbr_count = A.nbr_count + d.atom0;
  const uint32_t* nbr = A.nbr + d.nbr0;
  const real_t L0 = d.box.L[0], L1 = d.box.L[1], L2 = d.box.L[2];
  const real_t h0 = d.box.hL[0], h1 = d.box.hL[1], h2 = d.box.hL[2];
  const int T = d.n_types;
  const LjCoeff c0 = lj[0];
  // A team of kReplicaTeam adjacent lanes per atom: lane t takes list slots
  // base + u*kReplicaTeam + t, kReplicaSlots of them in flight per round; the
  // team's FP64 partial forces meet in a butterfly and lane 0 stores.
  const uint32_t t = threadIdx.x & (kReplicaTeam - 1);
  for (uint32_t i = threadIdx.x / kReplicaTeam; i < n; i += BLOCK / kReplicaTeam) {
    const real4 xi = pos[i];
    const int ti = static_cast<int>(id_of(xi.w));
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0, w = 0.0;
    const uint32_t cnt = nbr_count[i];
    for (uint32_t base = 0; base < cnt; base += kReplicaTeam * kReplicaSlots) {
      uint32_t jj[kReplicaSlots];
#pragma unroll
      for (int u = 0; u < kReplicaSlots; ++u) {
        const uint32_t s = base + u * kReplicaTeam + t;
        jj[u] = s < cnt ? nbr[static_cast<size_t>(s) * n + i] : i;
      }
      real4 xj[kReplicaSlots];
#pragma unroll
      for (int u = 0; u < kReplicaSlots; ++u) {
        xj[u] = pos[jj[u]];
      }
#pragma unroll
      for (int u = 0; u < kReplicaSlots; ++u) {
        const real_t dx = mic_diff(xi.x, xj[u].x, L0, h0);
        const real_t dy = mic_diff(xi.y, xj[u].y, L1, h1);
        const real_t dz = mic_diff(xi.z, xj[u].z, L2, h2);
        if (base + u * kReplicaTeam + t < cnt && r2_strict(dx, dy, dz) < d.rc2) {
          const LjCoeff c = single ? c0 : lj[ti * T + static_cast<int>(id_of(xj[u].w))];
          double rr, uu;
          const double s = lj_pair(dx, dy, dz, c, rr, uu);
          fx += s * dx;
          fy += s * dy;
          fz += s * dz;
          e += uu;
          w += s * rr;
        }
      }
    }
    fx = team_sum<kReplicaTeam>(fx);
    fy = team_sum<kReplicaTeam>(fy);
    fz = team_sum<kReplicaTeam>(fz);
    pe += 0.5 * e;
    vir += 0.5 * w;
    if (t != 0) {
      continue;
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
    ke += 0.5 * static_cast<double>(v.w) * (vx * vx + vy * vy + vz * vz);
  }
}

template <int BLOCK>
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

  double ke = 0.0, pe = 0.0, vir = 0.0;
  if (st->initialized == 0) {
    // Initial build and step-0 forces (simulation_run.cpp:113-119).
    rebuild_replica<BLOCK>(d, A, st, sh, 0);
    force_replica<BLOCK>(d, A, false, single, sh.lj, ke, pe, vir);
    ke = block_sum<BLOCK>(ke, sh.red);
    pe = block_sum<BLOCK>(pe, sh.red);
    vir = block_sum<BLOCK>(vir, sh.red);
    if (tid == 0) {
      st->initialized = 1;
      st->ke = ke;
      st->pe = pe;
      st->virial = vir;
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
      if (check) {
#pragma clang fp contract(off)
        const real4 r = ref_pos[a];
        const int4 ri = ref_image[a];
        const real_t dx = (x.x - r.x) + static_cast<real_t>(im.x - ri.x) * d.box.L[0];
        const real_t dy = (x.y - r.y) + static_cast<real_t>(im.y - ri.y) * d.box.L[1];
        const real_t dz = (x.z - r.z) + static_cast<real_t>(im.z - ri.z) * d.box.L[2];
        if (dx * dx + dy * dy + dz * dz > d.thr2) exceeded = 1;
      }
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
      rebuild_replica<BLOCK>(d, A, st, sh, physical);
    }
    const bool obs = observables_on_last && s + 1 == n_steps;
    ke = pe = vir = 0.0;
    force_replica<BLOCK>(d, A, true, single, sh.lj, ke, pe, vir);
    ke = block_sum<BLOCK>(ke, sh.red);
    if (obs) {
      pe = block_sum<BLOCK>(pe, sh.red);
      vir = block_sum<BLOCK>(vir, sh.red);
    }
    if (tid == 0) {
      double lambda = 1.0;
      if (d.berendsen && ke > 0.0 && d.dof > 0.0) {
        const double current = 2.0 * ke / d.dof;
        const double arg = 1.0 + d.dt_over_tau * (d.target_thermal / current - 1.0);
        lambda = sqrt(arg > 0.0 ? arg : 0.0);
      }
      sh.lambda = lambda;
      st->ke = ke;
      if (obs) {
        st->pe = pe;
        st->virial = vir;
      }
    }
    __syncthreads();
    step = physical;
  }
  if (tid == 0) {
    st->step = step;
    st->lambda = sh.lambda;
  }
}

}  // namespace

void launch_replicas(const ReplicaParams& p, uint32_t count, int block, uint64_t n_steps,
                     bool observables_on_last, hipStream_t stream) {
  const int obs = observables_on_last ? 1 : 0;
  switch (block) {
    case 256:
      replica_kernel<256><<<count, 256, 0, stream>>>(p, n_steps, obs);
      break;
    case 512:
      replica_kernel<512><<<count, 512, 0, stream>>>(p, n_steps, obs);
      break;
    default:
      replica_kernel<1024><<<count, 1024, 0, stream>>>(p, n_steps, obs);
      break;
  }
}

}  // namespace beads_amd
