// gfx950 kernels of the per-step MD path.
//
// Design (CDNA4): 256-wide workgroups (4 wave64s), one work-item per atom
// except in the force pass (a team of kTeam lanes per atom), real4 SoA
// streams (float4, or double4 in the BEADS_REAL64 build), FP64 for LJ pair
// values and global reductions. Kernels that gather
// neighbour positions map workgroups onto atoms XCD by XCD so each XCD's L2
// holds one contiguous (spatially sorted) slab of positions. Every kernel of
// the neighbour-rebuild sequence reads the device control block first and
// exits when no rebuild is due, so a step never waits on the host. No kernel
// fences at device scope (on gfx950 that is an L2 write-back and invalidate
// of the XCD): cross-workgroup totals are formed by a one-workgroup kernel
// launched after the workgroups that wrote the partials.
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

#include <type_traits>

#include "beads_amd.h"
#include "device_common.hpp"

namespace beads_amd {
namespace {

using namespace dev;

// ---------------------------------------------------------------- rebuild --

// Rebuild bookkeeping of NeighborSystem::update_after_position_change /
// rebuild_now (neighbor_system.cpp:79-146, 198-207), on the device. The sort
// cadence counts builds (neighbor_system.cpp:198-202) or, with
// sort_every_steps, steps: a build sorts when at least that many steps have
// passed since the last sort (the sort needs the rebuild that follows it).
__global__ void control_kernel(Ctrl* ctrl, NeighborControl nc, int forced, uint64_t steps) {
  if (threadIdx.x != 0 || blockIdx.x != 0) {
    return;
  }
  ctrl->clock += steps;
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
    const bool due = nc.sort_every_steps != 0
                         ? ctrl->clock - ctrl->last_sort >= nc.sort_every_steps
                         : (ctrl->build_count % nc.sort_every) == 0;
    const bool reorder = nc.path != 0 && due;
    ctrl->reorder = reorder ? 1u : 0u;
    ctrl->reorder_count += reorder ? 1u : 0u;
    if (reorder) ctrl->last_sort = ctrl->clock;
    ctrl->checked = 0;
    ctrl->neighbor_total = 0;
  }
}

// cell_list_build.cu:135-166: binary32 cell index and atomic counts.
__global__ __launch_bounds__(kBlock) void bin_kernel(BinParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) {
    return;
  }
  const uint32_t c = cell_of_position(p.pos[i], p.box, p.grid);
  p.cell_of[i] = c;
  wave_increment(p.cell_count, c);  // one atomic per distinct cell of the wave
}

// One workgroup: exclusive scan of cell counts into cell_start[0..n_cells],
// copy to the scatter cursor and clear the counts for the next rebuild. The
// counts pass through LDS in tiles of kScanTile cells: coalesced loads (all
// of a thread's loads in flight), a scan of kScanPer consecutive cells per
// thread, a shuffle scan of the thread sums within each wave and the wave
// totals added from LDS (four barriers per tile), coalesced stores. Rows
// are padded by one word per kScanPer so a thread's row starts in its own
// bank.
constexpr int kScanBlock = 1024;
constexpr uint32_t kScanPer = 8;
constexpr uint32_t kScanTile = kScanBlock * kScanPer;

__device__ __forceinline__ uint32_t scan_slot(uint32_t k) { return k + k / kScanPer; }

__global__ __launch_bounds__(kScanBlock) void cell_scan_kernel(CellScanParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  constexpr uint32_t kWaves = kScanBlock / 64;
  __shared__ uint32_t tile[kScanTile + kScanBlock];
  __shared__ uint32_t wave_total[kWaves];
  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & 63u, wave = t >> 6;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < p.n_cells; base += kScanTile) {
    const uint32_t len = min(kScanTile, p.n_cells - base);
#pragma unroll
    for (uint32_t m = 0; m < kScanPer; ++m) {
      const uint32_t k = m * kScanBlock + t;
      tile[scan_slot(k)] = k < len ? p.cell_count[base + k] : 0u;
    }
    __syncthreads();
    uint32_t local = 0;
#pragma unroll
    for (uint32_t m = 0; m < kScanPer; ++m) {
      local += tile[scan_slot(t * kScanPer + m)];
    }
    uint32_t incl = local;  // inclusive scan of the thread sums within the wave
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, 64);
      incl += lane >= off ? v : 0u;
    }
    if (lane == 63) wave_total[wave] = incl;
    __syncthreads();
    uint32_t before = 0, tile_total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
      const uint32_t v = wave_total[w];
      before += w < wave ? v : 0u;
      tile_total += v;
    }
    uint32_t run = carry + before + incl - local;
#pragma unroll
    for (uint32_t m = 0; m < kScanPer; ++m) {
      const uint32_t s = scan_slot(t * kScanPer + m);
      const uint32_t cnt = tile[s];
      tile[s] = run;
      run += cnt;
    }
    carry += tile_total;
    __syncthreads();
#pragma unroll
    for (uint32_t m = 0; m < kScanPer; ++m) {
      const uint32_t k = m * kScanBlock + t;
      if (k < len) {
        const uint32_t start = tile[scan_slot(k)];
        p.cell_start[base + k] = start;
        p.cell_cursor[base + k] = start;
        p.cell_count[base + k] = 0;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    p.cell_start[p.n_cells] = carry;
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
  // One atomic per distinct cell of the wave; cell_sort orders the members.
  const uint32_t slot = wave_increment(p.cell_cursor, p.cell_of[i]);
  p.cell_atoms[slot] = i;
}

// Order each cell's members by (x, slot): with the counting sort this is a
// deterministic order, where the reference's atomic scatter
// (cell_list_build.cu:168-181) leaves each cell's members in arbitrary order;
// the sets a build finds are the same for any order inside a cell. Ordering
// by x makes the members of an x-row of cells ascending in x, so a build can
// scan a row as one run and skip its ends beyond r_s (cell_xbins). One wave
// per cell: for up to 64 members lane l holds member l, whose rank is the
// number of members before it in (x, slot) order, and stores it at start +
// rank; the x-bin prefix counts come from one ballot per bin. A larger cell
// is insertion-sorted by lane 0. When no reorder follows, the kernel also
// writes the members' positions in cell order with the slot in .w (the
// build's candidate stream); after a reorder, reorder_commit writes them.
constexpr uint32_t kSortCellsPerBlock = kBlock / 64;

__device__ __forceinline__ real4 cell_entry(real4 x, uint32_t slot) {
  return make_real4(x.x, x.y, x.z, real_of_id(slot));
}

// v of lane `lane` (uniform over the wave).
__device__ __forceinline__ real_t read_lane(real_t v, int lane) {
#ifdef BEADS_REAL64
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b)), lane));
  const uint32_t hi = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b >> 32)), lane));
  return __builtin_bit_cast(real_t, (static_cast<uint64_t>(hi) << 32) | lo);
#else
  return __builtin_bit_cast(real_t, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
#endif
}

// Bin of coordinate x inside x-cell cx (0 .. kXBins - 1), monotone in x: the
// cell coordinate of axis_cell, offset by the cell and scaled by the bin count
// (a power of two).
__device__ __forceinline__ int x_bin(real_t x, real_t lo, real_t inv, uint32_t cx) {
#pragma clang fp contract(off)
  const real_t f = ((x - lo) * inv - static_cast<real_t>(cx)) * static_cast<real_t>(kXBins);
  return f <= real_t(0) ? 0
                        : (f >= static_cast<real_t>(kXBins - 1) ? kXBins - 1 : static_cast<int>(f));
}

__device__ __forceinline__ bool before_in_cell(real_t xa, uint32_t sa, real_t xb, uint32_t sb) {
  return xa < xb || (xa == xb && sa < sb);
}

__global__ __launch_bounds__(kBlock) void cell_sort_kernel(CellScanParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  const uint32_t lane = threadIdx.x & 63u;
  const bool stream = p.ctrl->reorder == 0;
  const real_t lo = p.box.lo[0], inv = p.grid.inv_cell[0];
  // A capped grid loops over the cells, so the launch stays small when the
  // rebuild flag makes it exit at once.
  for (uint32_t c = blockIdx.x * kSortCellsPerBlock + threadIdx.x / 64; c < p.n_cells;
       c += gridDim.x * kSortCellsPerBlock) {  // uniform over the wave
    const uint32_t b = p.cell_start[c], e = p.cell_start[c + 1];
    const uint32_t cnt = e - b;
    const uint32_t cx = c % p.grid.nx;
    uint32_t* a = p.cell_atoms;
    if (cnt <= 64) {
      const uint32_t v = lane < cnt ? a[b + lane] : 0xffffffffu;
      const real4 xv4 = lane < cnt ? p.pos[v] : make_real4(0, 0, 0, 0);
      const real_t xv = xv4.x;
      uint32_t rank = 0;
      for (uint32_t m = 0; m < cnt; ++m) {
        const uint32_t w = static_cast<uint32_t>(
            __builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(m)));
        const real_t xw = read_lane(xv, static_cast<int>(m));
        rank += before_in_cell(xw, w, xv, v) ? 1u : 0u;
      }
      if (lane < cnt) {
        a[b + rank] = v;
        if (stream) p.cell_pos[b + rank] = cell_entry(xv4, v);
      }
      const int bin = lane < cnt ? x_bin(xv, lo, inv, cx) : kXBins;
      uint32_t pre[kXBins];
#pragma unroll
      for (int m = 1; m <= kXBins; ++m) {
        pre[m - 1] = static_cast<uint32_t>(__builtin_popcountll(__ballot(bin < m)));
      }
      if (lane == 0) {
        p.cell_xbins[c] = make_uint4(pre[0] | (pre[1] << 16), pre[2] | (pre[3] << 16),
                                     pre[4] | (pre[5] << 16), pre[6] | (pre[7] << 16));
      }
    } else if (lane == 0) {
      for (uint32_t k = b + 1; k < e; ++k) {
        const uint32_t v = a[k];
        const real_t xv = p.pos[v].x;
        uint32_t m = k;
        while (m > b && before_in_cell(xv, v, p.pos[a[m - 1]].x, a[m - 1])) {
          a[m] = a[m - 1];
          --m;
        }
        a[m] = v;
      }
      uint32_t pre[kXBins] = {};
      for (uint32_t k = b; k < e; ++k) {
        const uint32_t v = a[k];
        const real4 xv4 = p.pos[v];
        if (stream) p.cell_pos[k] = cell_entry(xv4, v);
        const int bin = x_bin(xv4.x, lo, inv, cx);
        for (int m = bin + 1; m <= kXBins; ++m) ++pre[m - 1];
      }
      if (cnt <= kXBinsMaxCount) {
        p.cell_xbins[c] = make_uint4(pre[0] | (pre[1] << 16), pre[2] | (pre[3] << 16),
                                     pre[4] | (pre[5] << 16), pre[6] | (pre[7] << 16));
      }
    }
  }
}

// trajectory_output.cu:183-250: one record per atom in slot order, fields in
// the file's order (tag, type, position, image, velocity, force); the type is
// stored 1-based as in the spec.
__global__ __launch_bounds__(kBlock) void pack_records_kernel(PackParams p) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= p.n) {
    return;
  }
  uint32_t* r = p.out + static_cast<size_t>(i) * p.words;
  const real4 x = p.pos[i];
  uint32_t k = 0;
  // A real as one word (binary32) or two, low word first (binary64).
  auto put = [&](real_t v) {
#ifdef BEADS_REAL64
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    r[k++] = static_cast<uint32_t>(b);
    r[k++] = static_cast<uint32_t>(b >> 32);
#else
    r[k++] = __builtin_bit_cast(uint32_t, v);
#endif
  };
  r[k++] = p.tag[i];
  if (p.mask & BEADS_TRAJ_TYPE) r[k++] = id_of(x.w) + 1u;
  if (p.mask & BEADS_TRAJ_POSITION) {
    put(x.x);
    put(x.y);
    put(x.z);
  }
  if (p.mask & BEADS_TRAJ_IMAGE) {
    const int4 m = p.image[i];
    r[k++] = static_cast<uint32_t>(m.x);
    r[k++] = static_cast<uint32_t>(m.y);
    r[k++] = static_cast<uint32_t>(m.z);
  }
  if (p.mask & BEADS_TRAJ_VELOCITY) {
    const real4 v = p.vel[i];
    put(v.x);
    put(v.y);
    put(v.z);
  }
  if (p.mask & BEADS_TRAJ_FORCE) {
    const real4 f = p.force[i];
    put(f.x);
    put(f.y);
    put(f.z);
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
  if (p.spill_index != nullptr) p.s_spill_index[s] = p.spill_index[o];
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
  p.cell_pos[s] = cell_entry(p.s_pos[s], s);
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
  if (p.spill_index != nullptr) {
    // The particle's spill segments stay where they are; their partner slots
    // are remapped here by their one owner.
    const uint32_t si = p.s_spill_index[s];
    p.spill_index[s] = si;
    if (si != kNoSpill) {
      const SpillEntry e = p.spill[si];
      for (uint32_t k = 0; k < e.cnt[0]; ++k) {
        uint2& r = p.spill_bond[e.off[0] + k];
        r.x = m[r.x];
      }
      for (uint32_t k = 0; k < e.cnt[1]; ++k) {
        uint4& r = p.spill_angle[e.off[1] + k];
        r.x = m[r.x];
        r.y = m[r.y];
      }
      for (uint32_t k = 0; k < e.cnt[2]; ++k) {
        uint4& r = p.spill_dih[e.off[2] + k];
        r.x = m[r.x];
        r.y = m[r.y];
        r.z = m[r.z];
      }
      for (uint32_t k = 0; k < e.cnt[3]; ++k) {
        uint32_t& r = p.spill_excl[e.off[3] + k];
        r = m[r];
      }
    }
  }
}

constexpr int kBuildTeam = 4;
constexpr int kBuildSlots = 4;
constexpr uint32_t kBuildAtoms = kBlock / kBuildTeam;

// How the build forms a candidate's minimum image (build_kernel's accept).
constexpr int kShift = 0, kShiftYZ = 1, kGeneric = 2;
using kShiftMode = std::integral_constant<int, kShift>;
using kShiftYZMode = std::integral_constant<int, kShiftYZ>;
using kGenericMode = std::integral_constant<int, kGeneric>;

// Rows of three x-adjacent cells scanned per atom on the rows path (3 along
// z times 3 along y).
constexpr int kRows = 9;
// A row descriptor's end index (below n < 2^28) carries the row's shift codes
// in its top bits.
constexpr int kRowCodeShift = 28;
constexpr uint32_t kRowEndMask = (1u << kRowCodeShift) - 1u;
// LDS bytes a build workgroup may use to stage the union of its atoms' row
// runs (more LDS per workgroup costs occupancy).
constexpr uint32_t kStageBytesMax = 24576;

// Neighbour k (0, 1, 2) along an axis of count cells, in the order of
// unique_axis, and its image shift code (0: none, 1: +L, 2: -L): with at
// least 5 cells per axis a candidate in a wrapped neighbour cell lies more
// than L/2 away (mic adds +L or -L), and in any other neighbour cell less
// than L/2 away (mic adds 0).
__device__ inline uint32_t neighbour_cell(uint32_t cc, uint32_t count, int k, uint32_t& code) {
  code = 0;
  if (count <= 1) return 0u;
  if (count == 2) return k == 0 ? (cc == 0 ? 1u : 0u) : cc;
  if (k == 0) {
    if (cc == 0) {
      code = 1;
      return count - 1;
    }
    return cc - 1;
  }
  if (k == 2) {
    if (cc + 1 == count) {
      code = 2;
      return 0u;
    }
    return cc + 1;
  }
  return cc;
}

__device__ inline real_t shift_of(uint32_t code, real_t L) {
  return code == 0 ? static_cast<real_t>(0) : (code == 1 ? L : -L);
}

// Members of cell cc with x-bin below m (cell_xbins; m in 0..kXBins).
__device__ inline uint32_t members_below(const uint4* xbins, uint32_t cc, int m) {
  if (m <= 0) return 0u;
  const uint4 q = xbins[cc];
  const int f = m - 1;
  const uint32_t word = (f >> 1) == 0 ? q.x : ((f >> 1) == 1 ? q.y : ((f >> 1) == 2 ? q.z : q.w));
  return (word >> (16 * (f & 1))) & 0xffffu;
}

// Rows path: regular grid, at least 5 cells per axis, atom i not in an
// x-boundary cell. Each row of three x-adjacent cells is one run of the
// candidate stream, ascending in x (members are in (x, slot) order). A row
// whose cells lie at least gy along y and gz along z from the atom needs only
// members within ext = sqrt(r_s^2 - gy^2 - gz^2) along x, and none when
// gy^2 + gz^2 > r_s^2; the run starts and ends at the x-bins within ext of
// the atom, one bin of margin on each side (a bin is an eighth of a cell).
// The gaps are taken kGapSlack of a cell short, far above the rounding of the
// cell coordinates, so a skipped member is farther than ext along x, or
// farther than the gaps along y and z, by a margin that no rounding of r^2
// closes: its r^2 exceeds r_s^2 and the build would reject it. The accepted
// candidates and their order are those of the full scan. Writes row r (z
// neighbour r / 3, y neighbour r % 3) as {run begin, run end | (y shift code
// | z shift code << 2) << kRowCodeShift} for r = t, t + kBuildTeam, ... (the
// team's lanes share the rows), widens the block's span of row r, [lo[r],
// hi[r]), to cover a nonempty run, and returns false off the rows path.
__device__ inline bool row_runs(const BuildParams& p, uint32_t i, uint32_t t, uint2* rows,
                                uint32_t* lo_r, uint32_t* hi_r) {
  const CellGridView& g = p.grid;
  if (g.path != 1 || g.nx < 5 || g.ny < 5 || g.nz < 5) return false;
  const uint32_t c = p.cell_of[i];
  const uint32_t cx = c % g.nx, cyz = c / g.nx, cy = cyz % g.ny, cz = cyz / g.ny;
  if (cx < 1 || cx + 1 >= g.nx) return false;
  constexpr real_t kGapSlack = static_cast<real_t>(1.0 / 64.0);
  const real4 xi = p.pos[i];
  const real_t lo = p.box.lo[0], inv = g.inv_cell[0];
  const real_t fy = (xi.y - p.box.lo[1]) * g.inv_cell[1] - static_cast<real_t>(cy);
  const real_t fz = (xi.z - p.box.lo[2]) * g.inv_cell[2] - static_cast<real_t>(cz);
  // Gap to the row below (k = 0) and above (k = 2) along one axis.
  auto gap = [](real_t f, int k, real_t inv_cell) -> real_t {
    const real_t d = k == 0 ? f - kGapSlack : (static_cast<real_t>(1) - f) - kGapSlack;
    return k == 1 || d < 0 ? static_cast<real_t>(0) : d / inv_cell;
  };
#pragma unroll 1
  for (uint32_t r = t; r < static_cast<uint32_t>(kRows); r += kBuildTeam) {
    const int a = static_cast<int>(r / 3), b = static_cast<int>(r % 3);
    const real_t gz = gap(fz, a, g.inv_cell[2]);
    const real_t gy = gap(fy, b, g.inv_cell[1]);
    const real_t ext2 = p.rs2 - (gy * gy + gz * gz);
    if (ext2 < 0) {
      rows[r] = make_uint2(0u, 0u);
      continue;
    }
    const real_t ext = sqrt(ext2);
    const real_t fl = ((xi.x - ext) - lo) * inv - static_cast<real_t>(cx - 1);
    const real_t fr = ((xi.x + ext) - lo) * inv - static_cast<real_t>(cx + 1);
    const int m_left = static_cast<int>(floor(fl * static_cast<real_t>(kXBins))) - 1;
    const int m_right = static_cast<int>(floor(fr * static_cast<real_t>(kXBins))) + 2;
    const int ml = m_left < 0 ? 0 : (m_left > kXBins ? kXBins : m_left);
    const int mr = m_right < 0 ? 0 : (m_right > kXBins ? kXBins : m_right);
    uint32_t code_y, code_z;
    const uint32_t zc = neighbour_cell(cz, g.nz, a, code_z);
    const uint32_t rowc = g.nx * (neighbour_cell(cy, g.ny, b, code_y) + g.ny * zc);
    const uint32_t cl = rowc + cx - 1, cr = rowc + cx + 1;
    const uint32_t s_l = p.cell_start[cl], s_m = p.cell_start[cl + 1];
    const uint32_t s_r = p.cell_start[cr], e_r = p.cell_start[cr + 1];
    const uint32_t n_l = s_m - s_l, n_r = e_r - s_r;
    const uint32_t skip_l = n_l <= kXBinsMaxCount ? members_below(p.cell_xbins, cl, ml) : 0u;
    const uint32_t keep_r = n_r <= kXBinsMaxCount ? members_below(p.cell_xbins, cr, mr) : n_r;
    const uint32_t b_run = s_l + skip_l, e_run = s_r + keep_r;
    rows[r] = make_uint2(b_run, e_run | ((code_y | (code_z << 2)) << kRowCodeShift));
    if (b_run < e_run) {
      atomicMin(lo_r + r, b_run);
      atomicMax(hi_r + r, e_run);
    }
  }
  return true;
}

// Full-list build with the append rule of full_neighbor_path_common.cuh:49-94:
// skip self, skip excluded, keep r2 <= r_s^2 (reference minimum image and
// operation order), overflow past max_neighbors. Each atom has a team of
// kBuildTeam adjacent lanes; lane t takes candidates k + u*kBuildTeam + t
// (u < kBuildSlots) of every candidate run, so the team's loads are
// consecutive 16-byte entries of the cell-ordered position stream (slot in
// .w), one load per candidate with no index gather; slots past a run's end
// load whatever follows it (the streams are padded past the last atom) and
// test nothing. Accepted candidates are appended in
// candidate order (the order of a one-lane scan) through a ballot of the
// team's lanes and written straight into the atom's chunks. EXCL: 0 = no
// exclusions, 4 or 8 = up to that many partners held in registers (every
// candidate within r_s is compared with all of them), -1 = any count read
// from memory. Also captures the positions the displacement check compares
// against. HALF keeps only partners with a higher slot, so every pair is
// stored once.
template <int EXCL, bool HALF>
__global__ __launch_bounds__(kBlock) void build_kernel(BuildParams p) {
  if (p.ctrl->rebuild == 0) {
    return;
  }
  __shared__ double red_lds[4];
  const uint32_t t = threadIdx.x & (kBuildTeam - 1);
  const uint32_t team_lane0 = (threadIdx.x & 63u) & ~static_cast<uint32_t>(kBuildTeam - 1);
  const uint32_t below = (1u << t) - 1u;
  // The workgroup's atoms: kBuildAtoms consecutive entries of the candidate
  // stream (cell order) on the cell paths, consecutive slots on the direct
  // path.
  const uint32_t q = xcd_block(blockIdx.x, gridDim.x) * kBuildAtoms + threadIdx.x / kBuildTeam;
  const uint32_t i = q < p.n && p.grid.path != 0 ? id_of(p.cell_pos[q].w) : q;
  uint32_t count = 0;  // the same in every lane of a team
  // Rows path: each team's row runs, and the workgroup's span of each row,
  // staged in LDS (stage_off[r] - lo[r] is row r's stream-to-LDS offset)
  // when the spans fit; the teams then read their candidates from LDS.
  __shared__ uint2 row_lds[kBuildAtoms * kRows];
  extern __shared__ real4 stage[];  // p.stage_cap + kCandidatePad entries
  __shared__ uint32_t row_lo[kRows], row_hi[kRows], stage_off[kRows + 1];
  if (threadIdx.x < kRows) {
    row_lo[threadIdx.x] = 0xffffffffu;
    row_hi[threadIdx.x] = 0u;
  }
  __syncthreads();
  uint2* const team_rows = row_lds + (threadIdx.x / kBuildTeam) * kRows;
  const bool rows_path = i < p.n && row_runs(p, i, t, team_rows, row_lo, row_hi);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t off = 0;
    for (int r = 0; r < kRows; ++r) {
      stage_off[r] = off;
      off += row_hi[r] > row_lo[r] ? row_hi[r] - row_lo[r] : 0u;
    }
    stage_off[kRows] = off;
  }
  __syncthreads();
  const bool staged = stage_off[kRows] <= p.stage_cap;  // uniform over the workgroup
  if (staged) {
    for (int r = 0; r < kRows; ++r) {
      const uint32_t len = stage_off[r + 1] - stage_off[r];
      for (uint32_t k = threadIdx.x; k < len; k += kBlock) {
        stage[stage_off[r] + k] = p.cell_pos[row_lo[r] + k];
      }
    }
    if (rows_path) {
      // Re-base this lane's rows onto the stage (nonempty runs lie inside
      // their row's span).
      for (uint32_t r = t; r < static_cast<uint32_t>(kRows); r += kBuildTeam) {
        const uint2 d = team_rows[r];
        const uint32_t b_run = d.x, e_run = d.y & kRowEndMask;
        if (b_run < e_run) {
          const uint32_t shift = stage_off[r] - row_lo[r];
          team_rows[r] = make_uint2(b_run + shift, ((e_run + shift) & kRowEndMask) | (d.y & ~kRowEndMask));
        }
      }
    }
  }
  __syncthreads();
  if (i < p.n) {  // uniform over a team
    const real4 xi = p.pos[i];
    const uint32_t n = p.n;
    uint32_t ex[EXCL > 0 ? EXCL : 1];
    uint32_t ne = 0;
    uint32_t spill_ne = 0, spill_off = 0;  // exclusions past the inline capacity
    if (EXCL < 0 && p.spill_index != nullptr) {
      const uint32_t si = p.spill_index[i];
      if (si != kNoSpill) {
        spill_off = p.spill[si].off[3];
        spill_ne = p.spill[si].cnt[3];
      }
    }
    if (EXCL != 0) {
      ne = p.topo_count[i] >> 24;
      if (EXCL > 0) {
#pragma unroll
        for (int k = 0; k < EXCL; ++k) {
          ex[k] = k < static_cast<int>(ne) ? p.excl[static_cast<size_t>(k) * n + i] : 0xffffffffu;
        }
      }
    }
    const real_t L0 = p.box.L[0], L1 = p.box.L[1], L2 = p.box.L[2];
    const real_t h0 = p.box.hL[0], h1 = p.box.hL[1], h2 = p.box.hL[2];
    const uint32_t maxn = p.max_neighbors;
    // sh: the image shift of the candidate's cell, with mic(xi - xj) equal to
    // (xi - xj) + sh on every axis (cells of a regular grid with at least 5
    // cells per axis). mode: kShift adds sh on every axis, kShiftYZ only on y
    // and z (a row of x-adjacent cells has no x shift; adding a zero would
    // change nothing but the sign of a zero difference, which r^2 squares
    // away), kGeneric takes mic() per candidate.
    // Returns the wave's lanes that accept their candidate, as a lane mask:
    // every test is a compare whose ballot is its lane mask, so the tests
    // combine as scalar ANDs without branches.
    auto accept = [&](uint32_t j, real4 xj, auto mode, real3 sh) -> uint64_t {
      real_t dx, dy, dz;
      if (decltype(mode)::value == kGeneric) {
        dx = mic(xi.x - xj.x, L0, h0);
        dy = mic(xi.y - xj.y, L1, h1);
        dz = mic(xi.z - xj.z, L2, h2);
      } else {
        dx = decltype(mode)::value == kShift ? (xi.x - xj.x) + sh.x : xi.x - xj.x;
        dy = (xi.y - xj.y) + sh.y;
        dz = (xi.z - xj.z) + sh.z;
      }
      const real_t r2 = r2_strict(dx, dy, dz);
      uint64_t m = __builtin_amdgcn_ballot_w64(!(r2 > p.rs2)) &
                   __builtin_amdgcn_ballot_w64(HALF ? j > i : j != i);
      if (EXCL > 0) {
#pragma unroll
        for (int k = 0; k < EXCL; ++k) {
          m &= __builtin_amdgcn_ballot_w64(ex[k] != j);
        }
      } else if (EXCL < 0) {
        bool ok = __builtin_amdgcn_inverse_ballot_w64(m);
        if (ok) {
          for (uint32_t k = 0; k < ne; ++k) {
            if (p.excl[static_cast<size_t>(k) * n + i] == j) ok = false;
          }
          for (uint32_t k = 0; k < spill_ne; ++k) {
            if (p.spill_excl[spill_off + k] == j) ok = false;
          }
        }
        m = __builtin_amdgcn_ballot_w64(ok);
      }
      return m;
    };
    // Candidates [b, e) of a run (b < e) whose entry s is entry(s) = {x, y,
    // z, slot bits}, kBuildTeam * kBuildSlots per round.
    // Atom i's entries of chunk c start at c * chunk_words + row_word.
    const uint32_t row_word = i * kNbrChunk;
    const uint32_t chunk_words = n * kNbrChunk;
    // One round: candidates k + u * kBuildTeam + t (u < kBuildSlots) of the
    // run ending at e (k < e).
    auto round = [&](uint32_t k, uint32_t e, auto entry, auto mode, real3 sh) {
      // Entries past the run (at most kCandidatePad - 1 past the stream's
      // last atom) are loaded and then ignored.
      static_assert(kBuildTeam * kBuildSlots <= kCandidatePad, "candidate padding");
      real4 x[kBuildSlots];
#pragma unroll
      for (int u = 0; u < kBuildSlots; ++u) {
        x[u] = entry(k + t, u);
      }
      // Slot u of this lane is in the run while u * kBuildTeam < left.
      const int left = static_cast<int>(e - k) - static_cast<int>(t);
#pragma unroll
      for (int u = 0; u < kBuildSlots; ++u) {
        const uint32_t j = id_of(x[u].w);
        const uint64_t m =
            __builtin_amdgcn_ballot_w64(u * kBuildTeam < left) & accept(j, x[u], mode, sh);
        const uint32_t bits = static_cast<uint32_t>(m >> team_lane0) & ((1u << kBuildTeam) - 1u);
        if (__builtin_amdgcn_inverse_ballot_w64(m)) {
          const uint32_t slot = count + __builtin_popcount(bits & below);
          if (slot < maxn) {
            // Entry (chunk slot / kNbrChunk, atom i, slot % kNbrChunk);
            // a chunk holds n * kNbrChunk < 2^32 entries.
            const uint64_t w = static_cast<uint64_t>(slot / kNbrChunk) * chunk_words +
                               (row_word | (slot & (kNbrChunk - 1)));
            p.nbr[w] = j;
          }
        }
        count += __builtin_popcount(bits);
      }
    };
    auto scan_range = [&](uint32_t b, uint32_t e, auto entry, auto mode, real3 sh) {
      for (uint32_t k = b; k < e; k += kBuildTeam * kBuildSlots) {
        round(k, e, entry, mode, sh);
      }
    };
    const real3 no_shift = make_real3(0, 0, 0);
    // entry(k0, u): the stream's entry k0 + u * kBuildTeam, addressed by a
    // 32-bit byte offset from the stream's base ((n + kCandidatePad) *
    // sizeof(real4) < 2^32) plus u's constant, so that the slots of a round
    // are immediate offsets of one address.
    constexpr uint32_t kEntryBytes = sizeof(real4);
    constexpr size_t kSlotBytes = kBuildTeam * sizeof(real4);
    if (p.grid.path == 0) {
      const char* pos_bytes = reinterpret_cast<const char*>(p.pos);
      const auto entry = [pos_bytes](uint32_t k0, int u) {
        const real4 q = *reinterpret_cast<const real4*>(
            pos_bytes + static_cast<size_t>(k0 * kEntryBytes) + u * kSlotBytes);
        return make_real4(q.x, q.y, q.z, real_of_id(k0 + u * kBuildTeam));
      };
      scan_range(0, n, entry, kGenericMode{}, no_shift);
    } else {
      const uint32_t c = p.cell_of[i];
      const CellGridView& g = p.grid;
      const uint32_t cx = c % g.nx, cyz = c / g.nx, cy = cyz % g.ny, cz = cyz / g.ny;
      const char* cp_bytes = reinterpret_cast<const char*>(p.cell_pos);
      const auto entry = [cp_bytes](uint32_t k0, int u) {
        return *reinterpret_cast<const real4*>(
            cp_bytes + static_cast<size_t>(k0 * kEntryBytes) + u * kSlotBytes);
      };
      const int mx = g.nx < 3 ? static_cast<int>(g.nx) : 3;
      const int my = g.ny < 3 ? static_cast<int>(g.ny) : 3;
      const int mz = g.nz < 3 ? static_cast<int>(g.nz) : 3;
      // The team's rows in order, one run after another: a team moves to its
      // next row as soon as its run ends, so that the wave's lanes wait only
      // for the team with the most rounds in all, not in each row.
      auto scan_rows = [&](auto row_entry) {
        uint32_t r = 0, k = 0, e = 0;
        real3 sh = make_real3(0, 0, 0);
        for (;;) {
          if (k >= e) {
            bool more = true;
            do {
              if (r == static_cast<uint32_t>(kRows)) {
                more = false;
                break;
              }
              const uint2 d = team_rows[r++];
              k = d.x;
              e = d.y & kRowEndMask;
              const uint32_t code = d.y >> kRowCodeShift;
              sh = make_real3(0, shift_of(code & 3u, L1), shift_of(code >> 2, L2));
            } while (k >= e);
            if (!more) break;
          }
          round(k, e, row_entry, kShiftYZMode{}, sh);
          k += kBuildTeam * kBuildSlots;
        }
      };
      auto cells = [&](auto mode) {
#pragma unroll 1
        for (int a = 0; a < mz; ++a) {
          uint32_t code_z;
          const uint32_t zc = neighbour_cell(cz, g.nz, a, code_z);
          const real_t sz = shift_of(code_z, L2);
#pragma unroll 1
          for (int b = 0; b < my; ++b) {
            uint32_t code_y;
            const uint32_t rowc = g.nx * (neighbour_cell(cy, g.ny, b, code_y) + g.ny * zc);
            const real_t sy = shift_of(code_y, L1);
#pragma unroll 1
            for (int d = 0; d < mx; ++d) {
              uint32_t code_x;
              const uint32_t nc = neighbour_cell(cx, g.nx, d, code_x) + rowc;
              scan_range(p.cell_start[nc], p.cell_start[nc + 1], entry, mode,
                         make_real3(shift_of(code_x, L0), sy, sz));
            }
          }
        }
      };
      if (rows_path && staged) {
        scan_rows([](uint32_t k0, int u) { return stage[k0 + u * kBuildTeam]; });
      } else if (rows_path) {
        scan_rows(entry);
      } else if (g.path == 1 && g.nx >= 5 && g.ny >= 5 && g.nz >= 5) {
        cells(kShiftMode{});
      } else {
        cells(kGenericMode{});
      }
    }
    if (t == 0) {
      p.nbr_count[i] = min(count, maxn);
      p.ref_pos[i] = xi;
      p.ref_image[i] = p.image[i];
    }
  }
  const bool counted = t == 0 && i < p.n;
  const uint32_t stored = counted ? min(count, p.max_neighbors) : 0u;
  const double total = block_sum<kBlock>(static_cast<double>(stored), red_lds);
  if (threadIdx.x == 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&p.ctrl->neighbor_total),
              static_cast<unsigned long long>(total));
  }
  if (counted && count > p.max_neighbors) {
    atomicOr(&p.ctrl->overflow, 1u);
  }
}

// ------------------------------------------------------------------ force --

// LJ over one atom's stored list by one lane (pair_lj.cu:21-167), used by the
// half-list pair pass. The loop reads two 16-byte quads of indices and issues
// the eight position gathers together before any arithmetic. The decision
// r^2 < r_c^2 is binary32 on the exact-intermediate minimum image; the pair
// value is FP64. REACT adds the reaction -s d of every pair to pair_acc[j].
// TYPES: 0 = single type (coefficients in registers), 1 = table in LDS,
// 2 = table in global memory.
template <bool OBS, int TYPES, bool REACT>
__device__ __forceinline__ void lj_list(const ForceParams& p, uint32_t i, real4 xi4,
                                        const LjCoeff* lj_lds, const LjCoeff& c0, double& fxd,
                                        double& fyd, double& fzd, double& e_pair,
                                        double& w_pair) {
  const int T = p.n_types;
  const int ti = static_cast<int>(id_of(xi4.w));
  const uint32_t n = p.n;
  const real_t L0 = p.box.L[0], L1 = p.box.L[1], L2 = p.box.L[2];
  const real_t h0 = p.box.hL[0], h1 = p.box.hL[1], h2 = p.box.hL[2];
  const uint32_t cnt = p.nbr_count[i];
  // Eight slots per iteration: both index quads of a half chunk, then eight
  // gathers in flight before any arithmetic. Slots at or past cnt gather
  // atom i itself (never used), so every load stays inside pos and inside the
  // atom's allocated chunks (kNbrChunk is a multiple of 8).
  static_assert(kNbrChunk % 8 == 0, "neighbour chunks hold whole 8-slot groups");
  for (uint32_t base = 0; base < cnt; base += 8) {
    const uint32_t* cp = p.nbr + nbr_chunk_base(base / kNbrChunk, i, n) + (base & (kNbrChunk - 1));
    const uint4 qa = *reinterpret_cast<const uint4*>(cp);
    const uint4 qb = *reinterpret_cast<const uint4*>(cp + 4);
    const uint32_t jj[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
    real4 xj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      xj[u] = p.pos[base + u < cnt ? jj[u] : i];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const real_t dx = mic_diff(xi4.x, xj[u].x, L0, h0);
      const real_t dy = mic_diff(xi4.y, xj[u].y, L1, h1);
      const real_t dz = mic_diff(xi4.z, xj[u].z, L2, h2);
      const real_t r2 = r2_strict(dx, dy, dz);
      if (base + u < cnt && r2 < p.rc2) {
        LjCoeff c;
        if (TYPES == 0) {
          c = c0;
        } else if (TYPES == 1) {
          c = lj_lds[ti * T + static_cast<int>(id_of(xj[u].w))];
        } else {
          c = p.lj[ti * T + static_cast<int>(id_of(xj[u].w))];
        }
        double rr, u_pair;
        const double s = lj_pair(dx, dy, dz, c, rr, u_pair);
        fxd += s * dx;
        fyd += s * dy;
        fzd += s * dz;
        if (REACT) {
          double* acc = p.pair_acc + 3 * static_cast<size_t>(jj[u]);
          unsafeAtomicAdd(acc, -s * dx);
          unsafeAtomicAdd(acc + 1, -s * dy);
          unsafeAtomicAdd(acc + 2, -s * dz);
        }
        if (OBS) {
          e_pair += u_pair;
          w_pair += s * rr;
        }
      }
    }
  }
}

// LJ over atom i's stored list, shared by the kTeam lanes of its team: lane t
// takes slots base + u*kTeam + t, u < kTeamSlots, and keeps FP64 partial sums
// (pair_lj.cu:21-167; same decision and pair value as lj_list). Slots at or
// past cnt neither load an index nor count; they gather atom i itself.
template <bool OBS, int TYPES>
__device__ __forceinline__ void lj_team(const ForceParams& p, uint32_t i, uint32_t t, real4 xi4,
                                        const LjCoeff* lj_lds, const LjCoeff& c0, double& fxd,
                                        double& fyd, double& fzd, double& e_pair,
                                        double& w_pair) {
  static_assert(kNbrChunk % kTeam == 0, "a team's slots of one round share a chunk");
  static_assert(kTeam * kTeamSlots == kNbrRound, "a round covers the padded capacity's unit");
  const int T = p.n_types;
  const int ti = static_cast<int>(id_of(xi4.w));
  const uint32_t n = p.n;
  const real_t L0 = p.box.L[0], L1 = p.box.L[1], L2 = p.box.L[2];
  const real_t h0 = p.box.hL[0], h1 = p.box.hL[1], h2 = p.box.hL[2];
  const real_t xm0 = xi4.x - L0, xm1 = xi4.y - L1, xm2 = xi4.z - L2;
  const uint32_t cnt = p.nbr_count[i];
  for (uint32_t base = 0; base < cnt; base += kTeam * kTeamSlots) {
    // Every slot of the round lies inside the atom's padded capacity; slots
    // at or past the count read whatever the chunk holds and become i.
    uint32_t jj[kTeamSlots];
#pragma unroll
    for (int u = 0; u < kTeamSlots; ++u) {
      const uint32_t s = base + u * kTeam + t;
      jj[u] = p.nbr[nbr_chunk_base(s / kNbrChunk, i, n) + (s & (kNbrChunk - 1))];
    }
    // slot base + u * kTeam + t holds a neighbour iff u * kTeam < rem
    const int rem = static_cast<int>(cnt) - static_cast<int>(base + t);
#pragma unroll
    for (int u = 0; u < kTeamSlots; ++u) {
      jj[u] = u * kTeam < rem ? jj[u] : i;
    }
    // Gathers as a 32-bit byte offset from the array base (n * sizeof(real4)
    // < 2^32, checked at set-up).
    const char* pos_bytes = reinterpret_cast<const char*>(p.pos);
    real4 xj[kTeamSlots];
#pragma unroll
    for (int u = 0; u < kTeamSlots; ++u) {
      xj[u] = *reinterpret_cast<const real4*>(
          pos_bytes + jj[u] * static_cast<uint32_t>(sizeof(real4)));
    }
#pragma unroll
    for (int u = 0; u < kTeamSlots; ++u) {
      // mic_diff_sel returns the raw difference unless it exceeds half the
      // box; the folded forms are formed only when some lane of the wave
      // needs them (pairs across a periodic boundary), the same values.
      real_t dx = xi4.x - xj[u].x;
      real_t dy = xi4.y - xj[u].y;
      real_t dz = xi4.z - xj[u].z;
      if ((__builtin_amdgcn_ballot_w64(fabs(dx) > h0) | __builtin_amdgcn_ballot_w64(fabs(dy) > h1) |
           __builtin_amdgcn_ballot_w64(fabs(dz) > h2)) != 0) {
        dx = mic_diff_sel(xi4.x, xm0, xj[u].x, L0, h0);
        dy = mic_diff_sel(xi4.y, xm1, xj[u].y, L1, h1);
        dz = mic_diff_sel(xi4.z, xm2, xj[u].z, L2, h2);
      }
      const real_t r2 = r2_strict(dx, dy, dz);
      if (u * kTeam < rem && r2 < p.rc2) {
        LjCoeff c;
        if (TYPES == 0) {
          c = c0;
        } else if (TYPES == 1) {
          c = lj_lds[ti * T + static_cast<int>(id_of(xj[u].w))];
        } else {
          c = p.lj[ti * T + static_cast<int>(id_of(xj[u].w))];
        }
        double rr, u_pair;
        const double s = lj_pair(dx, dy, dz, c, rr, u_pair);
        fxd += s * dx;
        fyd += s * dy;
        fzd += s * dz;
        if (OBS) {
          e_pair += u_pair;
          w_pair += s * rr;
        }
      }
    }
  }
}

template <int TYPES>
__device__ __forceinline__ void load_lj_table(const ForceParams& p, LjCoeff* lj_lds) {
  if (TYPES == 1) {
    for (int k = threadIdx.x; k < p.n_types * p.n_types; k += kBlock) {
      lj_lds[k] = p.lj[k];
    }
    __syncthreads();
  }
}

// Half-list pair pass (no reference counterpart; the reference keeps a full
// list): every stored pair (i, j > i) is evaluated once, the force on i is
// summed in registers and the reaction on j added to pair_acc with FP64
// atomics; i's own sum is added last. Pair energy and virial are summed per
// workgroup and added to ctrl->half_pair. The force pass that follows adds the
// listed terms, kicks, reduces and clears pair_acc.
template <bool OBS, int TYPES>
__global__ __launch_bounds__(kBlock) void half_pair_kernel(ForceParams p) {
  __shared__ LjCoeff lj_lds[TYPES == 1 ? kMaxLdsTypes * kMaxLdsTypes : 1];
  __shared__ double red_lds[4];
  load_lj_table<TYPES>(p, lj_lds);
  const LjCoeff c0 = TYPES == 0 ? p.lj[0] : LjCoeff{0.0, 0.0, 0.0, 0.0};
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  double e_pair = 0.0, w_pair = 0.0;
  if (i < p.n) {
    double fxd = 0.0, fyd = 0.0, fzd = 0.0;
    lj_list<OBS, TYPES, true>(p, i, p.pos[i], lj_lds, c0, fxd, fyd, fzd, e_pair, w_pair);
    double* acc = p.pair_acc + 3 * static_cast<size_t>(i);
    unsafeAtomicAdd(acc, fxd);
    unsafeAtomicAdd(acc + 1, fyd);
    unsafeAtomicAdd(acc + 2, fzd);
  }
  if (OBS) {
    const double e = block_sum<kBlock>(e_pair, red_lds);
    const double w = block_sum<kBlock>(w_pair, red_lds);
    if (threadIdx.x == 0) {
      unsafeAtomicAdd(&p.ctrl->half_pair[0], e);
      unsafeAtomicAdd(&p.ctrl->half_pair[1], w);
    }
  }
}

// Atom i's total force (rounded once from FP64), the force store, optionally
// the post-force half-kick (velocity_verlet.cu:97-111), and the atom's
// kinetic energy. Run by one lane of the team.
template <bool KICK>
__device__ __forceinline__ void finish_atom(const ForceParams& p, uint32_t i, double fxd,
                                            double fyd, double fzd, StepSums& sums) {
  const real_t fx = static_cast<real_t>(fxd);
  const real_t fy = static_cast<real_t>(fyd);
  const real_t fz = static_cast<real_t>(fzd);
  p.force[i] = make_real4(fx, fy, fz, 0);
  real4 v = p.vel[i];
  if (KICK) {
    const real_t h = real_t(0.5) * p.dt / v.w;
    v.x += h * fx;
    v.y += h * fy;
    v.z += h * fz;
    p.vel[i] = v;
  }
  const double vx = v.x, vy = v.y, vz = v.z;
  sums.ke += 0.5 * static_cast<double>(v.w) * (vx * vx + vy * vy + vz * vz);
}

// Workgroup sums of one force pass: thread 0 writes partials[q * nbk + block]
// (q = 0 kinetic energy; with OBS 1..4 energies, 5..8 virials).
template <bool OBS>
__device__ __forceinline__ void store_step_partials(const ForceParams& p, const StepSums& sums,
                                                    double* red_lds) {
  const uint32_t nbk = gridDim.x;
  double* part = p.partials;
  double s = block_sum<kBlock>(sums.ke, red_lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (OBS) {
    const double vals[8] = {0.5 * sums.e_pair, sums.e_bond, sums.e_angle, sums.e_dih,
                            0.5 * sums.w_pair, sums.w_bond, sums.w_angle, sums.w_dih};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      s = block_sum<kBlock>(vals[q], red_lds);
      if (threadIdx.x == 0) part[(q + 1) * static_cast<size_t>(nbk) + blockIdx.x] = s;
    }
  }
}

// One workgroup after the force pass: FP64 totals of the nbk workgroups'
// partials, the Berendsen factor (berendsen_thermostat.cu:109-163), and the
// per-step flags cleared. A kernel of its own, so the force pass publishes
// its partials at the kernel boundary instead of with a device-scope fence
// (an L2 write-back and invalidate on gfx950) in every workgroup.
constexpr int kReduceBlock = 1024;

// This thread's share of v[0..n) for a kReduceBlock-wide sum: elements
// threadIdx.x + m*kReduceBlock, four independent loads in flight.
__device__ __forceinline__ double strided_sum(const double* v, uint32_t n) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  uint32_t b = threadIdx.x;
  for (; b + 3 * kReduceBlock < n; b += 4 * kReduceBlock) {
    a0 += v[b];
    a1 += v[b + kReduceBlock];
    a2 += v[b + 2 * kReduceBlock];
    a3 += v[b + 3 * kReduceBlock];
  }
  for (; b < n; b += kReduceBlock) {
    a0 += v[b];
  }
  return (a0 + a1) + (a2 + a3);
}

template <bool KICK, bool OBS, bool HALF>
__global__ __launch_bounds__(kReduceBlock) void step_reduce_kernel(ForceParams p, uint32_t nbk) {
  __shared__ double red_lds[kReduceBlock / 64];
  const double* part = p.partials;
  const int nq = OBS ? 9 : 1;
  for (int q = 0; q < nq; ++q) {
    const double acc = block_sum<kReduceBlock>(strided_sum(part + q * static_cast<size_t>(nbk), nbk),
                                               red_lds);
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
        ctrl->pe[q - 1] = (HALF && q == 1) ? ctrl->half_pair[0] : acc;
      } else {
        ctrl->virial[q - 5] = (HALF && q == 5) ? ctrl->half_pair[1] : acc;
      }
    }
  }
  if (threadIdx.x == 0) {
    p.ctrl->rebuild = 0;
    p.ctrl->reorder = 0;
    if (HALF) {
      p.ctrl->half_pair[0] = 0.0;
      p.ctrl->half_pair[1] = 0.0;
    }
  }
}

// Fused force pass: LJ over the stored full list (pair_lj.cu:21-167), or the
// half-list pair pass's sums (HALF), plus the listed terms, then optionally
// the post-force half-kick (velocity_verlet.cu:97-111) and the kinetic-energy
// reduction for the Berendsen factor (berendsen_thermostat.cu:109-163) as
// per-workgroup partials, which step_reduce_kernel totals in FP64.
//
// Each atom has a team of kTeam adjacent lanes. Lane t reads list slots
// base + u*kTeam + t (u < kTeamSlots), so one gather instruction fetches
// kTeam consecutive entries of each of the wave's 64/kTeam lists: mostly
// consecutive slots of one cell, a few cache lines per atom instead of one
// line per lane. The team's FP64 partial sums meet in a shuffle butterfly and
// the atom's total is rounded to binary32 once.
template <bool KICK, bool OBS, int TYPES, bool HALF, bool SPILL>
__global__ __launch_bounds__(kBlock) void force_kernel(ForceParams p) {
  __shared__ LjCoeff lj_lds[TYPES == 1 && !HALF ? kMaxLdsTypes * kMaxLdsTypes : 1];
  __shared__ double red_lds[4];
  if (!HALF) load_lj_table<TYPES>(p, lj_lds);
  const LjCoeff c0 = TYPES == 0 && !HALF ? p.lj[0] : LjCoeff{0.0, 0.0, 0.0, 0.0};
  const uint32_t t = threadIdx.x & (kTeam - 1);
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * kForceAtoms + threadIdx.x / kTeam;
  StepSums sums;
  if (i < p.n) {  // uniform over a team
    const real4 xi4 = p.pos[i];
    double fxd = 0.0, fyd = 0.0, fzd = 0.0;
    if (HALF) {
      if (t == 0) {
        double* acc = p.pair_acc + 3 * static_cast<size_t>(i);
        fxd = acc[0];
        fyd = acc[1];
        fzd = acc[2];
        acc[0] = 0.0;
        acc[1] = 0.0;
        acc[2] = 0.0;
      }
    } else {
      lj_team<OBS, TYPES>(p, i, t, xi4, lj_lds, c0, fxd, fyd, fzd, sums.e_pair, sums.w_pair);
    }
    if (p.topo.has_listed) {
      listed_team<OBS, SPILL, kTeam>(p.pos, p.topo, p.box, p.n, i, t, xi4, fxd, fyd, fzd, sums);
    }
    fxd = team_sum<kTeam>(fxd);
    fyd = team_sum<kTeam>(fyd);
    fzd = team_sum<kTeam>(fzd);
    if (t == 0) {
      finish_atom<KICK>(p, i, fxd, fyd, fzd, sums);
    }
  }
  store_step_partials<OBS>(p, sums, red_lds);
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
    real4 v = p.vel[i];
    if (THERMO) {
      const real_t lam = static_cast<real_t>(p.ctrl->lambda);
      v.x *= lam;
      v.y *= lam;
      v.z *= lam;
      p.vel[i] = v;
    }
    const double vx = v.x, vy = v.y, vz = v.z;
    ke = 0.5 * static_cast<double>(v.w) * (vx * vx + vy * vy + vz * vz);
  }
  const double s = block_sum<kBlock>(ke, red_lds);
  if (threadIdx.x == 0) p.partials[blockIdx.x] = s;
}

// One workgroup after scale_kernel: FP64 total of its nbk partials.
__global__ __launch_bounds__(kReduceBlock) void scale_reduce_kernel(ScaleParams p, uint32_t nbk) {
  __shared__ double red_lds[kReduceBlock / 64];
  const double acc = block_sum<kReduceBlock>(strided_sum(p.partials, nbk), red_lds);
  if (threadIdx.x == 0) {
    p.ctrl->ke = acc;
    p.ctrl->lambda = 1.0;
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
  real4 v = p.vel[i];
  const real4 f = p.force[i];
  real4 x = p.pos[i];
  if (THERMO) {
    // Rounded on its own, as the reference's separate rescale kernel and
    // scale_kernel (output steps) round it: never fused into the kick.
#pragma clang fp contract(off)
    const real_t lam = static_cast<real_t>(p.ctrl->lambda);
    v.x *= lam;
    v.y *= lam;
    v.z *= lam;
  }
  const real_t h = real_t(0.5) * p.dt / v.w;
  v.x += h * f.x;
  v.y += h * f.y;
  v.z += h * f.z;
  x.x += p.dt * v.x;
  x.y += p.dt * v.y;
  x.z += p.dt * v.z;
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
  p.vel[i] = v;
  const bool wrapped = c[0] != real_t(0) || c[1] != real_t(0) || c[2] != real_t(0);
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


}  // namespace

void launch_control(Ctrl* ctrl, NeighborControl control, bool forced, uint64_t steps,
                    hipStream_t stream) {
  control_kernel<<<1, 64, 0, stream>>>(ctrl, control, forced ? 1 : 0, steps);
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
  constexpr unsigned kMaxSortBlocks = 4096;
  const unsigned g = grid_for(p.n_cells, kSortCellsPerBlock);
  cell_sort_kernel<<<g < kMaxSortBlocks ? g : kMaxSortBlocks, kBlock, 0, stream>>>(p);
}

void launch_reorder(const ReorderParams& p, hipStream_t stream) {
  reorder_gather_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
  reorder_commit_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
}

void launch_pack_records(const PackParams& p, hipStream_t stream) {
  pack_records_kernel<<<grid_for(p.n), kBlock, 0, stream>>>(p);
}

void launch_build(const BuildParams& params, hipStream_t stream) {
  const unsigned g = grid_for(params.n, kBuildAtoms);
  // Stage capacity from the mean cell occupancy m: a workgroup's kBuildAtoms
  // atoms span about kBuildAtoms / m cells of one row of cells, so each of the
  // kRows rows of the stencil holds about kBuildAtoms + 2m members; 15 %
  // margin, whole multiples of 64 entries, at most kStageBytesMax.
  BuildParams p = params;
  const double m = p.grid.n_cells > 0 ? static_cast<double>(p.n) / p.grid.n_cells : 0.0;
  const uint32_t want = static_cast<uint32_t>(1.15 * kRows * (kBuildAtoms + 2.0 * m) + 63.0) & ~63u;
  const uint32_t most = kStageBytesMax / sizeof(real4) - kCandidatePad;
  p.stage_cap = p.grid.path == 1 ? (want < most ? want : most) : 0u;
  const size_t lds = p.grid.path == 1 ? (p.stage_cap + kCandidatePad) * sizeof(real4) : 0;
  // Exclusion registers: the fewest of 4 / 8 that hold every slot's
  // exclusions (ecap is the largest count), else read from memory.
  const int ex = p.spill_index != nullptr
                     ? -1
                     : (p.ecap == 0 ? 0 : (p.ecap <= 4 ? 4 : (p.ecap <= 8 ? 8 : -1)));
#define BEADS_BUILD_CASE(E, H) \
  if (ex == E && (p.half != 0) == H) build_kernel<E, H><<<g, kBlock, lds, stream>>>(p);
  BEADS_BUILD_CASE(0, false)
  BEADS_BUILD_CASE(4, false)
  BEADS_BUILD_CASE(8, false)
  BEADS_BUILD_CASE(-1, false)
  BEADS_BUILD_CASE(0, true)
  BEADS_BUILD_CASE(4, true)
  BEADS_BUILD_CASE(8, true)
  BEADS_BUILD_CASE(-1, true)
#undef BEADS_BUILD_CASE
}

void launch_force(const ForceParams& p, bool kick, bool obs, hipStream_t stream) {
  const unsigned g = grid_for(p.n);
  const int types = p.n_types == 1 ? 0 : (p.n_types <= kMaxLdsTypes ? 1 : 2);
  const bool half = p.half != 0;
  if (half) {
#define BEADS_HALF_CASE(O, T)                                           \
    if (obs == O && types == T) half_pair_kernel<O, T><<<g, kBlock, 0, stream>>>(p);
    BEADS_HALF_CASE(false, 0)
    BEADS_HALF_CASE(false, 1)
    BEADS_HALF_CASE(false, 2)
    BEADS_HALF_CASE(true, 0)
    BEADS_HALF_CASE(true, 1)
    BEADS_HALF_CASE(true, 2)
#undef BEADS_HALF_CASE
  }
  const unsigned gf = force_grid(p.n);
  const bool spill = p.topo.spill_index != nullptr;
#define BEADS_FORCE_CASE(K, O, T)                                                      \
  if (kick == K && obs == O && (half || types == T)) {                                 \
    if (half) {                                                                        \
      if (spill) {                                                                     \
        force_kernel<K, O, 0, true, true><<<gf, kBlock, 0, stream>>>(p);               \
      } else {                                                                         \
        force_kernel<K, O, 0, true, false><<<gf, kBlock, 0, stream>>>(p);              \
      }                                                                                \
      step_reduce_kernel<K, O, true><<<1, kReduceBlock, 0, stream>>>(p, gf);           \
    } else {                                                                           \
      if (spill) {                                                                     \
        force_kernel<K, O, T, false, true><<<gf, kBlock, 0, stream>>>(p);              \
      } else {                                                                         \
        force_kernel<K, O, T, false, false><<<gf, kBlock, 0, stream>>>(p);             \
      }                                                                                \
      step_reduce_kernel<K, O, false><<<1, kReduceBlock, 0, stream>>>(p, gf);          \
    }                                                                                  \
    return;                                                                            \
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
  const unsigned g = grid_for(p.n);
  if (thermostat) scale_kernel<true><<<g, kBlock, 0, stream>>>(p);
  else scale_kernel<false><<<g, kBlock, 0, stream>>>(p);
  scale_reduce_kernel<<<1, kReduceBlock, 0, stream>>>(p, g);
}

}  // namespace beads_amd
