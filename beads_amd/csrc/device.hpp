// Parameter blocks shared by the host engine and the gfx950 kernels.
//
// Device layout (slot order, N atoms, all records coalesced vectors):
//   pos    real4[N]   x, y, z, w = bits of the 0-based type id (id_of)
//   vel    real4[N]   vx, vy, vz, w = mass
//   force  real4[N]   fx, fy, fz, w unused
//   (real4 is float4, 16 bytes; double4, 32 bytes, in the BEADS_REAL64 build)
//   image  int4[N]    ix, iy, iz, unused
//   nbr_count u32[N], nbr u32[ceil(max_neighbors / 16) * 16 * N] in 64-byte
//     chunks: entries 16c..16c+15 of atom i at ((c * N) + i) * 16 (the
//     reference's slot-major k * N + i, neighbor_list.cuh:12-19, is what
//     beads_engine_read_neighbors returns)
//   pair_acc double[3 N] (half list only): partner-side LJ sums, zeroed by
//     the force pass that consumes them
//   per-slot listed-term roles (rebuilt on every reorder):
//     topo_count u32[N] = bonds | angles << 8 | dihedrals << 16 | exclusions << 24
//     bond_rec  uint2[bcap * N]  {partner slot, type | first-role << 31}
//     angle_rec uint4[acap * N]  {other slot, other slot, type, role 0..2}
//     dih_rec   uint4[dcap * N]  {other, other, other, type | role << 30}
//     excl      u32[ecap * N]    excluded partner slots
#pragma once

#include <cstdint>

#include "real.hpp"

namespace beads_amd {

constexpr int kMaxLdsTypes = 16;  // LJ type tables up to 16x16 live in LDS
constexpr uint32_t kNbrChunk = 16;  // neighbour entries per stored chunk (64 bytes)

// Entries allocated per atom for a capacity of max_neighbors: whole pairs of
// chunks, so that the force pass's rounds of 2 * kNbrChunk slots can load
// every slot of a round without a bounds test (entries past the count are
// read and ignored).
constexpr uint32_t kNbrRound = 2 * kNbrChunk;
inline uint32_t nbr_padded_capacity(uint32_t max_neighbors) {
  return (max_neighbors + kNbrRound - 1) / kNbrRound * kNbrRound;
}

// LJ coefficients of one ordered type pair, formed in FP64 from the binary32
// epsilon, sigma and cutoff (pair_lj.cu:171-192 packs the same quantities in
// binary32): sigma^2, 24 eps, 4 eps and the energy shift
// 4 eps ((sigma/rc)^12 - (sigma/rc)^6) (0 when unshifted).
struct LjCoeff {
  double sigma2, eps24, eps4, shift;
};

struct BoxView {
  real_t lo[3];
  real_t L[3];
  real_t hL[3];
  real_t invL[3];
};

// Members of a cell are ordered by (x, slot). Each cell's x extent is cut
// into kXBins equal bins; cell_xbins holds, for m = 1..kXBins, the number of
// members whose bin is below m, as 16-bit counts (word (m-1)/2, half (m-1)%2),
// so that the build can skip the members of an x-neighbour cell that lie
// farther than r_s along x. Cells of more than kXBinsMaxCount members carry no
// table and are scanned whole.
constexpr int kXBins = 8;
constexpr uint32_t kXBinsMaxCount = 65535;

struct CellGridView {
  uint32_t nx, ny, nz, n_cells;
  real_t inv_cell[3];
  int32_t path;  // 0 direct, 1 regular, 2 small dimension
};

// Device-resident control block. Written only by kernels (and zeroed once at
// set-up), read back asynchronously by the host.
struct alignas(256) Ctrl {
  uint32_t rebuild;         // a rebuild is due in this step
  uint32_t reorder;         // the rebuild also sorts particles into cell order
  uint32_t checked;         // a scheduled check ran since the last build
  uint32_t overflow;        // a neighbour list overflowed
  uint64_t build_count;
  uint64_t dangerous;
  uint64_t reorder_count;
  uint64_t neighbor_total;  // sum of neighbour counts of the current list
  uint64_t clock;           // steps advanced (counted by the control kernel)
  uint64_t last_sort;       // clock at the last reorder
  double lambda;            // Berendsen factor still to be applied to vel
  double ke_pre;            // kinetic energy after the post-force half-kick
  double ke;                // kinetic energy of the velocities as stored
  double pe[4];             // pair, bond, angle, dihedral
  double virial[4];
  double half_pair[2];      // half list: pair energy and virial summed by the pair pass
};

struct NeighborControl {
  uint64_t check_every;
  uint64_t sort_every;       // reorder every sort_every-th build ...
  uint64_t sort_every_steps; // ... or, when nonzero, at the first build at least this
                             // many steps after the last reorder
  int32_t path;
};

// Listed-term roles and exclusions are stored per slot, slot-major, up to an
// inline capacity per kind (bond, angle, dihedral, exclusion); a particle
// with more records of a kind keeps the rest in its own contiguous segment of
// that kind's spill array. spill_index[slot] names the particle's SpillEntry
// (kNoSpill when it has none); the entry moves with the particle's slot and
// its records' partner slots are remapped by every reorder.
constexpr uint32_t kNoSpill = 0xffffffffu;
struct SpillEntry {
  uint32_t off[4];  // first record in the spill array of each kind
  uint32_t cnt[4];  // records of each kind beyond the inline capacity
};

struct TopologyView {
  const uint32_t* count;
  const uint2* bond;
  const uint4* angle;
  const uint4* dih;
  const uint32_t* excl;
  uint32_t bcap, acap, dcap, ecap;
  const real2* bond_coeff;   // k, r0
  const real2* angle_coeff;  // k, theta0 (rad)
  const real4* dih_coeff;    // k, d, n, unused
  int32_t has_listed;
  // null when no particle spills
  const uint32_t* spill_index;
  const SpillEntry* spill;
  const uint2* spill_bond;
  const uint4* spill_angle;
  const uint4* spill_dih;
};

struct ForceParams {
  uint32_t n;
  const real4* pos;
  real4* vel;
  real4* force;
  const uint32_t* nbr_count;
  const uint32_t* nbr;
  BoxView box;
  real_t rc2;
  int32_t n_types;
  const LjCoeff* lj;  // [T*T]
  TopologyView topo;
  real_t dt;
  int32_t berendsen;
  double target_thermal;  // kB * T0
  double dt_over_tau;
  double dof;
  double* partials;       // per block: ke, pe[4], vir[4]
  Ctrl* ctrl;
  uint32_t* cell_count;   // zeroed after a rebuild consumed it
  uint32_t n_cells;
  int32_t half;           // stored list holds each pair once (j > i)
  double* pair_acc;       // [3 N] LJ sums of the half-list pair pass
};

struct IntegrateParams {
  uint32_t n;
  real4* pos;
  real4* vel;
  const real4* force;
  int4* image;
  const real4* ref_pos;
  const int4* ref_image;
  BoxView box;
  real_t dt;
  real_t threshold_sq;
  Ctrl* ctrl;
};

struct BuildParams {
  uint32_t n;
  const real4* pos;
  const int4* image;
  real4* ref_pos;
  int4* ref_image;
  BoxView box;
  CellGridView grid;
  real_t rs2;
  const uint32_t* cell_of;
  const uint32_t* cell_start;
  const uint32_t* cell_count;
  const uint32_t* cell_atoms;
  const real4* cell_pos;  // cell_atoms' positions, .w = the slot's bits
  const uint4* cell_xbins; // per cell: x-bin prefix counts (cell_sort_kernel)
  const uint32_t* topo_count;
  const uint32_t* excl;
  uint32_t ecap;
  const uint32_t* spill_index;  // null unless some particle spills exclusions
  const SpillEntry* spill;
  const uint32_t* spill_excl;
  uint32_t max_neighbors;
  uint32_t* nbr_count;
  uint32_t* nbr;
  Ctrl* ctrl;
  int32_t half;  // keep only partners with a higher slot
  uint32_t stage_cap;  // candidate entries a workgroup may stage in LDS (set by launch_build)
};

struct ReorderParams {
  uint32_t n;
  // primary arrays
  real4* pos;
  real4* vel;
  int4* image;
  uint32_t* tag;
  uint32_t* molid;
  uint32_t* cell_of;
  uint32_t* topo_count;
  uint2* bond;
  uint4* angle;
  uint4* dih;
  uint32_t* excl;
  // scratch
  real4* s_pos;
  real4* s_vel;
  int4* s_image;
  uint32_t* s_tag;
  uint32_t* s_molid;
  uint32_t* s_cell_of;
  uint32_t* s_topo_count;
  uint2* s_bond;
  uint4* s_angle;
  uint4* s_dih;
  uint32_t* s_excl;
  uint32_t bcap, acap, dcap, ecap;
  uint32_t* spill_index;        // null when no particle spills
  uint32_t* s_spill_index;
  const SpillEntry* spill;
  uint2* spill_bond;
  uint4* spill_angle;
  uint4* spill_dih;
  uint32_t* spill_excl;
  uint32_t* cell_atoms;
  real4* cell_pos;
  uint32_t* new_of_old;
  Ctrl* ctrl;
};

}  // namespace beads_amd
