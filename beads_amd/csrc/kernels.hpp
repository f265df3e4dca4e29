// Host-side launchers for the gfx950 kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace beads_amd {

constexpr int kBlock = 256;

struct BinParams {
  uint32_t n;
  const real4* pos;
  BoxView box;
  CellGridView grid;
  uint32_t* cell_of;
  uint32_t* cell_count;
  Ctrl* ctrl;
  NeighborControl control;
};

struct CellScanParams {
  uint32_t n;
  uint32_t n_cells;
  const uint32_t* cell_of;
  uint32_t* cell_count;
  uint32_t* cell_start;
  uint32_t* cell_cursor;
  uint32_t* cell_atoms;
  const real4* pos;
  real4* cell_pos;  // written in cell order when no reorder follows
  uint4* cell_xbins; // per cell: x-bin prefix counts of its x-sorted members
  BoxView box;
  CellGridView grid;
  Ctrl* ctrl;
};

struct ScaleParams {
  uint32_t n;
  real4* vel;
  double* partials;
  Ctrl* ctrl;
};

inline unsigned grid_for(uint32_t n, int block = kBlock) {
  return static_cast<unsigned>((static_cast<uint64_t>(n) + block - 1) / block);
}

// The force pass gives every atom a team of kTeam adjacent lanes; each lane
// takes kTeamSlots neighbour slots per round. A workgroup covers kForceAtoms
// atoms.
constexpr int kTeam = 4;
constexpr int kTeamSlots = 8;
constexpr int kForceAtoms = kBlock / kTeam;

inline unsigned force_grid(uint32_t n) { return grid_for(n, kForceAtoms); }

// Entries allocated past the last atom in the build's candidate streams (the
// position array and its cell-ordered copy): a build round loads 16
// consecutive candidates without a bounds test and ignores those past the run.
constexpr uint32_t kCandidatePad = 16;

// Neighbour-rebuild sequence; every kernel is predicated on ctrl->rebuild so
// the host never waits for the displacement check.
// steps: steps advanced since the previous control launch (0 at set-up).
void launch_control(Ctrl* ctrl, NeighborControl control, bool forced, uint64_t steps,
                    hipStream_t stream);
void launch_bin(const BinParams& p, hipStream_t stream);
void launch_cell_scan(const CellScanParams& p, hipStream_t stream);
void launch_cell_scatter(const CellScanParams& p, hipStream_t stream);
void launch_cell_sort(const CellScanParams& p, hipStream_t stream);
void launch_reorder(const ReorderParams& p, hipStream_t stream);
void launch_build(const BuildParams& p, hipStream_t stream);

// Trajectory records in file order (slot order), one `words`-word record per
// atom: tag, then type, position, image, velocity, force as `mask` selects.
struct PackParams {
  uint32_t n;
  uint32_t mask;   // BEADS_TRAJ_* bits
  uint32_t words;  // record size in 32-bit words
  const real4* pos;
  const real4* vel;
  const real4* force;
  const int4* image;
  const uint32_t* tag;
  uint32_t* out;
};
void launch_pack_records(const PackParams& p, hipStream_t stream);

// One step's work.
void launch_force(const ForceParams& p, bool kick, bool observables, hipStream_t stream);
void launch_integrate(const IntegrateParams& p, bool thermostat, bool check,
                      bool force_rebuild, hipStream_t stream);
void launch_scale(const ScaleParams& p, bool thermostat, hipStream_t stream);

}  // namespace beads_amd
