// Batched replicas: device descriptors and the persistent-launch entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace beads_amd {

// One replica's geometry, parameters and offsets into the batch's arrays.
struct ReplicaDesc {
  uint32_t n;             // atoms
  uint32_t atom0;         // first atom in the concatenated per-atom arrays
  uint32_t cell0;         // first cell entry (each replica owns n_cells + 1)
  uint32_t max_neighbors;
  uint64_t nbr0;          // first neighbour-index entry (chunks of kNbrChunk, see replica_nbr_index)
  BoxView box;
  CellGridView grid;
  real_t rs2, rc2, thr2, dt;
  real_t cutoff_buffer;
  int32_t berendsen;
  int32_t n_types;
  uint32_t lj0;           // first LJ table entry (n_types^2 entries)
  double target_thermal;  // kB * T0
  double dt_over_tau;
  double dof;
  uint64_t check_every;
  uint64_t sort_every;
  uint64_t sort_every_steps;  // nonzero: step-based sort cadence (see NeighborControl)
  uint32_t chunk0;            // first chunk of the replica's atoms (phased mode)
  uint32_t n_chunks;          // ceil(n / kReplicaChunk)
  // Listed terms and exclusions: the replica's per-slot records start at
  // cap * atom0 in each kind's array (stride n, partner slots local to the
  // replica); its types index the coefficient tables from these offsets.
  int32_t has_listed;         // some bond, angle or dihedral record
  int32_t has_excl;           // some excluded pair
  uint32_t bond_t0, angle_t0, dih_t0;
};

// Phased mode: the replicas' atoms in chunks of kReplicaChunk (one wave's
// worth), each inside one replica.
constexpr uint32_t kReplicaChunk = 64;
// Per-chunk partial sums: kinetic energy, pair/bond/angle/dihedral energies
// and virials.
constexpr uint32_t kReplicaSums = 9;

// Per-replica run state carried across launches.
struct ReplicaState {
  uint64_t step;
  uint64_t build_count;
  uint64_t reorder_count;
  uint64_t dangerous;
  uint32_t checked;
  uint32_t overflow;
  uint32_t initialized;
  uint32_t pad;
  double lambda;  // Berendsen factor still to be applied to the stored velocities
  double ke;      // kinetic energy after the last second half-kick
  double pe[4];   // pair, bond, angle, dihedral energy (last observable step)
  double virial[4];
  uint64_t last_sort_step;  // step of the last reorder (step-based cadence)
  // Phased mode, within a step: displacement check result, and the decisions
  // the rebuild kernels follow.
  uint32_t exceeded;
  uint32_t rebuild;
  uint32_t reorder;
  uint32_t pad2;
};

struct ReplicaArrays {
  real4* pos;
  real4* vel;
  real4* force;
  int4* image;
  uint32_t* tag;
  real4* ref_pos;
  int4* ref_image;
  uint32_t* cell_of;
  uint32_t* cell_atoms;
  uint32_t* nbr_count;
  uint32_t* nbr;
  uint32_t* cell_count;
  uint32_t* cell_start;
  uint32_t* cell_cursor;
  real4* s_pos;
  real4* s_vel;
  int4* s_image;
  uint32_t* s_tag;
  uint32_t* s_cell_of;
  uint32_t* cell_tmp;  // phased mode: members of each cell before the slot sort
  real4* cell_pos;     // phased mode: positions in cell order, .w = the slot's bits
  // Per-slot listed-term roles and exclusions (the single-system layout of
  // device.hpp; inline only) and their reorder staging.
  uint32_t* topo_count;
  uint2* bond;
  uint4* angle;
  uint4* dih;
  uint32_t* excl;
  uint32_t* s_topo_count;
  uint2* s_bond;
  uint4* s_angle;
  uint4* s_dih;
  uint32_t* s_excl;
  uint32_t* new_of_old;  // reorder: new local slot of each old local slot
};

struct ReplicaParams {
  const ReplicaDesc* desc;
  ReplicaState* state;
  const LjCoeff* lj;
  ReplicaArrays arrays;
  uint32_t bcap, acap, dcap, ecap;  // records per slot of each kind
  const real2* bond_coeff;          // k, r0
  const real2* angle_coeff;         // k, theta0 (rad)
  const real4* dih_coeff;           // k, d, n, unused
  // Phased mode.
  uint32_t count;                // replicas
  uint32_t total_chunks;
  const uint32_t* chunk_replica; // replica of each chunk
  double* partials;              // kReplicaSums x total_chunks: kinetic energy, energies[4],
                                 // virials[4]
  uint32_t* work;                // chunks of the replicas that rebuild this step
  uint32_t* work_reorder;        // chunks of the replicas that also reorder
  uint32_t* work_count;          // [0]: entries of work, [1]: of work_reorder
};

// Phased mode force layout: kReplicaChunk atoms per workgroup, team lanes per
// atom, list slots per lane per round.
struct PhasedShape {
  int team;
  int slots;
};

// One workgroup per replica, all n_steps inside the launch.
void launch_replicas(const ReplicaParams& p, uint32_t count, int block, uint64_t n_steps,
                     bool observables_on_last, hipStream_t stream);

// Phased mode: one step of every replica as a sequence of batched kernels
// (initial: the first build and step-0 forces instead). Captured into a graph
// by the caller.
bool phased_shape_supported(PhasedShape shape);
void launch_phased_step(const ReplicaParams& p, PhasedShape shape, bool initial, bool observables,
                        hipStream_t stream);

}  // namespace beads_amd
