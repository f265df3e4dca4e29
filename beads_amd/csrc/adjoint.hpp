// Adjoint (force matching / trajectory backprop) kernels. See adjoint.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace beads_amd {

constexpr int kMaxBondTypesAdjoint = 64;
constexpr int kMaxAngleTypesAdjoint = 64;
constexpr int kMaxDihedralTypesAdjoint = 64;

// Largest system frame_backward's allocation-free variant serves: its dL/dF
// staging lives in a static device array of this many real4 (16 MiB per GPU,
// 32 MiB in the BEADS_REAL64 build).
constexpr uint32_t kNoAllocMaxAtoms = 1u << 20;

struct AdjointParams {
  uint32_t n;
  const real4* pos;        // positions, .w = type bits
  const real4* g;          // dL/dF per atom
  const uint32_t* nbr_count;
  const uint32_t* nbr;      // chunked full list
  BoxView box;
  real_t rc2;               // cutoff decision in real_t, as in the forward
  double rc2_f64;           // FP64 rc^2 for the energy-shift derivative
  int32_t lj_shift;
  int32_t n_types;
  const double* eps;        // [T*T], symmetric
  const double* sigma;      // [T*T], symmetric
  const uint32_t* topo_count;
  const uint2* bond;
  uint32_t bcap;
  const uint32_t* spill_index;  // null when no particle spills
  const SpillEntry* spill;
  const uint2* spill_bond;
  int32_t n_bond_types;
  const double* bond_k;
  const double* bond_r0;
  // Harmonic angles and dihedrals (per-slot role records as in the force
  // pass; acap/dcap 0 when the system has none).
  const uint4* angle;
  const uint4* dih;
  uint32_t acap, dcap;
  const uint4* spill_angle;
  const uint4* spill_dih;
  int32_t n_angle_types;
  const double* angle_k;
  const double* angle_theta0;  // radians
  int32_t n_dihedral_types;
  const double* dih_k;
  const double* dih_d;
  const double* dih_n;
  double grad_energy;       // dL/dE for the total potential energy
  real_t* grad_pos;          // [n][3]
  // eps[T*T], sigma[T*T], bond k[nb], bond r0[nb], angle k[na], angle
  // theta0[na] (per radian), dihedral k[nd]; zeroed by the caller
  double* grad_params;
};

// Caller positions [n][3] (unwrapped allowed) -> wrapped pos (type kept in .w)
// and image counts. With check set, an atom farther than the half skin from its
// position at the last build raises ctrl->rebuild (integrate_kernel's test).
struct FrameLoadParams {
  uint32_t n;
  const real_t* xyz;
  real4* pos;
  int4* image;
  const real4* ref_pos;
  const int4* ref_image;
  BoxView box;
  real_t threshold_sq;
  Ctrl* ctrl;
};

void launch_frame_load(const FrameLoadParams& p, bool check, hipStream_t stream);
void launch_frame_store(uint32_t n, const real4* force, real_t* xyz, hipStream_t stream);
void launch_pack_xyz(uint32_t n, const real_t* xyz, real4* out, hipStream_t stream);
// out[0, count) = 0 with a kernel (no memset node when a caller captures the stream).
void launch_fill_f64(double* out, uint32_t count, hipStream_t stream);
// This GPU's static staging array (kNoAllocMaxAtoms real4), address resolved
// once per device; allocates nothing.
real4* scratch_address();
void launch_fm_backward(const AdjointParams& p, hipStream_t stream);

}  // namespace beads_amd
