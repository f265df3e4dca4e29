/*
 * FP64 host oracle for the BEADS per-step MD path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference's formulas on
 * the CPU so that tests, the smoke check and the CPU-baseline leg of the bench
 * can check the HIP engine. The product path never links, loads or calls it.
 *
 * Two evaluation modes (see oracle.cpp for the per-function citations):
 *  - discrete decisions (cell index, r^2 <= r_s^2 neighbour test, r^2 < r_c^2
 *    LJ test, wrap crossing count) are evaluated in IEEE binary32 with the
 *    reference's operation order and no FMA contraction, so neighbour sets and
 *    cell assignments can be compared bit-exactly with the device (the *_f64
 *    entry points evaluate them in binary64, as the BEADS_REAL64 build does);
 *  - values (forces, energies, virials, kinetic energy, Berendsen lambda,
 *    integrator updates) are evaluated in binary64.
 */
#ifndef BEADS_ORACLE_H
#define BEADS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PATH_DIRECT = 0, ORC_PATH_REGULAR = 1, ORC_PATH_SMALL = 2 };

typedef struct orc_plan {
  uint32_t nx, ny, nz, n_cells;
  float cell_size[3];
  float inv_cell[3];
  float force_cutoff;
  float search_cutoff;
  float search_cutoff_sq;
  int32_t path;
} orc_plan_t;

/* The plan of the BEADS_REAL64 build: cell sizes, search cutoff in binary64. */
typedef struct orc_plan_f64 {
  uint32_t nx, ny, nz, n_cells;
  double cell_size[3];
  double inv_cell[3];
  double force_cutoff;
  double search_cutoff;
  double search_cutoff_sq;
  int32_t path;
} orc_plan_f64_t;

/* A particle system in slot order. Positions/velocities are AoS [n][3]. */
typedef struct orc_system {
  uint32_t n;
  const float* pos;
  const float* vel;
  const float* mass;
  const int32_t* type;   /* 1-based */
  const uint32_t* tag;   /* permutation of 1..n */
  const int32_t* image;  /* [n][3] */
  float lo[3], hi[3];
  int32_t n_types;
  const float* lj_eps;   /* [n_types][n_types] symmetric */
  const float* lj_sig;
  float lj_cutoff;
  int32_t lj_shift;
  uint32_t n_bonds;
  const uint32_t* bonds;     /* [n_bonds][3] tag_i, tag_j, type */
  const float* bond_k;       /* indexed by type-1 */
  const float* bond_r0;
  uint32_t n_angles;
  const uint32_t* angles;    /* [n_angles][4] tag_i, tag_j, tag_k, type */
  const float* angle_k;
  const float* angle_theta0_deg;
  uint32_t n_dihedrals;
  const uint32_t* dihedrals; /* [n][5] tag_i, tag_j, tag_k, tag_l, type */
  const float* dihedral_k;
  const int32_t* dihedral_d;
  const int32_t* dihedral_n;
  int32_t exclusion_distance; /* 0 = no bonded exclusions */
} orc_system_t;

typedef struct orc_observables {
  double pair_pe, bond_pe, angle_pe, dihedral_pe;
  double pair_virial, bond_virial, angle_virial, dihedral_virial;
} orc_observables_t;

typedef struct orc_run_params {
  double dt;
  int32_t thermostat;     /* 0 none, 1 berendsen */
  double temperature;     /* public temperature units */
  double tau;
  double kb;              /* Boltzmann constant of the unit style */
  float cutoff_buffer;
  uint64_t rebuild_check_every;
  uint64_t sort_every_rebuild;
  uint64_t sort_every_steps; /* nonzero: sort at the first build >= this many steps after the last */
  uint32_t max_neighbors;
  uint64_t steps;
  int32_t threads;        /* 0 = OpenMP default */
} orc_run_params_t;

typedef struct orc_run_stats {
  uint64_t build_count;
  uint64_t dangerous_rebuild_count;
  double wall_seconds;     /* step loop only */
  double mean_neighbors;   /* mean stored neighbours per atom over builds */
  int32_t threads_used;
  double final_ke;
  orc_observables_t final_obs;
  uint64_t reorder_count;
} orc_run_stats_t;

/* neighbor_plan.cpp:182-236. Returns 0, or 1 with a message for an invalid plan. */
int orc_plan(const float lo[3], const float hi[3], float force_cutoff,
             float cutoff_buffer, uint32_t n, orc_plan_t* out);
const char* orc_last_error(void);

/* cell_list_build.cu:110-166 -- cell index per slot (binary32 decision). */
void orc_cell_ids(uint32_t n, const float* pos, const float lo[3],
                  const orc_plan_t* plan, uint32_t* cell);

/* Counting sort by cell, inside a cell by (x, old slot) when pos ([n][3]) is
 * given, else by old slot: order[new] = old. */
void orc_stable_cell_order(uint32_t n, const uint32_t* cell, uint32_t n_cells, const float* pos,
                           uint32_t* order, uint32_t* counts, uint32_t* offsets);

/* Bond-graph exclusions, exclusions.cu:514-598: all tag pairs (a<b) within
 * graph distance <= distance. Returns the pair count; writes at most cap pairs. */
uint64_t orc_exclusion_pairs(uint32_t n, uint32_t n_bonds, const uint32_t* bonds,
                             int32_t distance, uint32_t* pairs, uint64_t cap);

/* Full neighbour list (the three builders under simulation/neighbor/full).
 * path < 0 selects the plan's path.
 * list is atom-major [n][max_neighbors]. Returns 0 or 1 on plan error. */
int orc_neighbor_build(const orc_system_t* sys, float cutoff_buffer,
                       uint32_t max_neighbors, int32_t path, uint32_t* counts,
                       uint32_t* list, int32_t* overflow);

/* Forces (FP64 values) on binary32 positions. counts/list (atom-major) give
 * the stored neighbours; if they are NULL the oracle enumerates its own pairs
 * with a zero-buffer cell search. Pairs are kept if the binary32 r^2 < r_c^2
 * test passes. */
int orc_forces(const orc_system_t* sys, const uint32_t* counts, const uint32_t* list,
               uint32_t max_neighbors, double* forces, orc_observables_t* obs);

/* Same physics on binary64 positions (decisions in binary64); for finite
 * differences. pos64 is [n][3]. */
int orc_forces_f64(const orc_system_t* sys, const double* pos64, double* forces,
                   orc_observables_t* obs);

/* velocity_verlet.cu:59-111 and berendsen_thermostat.cu:104-163 in binary64. */
void orc_vv_pre(uint32_t n, const float* pos, const float* vel, const float* force,
                const float* mass, const int32_t* image, const float lo[3],
                const float hi[3], double dt, double* pos_out, double* vel_out,
                int32_t* image_out);
void orc_vv_post(uint32_t n, const float* vel, const float* force, const float* mass,
                 double dt, double* vel_out);
double orc_kinetic_energy(uint32_t n, const float* vel, const float* mass);
double orc_berendsen_lambda(double ke, uint32_t n, double kb, double temperature,
                            double dt, double tau);

/* rebuild_tracker.cu:80-92 + neighbor_system.cpp:213-216 in binary32: 1 if
 * any atom's unwrapped displacement from its reference exceeds half the
 * cutoff buffer (strict), else 0. pos/ref_pos [n][3], image/ref_image [n][3]. */
int orc_displacement_exceeds(uint32_t n, const float* pos, const int32_t* image,
                             const float* ref_pos, const int32_t* ref_image,
                             const float lo[3], const float hi[3], float cutoff_buffer);

/* A whole FP64 run (CPU baseline; NVE/NVT invariants). Final state in slot
 * order of the oracle's own reordering; tags_out gives the identity. */
int orc_run(const orc_system_t* sys, const orc_run_params_t* params,
            double* pos_out, double* vel_out, int32_t* image_out,
            uint32_t* tags_out, orc_run_stats_t* stats);

/* The BEADS_REAL64 build (real_t = double), for its parity tests: binary64
 * positions and velocities, and the box (hi - lo in binary64), cell grid,
 * search cutoff (force cutoff + buffer in binary64) and every decision (cell
 * index, r^2 <= r_s^2, r^2 < r_c^2, wrap, displacement test) in binary64 with
 * the reference's operation order. Box bounds and force-field parameters are
 * taken from sys (binary32 values, exact in binary64). */
int orc_plan_f64(const float lo[3], const float hi[3], double force_cutoff,
                 double cutoff_buffer, uint32_t n, orc_plan_f64_t* out);
void orc_cell_ids_f64(uint32_t n, const double* pos, const float lo[3],
                      const orc_plan_f64_t* plan, uint32_t* cell);
int orc_neighbor_build_f64(const orc_system_t* sys, const double* pos64, double cutoff_buffer,
                           uint32_t max_neighbors, int32_t path, uint32_t* counts,
                           uint32_t* list, int32_t* overflow);
/* Forces and observables at binary64 positions over the given lists
 * (atom-major), with the r^2 < r_c^2 test in binary64 on the binary64 box. */
int orc_forces_f64_list(const orc_system_t* sys, const double* pos64, const uint32_t* counts,
                        const uint32_t* list, uint32_t max_neighbors, double* forces,
                        orc_observables_t* obs);
/* orc_run from binary64 initial positions and velocities (pos_in, vel_in). */
int orc_run_f64(const orc_system_t* sys, const orc_run_params_t* params,
                const double* pos_in, const double* vel_in, double* pos_out,
                double* vel_out, int32_t* image_out, uint32_t* tags_out,
                orc_run_stats_t* stats);
/* orc_run_f64 with its tangent (forward-mode derivative) beside the state:
 * (dpos_in, dvel_in) is a perturbation of the initial state and
 * (dpos_out, dvel_out) its image under the linearised dynamics, in the slot
 * order of pos_out. LJ pairs and harmonic bonds under NVE only (else 1). */
int orc_run_tangent_f64(const orc_system_t* sys, const orc_run_params_t* params,
                        const double* pos_in, const double* vel_in, const double* dpos_in,
                        const double* dvel_in, double* pos_out, double* vel_out,
                        int32_t* image_out, uint32_t* tags_out, double* dpos_out,
                        double* dvel_out, orc_run_stats_t* stats);

#ifdef __cplusplus
}
#endif

#endif
