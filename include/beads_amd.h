/*
 * beads_amd C ABI -- the boundary between the Python front end and the
 * MI355X (gfx950) engine. Plain C types only: host pointers, sizes, status
 * codes. No exceptions cross this boundary; every entry point returns a
 * beads_status_t and leaves a thread-local message for beads_last_error().
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   beads_build_info        <- native/src/bindings/module.cpp:15-27 (_native.build_info)
 *   beads_execute_simulation<- native/src/bindings/module.cpp:36-44 (_native.execute_simulation),
 *                              i.e. parse_native_spec + SimulationRun::execute
 *                              (native/src/simulation/simulation_run.cpp:237-244)
 *   beads_engine_*          <- SimulationRun's construction, initialize_force_state and
 *                              advance_dynamics_step (simulation_run.cpp:60-133, 212-235),
 *                              split so tests and the bench can drive and inspect a run
 *   beads_spec_t            <- input::SimulationSpec (native/src/input/native_spec.hpp:16-176)
 *   beads_real_t            <- beads::real_t (native/include/beads/core/types.hpp:12-16)
 *
 * Precision is chosen at build time, like the reference's BEADS_REAL64 switch
 * (CMakeLists.txt:29-31): libbeads_amd.so stores positions, velocities,
 * forces and every real-valued parameter in binary32; libbeads_amd_real64.so
 * is built with BEADS_REAL64 defined and stores them in binary64. A program
 * that links or loads the binary64 library compiles this header with
 * BEADS_REAL64 defined; beads_build_info reports which precision a library
 * was built with.
 */
#ifndef BEADS_AMD_H
#define BEADS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEADS_AMD_ABI_VERSION 6

#ifdef BEADS_REAL64
typedef double beads_real_t;
#else
typedef float beads_real_t;
#endif

typedef enum beads_status {
  BEADS_OK = 0,
  BEADS_E_VALUE = 1,       /* -> ValueError  (std::invalid_argument in the reference) */
  BEADS_E_TYPE = 2,        /* -> TypeError */
  BEADS_E_OVERFLOW = 3,    /* -> OverflowError (std::overflow_error) */
  BEADS_E_INDEX = 4,       /* -> IndexError (std::out_of_range) */
  BEADS_E_RUNTIME = 5,     /* -> RuntimeError (std::runtime_error / logic_error) */
  BEADS_E_INTERRUPTED = 6, /* the interrupt callback asked to stop (KeyboardInterrupt) */
  BEADS_E_DEVICE = 7       /* HIP runtime error -> RuntimeError */
} beads_status_t;

typedef enum beads_units { BEADS_UNITS_REDUCED = 0, BEADS_UNITS_NM_KJMOL = 1 } beads_units_t;

typedef enum beads_log_echo {
  BEADS_LOG_SCREEN = 0,
  BEADS_LOG_FILE = 1,
  BEADS_LOG_BOTH = 2,
  BEADS_LOG_NONE = 3
} beads_log_echo_t;

/* Resolved bonded-exclusion policy (native_spec_parser.cpp:730-753); reported
 * in the run-start log. */
typedef enum beads_exclusion_policy {
  BEADS_EXCLUSION_NONE = 0,
  BEADS_EXCLUSION_DEFAULT = 1,
  BEADS_EXCLUSION_EXPLICIT = 2
} beads_exclusion_policy_t;

/* Trajectory field bits (trajectory_output.cu:39-46). */
enum {
  BEADS_TRAJ_TAG = 1u << 0,
  BEADS_TRAJ_TYPE = 1u << 1,
  BEADS_TRAJ_POSITION = 1u << 2,
  BEADS_TRAJ_IMAGE = 1u << 3,
  BEADS_TRAJ_VELOCITY = 1u << 4,
  BEADS_TRAJ_FORCE = 1u << 5
};

/* Returns 0 to continue, nonzero to stop the run with BEADS_E_INTERRUPTED. */
typedef int (*beads_interrupt_fn)(void* context);

/*
 * One simulation, fully typed. Arrays are host memory borrowed for the
 * duration of the call (the engine copies them in). Shapes in brackets.
 */
typedef struct beads_spec {
  /* system */
  int32_t units;                 /* beads_units_t */
  uint32_t n_particles;
  const beads_real_t* box_bound; /* [2][3] lower row, upper row */
  const beads_real_t* positions; /* [n][3] */
  const beads_real_t* velocities;/* [n][3] */
  const beads_real_t* masses;    /* [n] */
  const int32_t* types;          /* [n], dense 1-based */
  const uint32_t* tags;          /* [n], permutation of 1..n */
  const uint32_t* molecule_ids;  /* [n] */
  const int32_t* images;         /* [n][3] */
  uint32_t n_bonds;
  const uint32_t* bonds;         /* [n_bonds][3]: tag_i, tag_j, type */
  uint32_t n_angles;
  const uint32_t* angles;        /* [n_angles][4]: tag_i, tag_j, tag_k, type */
  uint32_t n_dihedrals;
  const uint32_t* dihedrals;     /* [n_dihedrals][5]: tag_i..tag_l, type */

  /* Style names (NULL: the default) select each model from its registry
   * (pair_factory.cpp:30-47 and the bond/angle/dihedral factories): pair
   * "lj" (default); bond, angle, dihedral "harmonic" (default when the
   * kind's type count is positive, else none); integrator "velocity_verlet"
   * (default); thermostat "berendsen" (default when berendsen is set, else
   * none). An unregistered name is rejected with BEADS_E_VALUE. The
   * parameter blocks below are those of the built-in styles. */
  const char* pair_style;
  const char* bond_style;
  const char* angle_style;
  const char* dihedral_style;
  const char* integrator_style;
  const char* thermostat_style;

  /* pair style "lj" */
  int32_t n_types;
  beads_real_t lj_cutoff;
  int32_t lj_shift;
  const beads_real_t* lj_epsilon; /* [n_types][n_types], symmetric */
  const beads_real_t* lj_sigma;  /* [n_types][n_types], symmetric */

  /* listed styles "harmonic"; a count of 0 means the style is absent */
  int32_t n_bond_types;
  const beads_real_t* bond_k;    /* [n_bond_types] */
  const beads_real_t* bond_r0;
  int32_t n_angle_types;
  const beads_real_t* angle_k;
  const beads_real_t* angle_theta0_deg;
  int32_t n_dihedral_types;
  const beads_real_t* dihedral_k;
  const int32_t* dihedral_d;     /* +1 or -1 */
  const int32_t* dihedral_n;     /* >= 0 */
  int32_t exclusion_distance;    /* 0 none, else 1..3 */
  int32_t exclusion_policy;      /* beads_exclusion_policy_t */

  /* dynamics "velocity_verlet" + optional thermostat "berendsen" */
  beads_real_t dt;
  int32_t berendsen;
  beads_real_t berendsen_temperature;
  beads_real_t berendsen_tau;

  /* neighbor */
  beads_real_t cutoff_buffer;
  uint64_t rebuild_check_every;
  uint64_t sort_every_rebuild;
  uint64_t sort_every_steps;     /* 0: sort every sort_every_rebuild-th build (the reference's
                                    rule); else sort at the first build at least this many
                                    steps after the last sort (no reference counterpart) */
  uint32_t max_neighbors;
  int32_t half_list;             /* 0: full list (the reference's); 1: each pair stored once,
                                    j > i, reactions added with FP64 atomics (no reference
                                    counterpart; LJ only, listed terms unchanged) */

  /* outputs (NULL path = disabled) */
  const char* thermo_path;
  uint64_t thermo_every;
  const char* trajectory_path;
  uint64_t trajectory_every;
  uint32_t trajectory_fields;    /* BEADS_TRAJ_* mask, TAG always set */
  const char* final_state_path;
  int32_t log_echo;              /* beads_log_echo_t */
  const char* log_path;

  uint64_t runsteps;

  beads_interrupt_fn interrupt;  /* may be NULL */
  void* interrupt_context;
} beads_spec_t;

typedef struct beads_build_info {
  const char* real_precision;  /* "single", or "double" in the BEADS_REAL64 build */
  const char* real_dtype;      /* "float32" / "float64" */
  const char* type_dtype;      /* "int32" */
  const char* index_dtype;     /* "uint32" */
  const char* image_dtype;     /* "int32" */
  const char* runstep_dtype;   /* "uint64" */
  const char* device_arch;     /* "gfx950" */
} beads_build_info_t;

const char* beads_last_error(void);
int beads_abi_version(void);
beads_status_t beads_build_info(beads_build_info_t* out);

/* Validate, run and write outputs: the whole of _native.execute_simulation. */
beads_status_t beads_execute_simulation(const beads_spec_t* spec);

/* ---- engine handle: a run that tests and the bench can step and inspect ---- */
typedef struct beads_engine beads_engine_t;

typedef struct beads_engine_stats {
  uint64_t step;                 /* physical steps advanced */
  uint64_t build_count;          /* neighbour builds, initial one included */
  uint64_t dangerous_rebuild_count;
  uint64_t reorder_count;
  int32_t overflow;              /* neighbour-list overflow seen */
  int32_t path;                  /* 0 direct, 1 regular cell, 2 small dimension */
  uint32_t nx, ny, nz;           /* cell grid */
  uint32_t max_neighbors;
  double mean_neighbors;         /* sum(neighbor_count)/N of the current list */
  double ke;                     /* kinetic energy after the last completed step */
  double lambda;                 /* last Berendsen factor */
  double pe[4];                  /* pair, bond, angle, dihedral (last observable step) */
  double virial[4];
  double temperature;
  double pressure;
} beads_engine_stats_t;

/* Host-side views filled by beads_engine_read_state (each pointer may be NULL). */
typedef struct beads_state_out {
  beads_real_t* positions;       /* [n][3] slot order */
  beads_real_t* velocities;      /* [n][3] */
  beads_real_t* forces;          /* [n][3] */
  int32_t* images;               /* [n][3] */
  uint32_t* tags;                /* [n] */
  int32_t* types;                /* [n] */
  uint32_t* molecule_ids;        /* [n] */
  beads_real_t* masses;          /* [n] */
} beads_state_out_t;

beads_status_t beads_engine_create(const beads_spec_t* spec, beads_engine_t** out);
/* Initial neighbour build and step-0 forces (simulation_run.cpp:113-119). */
beads_status_t beads_engine_initialize(beads_engine_t* engine, int32_t with_observables);
/* Advance n steps; observables are computed on the last one if requested. */
beads_status_t beads_engine_advance(beads_engine_t* engine, uint64_t n_steps,
                                    int32_t observables_on_last);
beads_status_t beads_engine_synchronize(beads_engine_t* engine);
beads_status_t beads_engine_stats(beads_engine_t* engine, beads_engine_stats_t* out);
beads_status_t beads_engine_read_state(beads_engine_t* engine, const beads_state_out_t* out);
/* counts[n]; list is slot-major [max_neighbors][n] (neighbor_list.cuh:12-19). */
beads_status_t beads_engine_read_neighbors(beads_engine_t* engine, uint32_t* counts,
                                           uint32_t* list, uint64_t list_capacity);
/* cell_of[n], cell_atoms[n] (atoms in cell order), cell_start/cell_count[n_cells]. */
beads_status_t beads_engine_read_cells(beads_engine_t* engine, uint32_t* cell_of,
                                       uint32_t* cell_atoms, uint32_t* cell_start,
                                       uint32_t* cell_count, uint64_t n_cells_capacity);
/* Device time of n steps measured with HIP events on the engine's stream.
 * Per-phase times are summed over the steps (events around each launch group). */
typedef struct beads_phase_times {
  uint64_t steps;
  double total_s;       /* whole advance, events at both ends */
  double integrate_s;   /* Berendsen factor + first half-kick + drift + wrap + check */
  double rebuild_s;     /* control + binning + sort + reorder + neighbour build */
  double force_s;       /* fused force + second half-kick + kinetic energy */
  uint64_t force_launches;
} beads_phase_times_t;
beads_status_t beads_engine_profile_advance(beads_engine_t* engine, uint64_t n_steps,
                                            int32_t per_phase, beads_phase_times_t* out);

/* ---- frame mode: forces and their adjoint at caller-supplied positions ----
 * No reference counterpart: the reference computes forces only inside its
 * dynamics loop (simulation_run.cpp:212-235) and has no gradients. Frame mode
 * serves force matching and trajectory backprop. Positions are given in slot
 * order (the order of the spec's arrays) and are never reordered; an engine
 * used in frame mode cannot also run dynamics. All [n][3] arrays are binary32
 * device memory; every call synchronises the engine's stream before returning.
 *
 * Parameter updates take host arrays: epsilon/sigma [T][T] (symmetric for the
 * adjoint), bond k/r0 [n_bond_types], angle k and theta0 in degrees
 * [n_angle_types], dihedral k [n_dihedral_types] (its d and n stay the spec's). */
beads_status_t beads_engine_set_lj_params(beads_engine_t* engine, const beads_real_t* epsilon,
                                          const beads_real_t* sigma);
beads_status_t beads_engine_set_bond_params(beads_engine_t* engine, const beads_real_t* k,
                                            const beads_real_t* r0);
beads_status_t beads_engine_set_angle_params(beads_engine_t* engine, const beads_real_t* k,
                                             const beads_real_t* theta0_deg);
beads_status_t beads_engine_set_dihedral_params(beads_engine_t* engine, const beads_real_t* k);
/* Number of doubles frame_backward writes to grad_params:
 * 2*T*T + 2*n_bond_types + 2*n_angle_types + n_dihedral_types. */
beads_status_t beads_engine_adjoint_param_count(beads_engine_t* engine, uint64_t* count);
/* Wrap positions into the box, update the neighbour list, compute forces.
 * reuse_list = 0 rebuilds the list at every call; reuse_list = 1 keeps the
 * previous frame's list while every atom stays within half the skin of its
 * position at the last build (the dynamics loop's test,
 * neighbor_system.cpp:107-146). energies (host, may be NULL): pair, bond,
 * angle, dihedral. */
beads_status_t beads_engine_frame_forces(beads_engine_t* engine, const beads_real_t* positions,
                                         beads_real_t* forces, double* energies,
                                         int32_t reuse_list);
/* Load positions and update the neighbour list exactly as frame_forces does,
 * without evaluating forces: re-establishes a frame for frame_backward after
 * other frames were evaluated. */
beads_status_t beads_engine_frame_load(beads_engine_t* engine, const beads_real_t* positions,
                                       int32_t reuse_list);
/* Adjoint of the last frame_forces (or frame_load) call for L with dL/dF = grad_forces and
 * dL/dE = grad_energy (E = total potential energy). Writes dL/dx [n][3] and
 * FP64 parameter gradients (device) laid out as
 * [epsilon T*T | sigma T*T | bond k nb | bond r0 nb | angle k na |
 *  angle theta0 na (per radian) | dihedral k nd]
 * (beads_engine_adjoint_param_count entries). LJ pairs, harmonic bonds,
 * harmonic angles and harmonic dihedrals. */
beads_status_t beads_engine_frame_backward(beads_engine_t* engine, const beads_real_t* grad_forces,
                                           double grad_energy, beads_real_t* grad_positions,
                                           double* grad_params);
/* Device time of the last frame_forces / frame_backward (HIP events); after
 * the async variants this waits for their end events. */
beads_status_t beads_engine_frame_timing(beads_engine_t* engine, double* forward_s,
                                         double* backward_s);
/* Stream-ordered frame mode (no reference counterpart; for callers that keep
 * tensors on their own stream, e.g. torch's current stream). set_stream puts
 * the engine in frame mode and makes later calls enqueue on `stream` (a
 * hipStream_t taken as given, NULL being HIP's null stream; until the first
 * call the engine uses a non-blocking stream of its own), after the work
 * already enqueued has completed. The _async variants take the same arguments as their
 * synchronous counterparts but return without waiting: `energies` is then
 * device memory (4 doubles, written in stream order), and a neighbour-list
 * overflow is reported by the next beads_engine_synchronize. */
beads_status_t beads_engine_set_stream(beads_engine_t* engine, void* stream);
beads_status_t beads_engine_frame_forces_async(beads_engine_t* engine,
                                               const beads_real_t* positions,
                                               beads_real_t* forces, double* energies,
                                               int32_t reuse_list);
beads_status_t beads_engine_frame_load_async(beads_engine_t* engine,
                                             const beads_real_t* positions, int32_t reuse_list);
beads_status_t beads_engine_frame_backward_async(beads_engine_t* engine,
                                                 const beads_real_t* grad_forces,
                                                 double grad_energy, beads_real_t* grad_positions,
                                                 double* grad_params);
/* Allocation-free variant of beads_engine_frame_backward_async. No counterpart
 * in the reference; it is for callers that cannot allocate device memory once
 * set up: code under HIP graph capture (a captured training step), tight
 * serving loops, harnesses that forbid hipMalloc. Same arguments, meaning and
 * stream order as the async call, but dL/dF is staged in the library's static
 * per-GPU scratch array (16 MiB, 32 MiB in the BEADS_REAL64 build; at most
 * 1,048,576 particles; a larger engine gets BEADS_E_VALUE) instead of an
 * engine buffer allocated on first use, grad_params is cleared by a kernel
 * rather than a memset, and no timing events are recorded
 * (beads_engine_frame_timing does not cover these calls).
 * The scratch keeps the last call's contents and is shared by every engine on
 * that GPU: calls that overlap on one GPU (engines on concurrent streams) race. */
beads_status_t beads_engine_frame_backward_noalloc(beads_engine_t* engine,
                                                   const beads_real_t* grad_forces,
                                                   double grad_energy,
                                                   beads_real_t* grad_positions,
                                                   double* grad_params);
void beads_engine_destroy(beads_engine_t* engine);

/* ---- batched replicas ----
 * No reference counterpart: the reference runs one system per process
 * (simulation_run.cpp). Many independent systems (LJ pairs, harmonic bonds,
 * angles and dihedrals, bonded exclusions), each a full beads_spec_t (outputs
 * ignored), advance together in one of two modes:
 *   persistent (mode 0): ONE launch per run, one workgroup of `block` lanes
 *     per replica running that replica's whole step loop (integrate, rebuild
 *     with stable sort, full-list LJ + listed terms + kick + Berendsen);
 *   phased (mode 1): per step, a sequence of batched kernels over 64-atom
 *     chunks of every replica, so a large replica spreads over many
 *     workgroups; the force kernel uses `team` lanes per atom and `slots` list
 *     slots per lane per round (`block` = 64 * team).
 * Both make the same decisions with the same arithmetic. Each launch sequence
 * is captured once into a hipGraph and replayed. */
typedef struct beads_batch beads_batch_t;

typedef struct {
  int32_t mode;  /* 0 persistent, 1 phased */
  int32_t block; /* lanes per workgroup (phased: of the force kernel) */
  int32_t team;  /* force kernel: lanes per atom */
  int32_t slots; /* force kernel: list slots per lane per round */
} beads_batch_config_t;

#define BEADS_BATCH_CONFIGS 9

beads_status_t beads_batch_create(const beads_spec_t* specs, uint32_t count,
                                  beads_batch_t** out);
/* Configuration index (0 .. BEADS_BATCH_CONFIGS - 1) described. */
beads_status_t beads_batch_config_info(uint32_t index, beads_batch_config_t* out);
/* Times each configuration for trial_steps from the initial state, keeps the
 * fastest (its index in chosen_config) and restores the initial state.
 * seconds may be NULL or hold BEADS_BATCH_CONFIGS doubles. */
beads_status_t beads_batch_autotune(beads_batch_t* batch, uint32_t trial_steps,
                                    int32_t* chosen_config, double* seconds);
beads_status_t beads_batch_set_config(beads_batch_t* batch, uint32_t index);
beads_status_t beads_batch_get_config(beads_batch_t* batch, uint32_t* index);
/* The persistent configuration with this workgroup size (256, 512 or 1024). */
beads_status_t beads_batch_set_block(beads_batch_t* batch, int32_t block);
/* The first call also performs every replica's initial build and forces. */
beads_status_t beads_batch_run(beads_batch_t* batch, uint64_t n_steps,
                               int32_t observables_on_last);
beads_status_t beads_batch_profile(beads_batch_t* batch, uint64_t n_steps, double* seconds);
beads_status_t beads_batch_synchronize(beads_batch_t* batch);
/* stats of replica r (index in the order given to create): build/reorder
 * counts, kinetic energy, pending lambda, and the pair, bond, angle and
 * dihedral energies and virials in pe[0..3] and virial[0..3] (last observable
 * step), as the single engine reports them. */
beads_status_t beads_batch_replica_stats(beads_batch_t* batch, uint32_t r,
                                         beads_engine_stats_t* out);
beads_status_t beads_batch_read_state(beads_batch_t* batch, uint32_t r,
                                      const beads_state_out_t* out);
void beads_batch_destroy(beads_batch_t* batch);

#ifdef __cplusplus
}
#endif

#endif
