// Host engine: owns device state for one simulation and drives the gfx950
// kernels. Counterpart of the reference's SimulationRun together with the
// NeighborSystem, ForceFieldModel and DynamicsProgram it holds
// (native/src/simulation/simulation_run.hpp:20-64).
#pragma once

#include <chrono>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "models.hpp"
#include "output.hpp"

namespace beads_amd {

// unit_system.hpp:27-49, unit_system.cpp:9-40.
struct UnitSystem {
  std::string name;
  double kb;
  double pressure_scale;
};

UnitSystem unit_system_from(int32_t units);

// neighbor_plan.cpp:182-236: search cutoff, cell grid, path choice.
struct NeighborPlan {
  real_t force_cutoff = 0;
  real_t cutoff_buffer = 0;
  real_t search_cutoff = 0;
  real_t search_cutoff_sq = 0;
  uint32_t nx = 1, ny = 1, nz = 1, n_cells = 1;
  real_t cell_size[3] = {0, 0, 0};
  real_t inv_cell[3] = {0, 0, 0};
  int32_t path = 0;  // 0 direct, 1 regular cell, 2 small dimension
  uint32_t max_neighbors = 0;
  uint64_t check_every = 1;
  uint64_t sort_every = 10;
};

NeighborPlan make_neighbor_plan(const BoxView& box, real_t force_cutoff, real_t buffer,
                                uint32_t n, uint32_t max_neighbors, uint64_t check_every,
                                uint64_t sort_every);
BoxView make_box_view(const real_t* box_bound);

// Per-slot listed-term roles and exclusion partners built on the host from the
// tag-based topology (the device keeps them current across reorders).
struct HostTopology {
  uint32_t bcap = 0, acap = 0, dcap = 0, ecap = 0;  // inline capacities
  std::vector<uint32_t> count;  // packed inline counts per slot (8 bits each)
  std::vector<uint2> bond;      // [bcap][n]
  std::vector<uint4> angle;     // [acap][n]
  std::vector<uint4> dih;       // [dcap][n]
  std::vector<uint32_t> excl;   // [ecap][n]
  uint64_t excluded_pairs = 0;
  // Records past the inline capacities (all empty when nothing spills).
  std::vector<uint32_t> spill_index;  // [n], kNoSpill or an entry index
  std::vector<SpillEntry> spill;
  std::vector<uint2> spill_bond;
  std::vector<uint4> spill_angle;
  std::vector<uint4> spill_dih;
  std::vector<uint32_t> spill_excl;
  bool spills_exclusions = false;
};

// Inline capacity for one kind of per-slot record: the largest count when it
// is small; otherwise the smallest capacity that at most max(16, n / 256)
// particles exceed (those spill the rest), never above 255.
uint32_t inline_capacity(const std::vector<uint32_t>& counts);

// inline_only: every record inline (capacities = the largest counts, at most
// 255 per kind and particle), for the batched replicas, which have no spill.
HostTopology build_host_topology(const beads_spec_t& spec, bool inline_only = false);

// Shell validation of one spec (throws Error with the matching status).
void validate_spec(const beads_spec_t& spec);



// Bond-graph exclusion pairs (tag_a < tag_b) within the given distance
// (exclusions.cu:514-598).
std::vector<std::pair<uint32_t, uint32_t>> bond_graph_exclusions(uint32_t n, uint32_t n_bonds,
                                                                 const uint32_t* bonds,
                                                                 int32_t distance);

class Engine {
 public:
  explicit Engine(const beads_spec_t& spec);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  void initialize(bool observables);
  void advance(uint64_t n_steps, bool observables_on_last);
  void run();  // the whole simulation with outputs
  void synchronize();
  void stats(beads_engine_stats_t* out);
  void read_state(const beads_state_out_t& out);
  void read_neighbors(uint32_t* counts, uint32_t* list, uint64_t capacity);
  void read_cells(uint32_t* cell_of, uint32_t* cell_atoms, uint32_t* cell_start,
                  uint32_t* cell_count, uint64_t capacity);
  void profile_advance(uint64_t n_steps, bool per_phase, beads_phase_times_t* out);

  // Frame mode (force matching, trajectory backprop): positions come from the
  // caller in slot order, so particles are never reordered.
  void set_lj_params(const real_t* eps, const real_t* sigma);
  void set_bond_params(const real_t* k, const real_t* r0);
  void set_angle_params(const real_t* k, const real_t* theta0_deg);
  void set_dihedral_params(const real_t* k);
  // Entries of frame_backward's parameter gradient: eps, sigma [T*T] each,
  // bond k, r0 [nb] each, angle k, theta0 [na] each, dihedral k [nd].
  size_t adjoint_param_count() const;
  // async: enqueue and return without waiting (energies_out is then device
  // memory, and a list overflow is reported by the next synchronize()).
  void frame_forces(const real_t* positions_dev, real_t* forces_dev, double* energies_out,
                    bool reuse_list, bool async = false);
  void frame_load(const real_t* positions_dev, bool reuse_list, bool async = false);
  // noalloc: stage dL/dF in the library's static per-GPU scratch array instead
  // of an engine buffer and clear grad_params with a kernel; implies async.
  void frame_backward(const real_t* grad_forces_dev, double grad_energy,
                      real_t* grad_positions_dev, double* grad_params_dev, bool async = false,
                      bool noalloc = false);
  void frame_timing(double* forward_s, double* backward_s);
  // Enqueue frame-mode work on the caller's stream (nullptr: the engine's own).
  void set_stream(hipStream_t stream);
  uint32_t n_particles() const { return n_; }
  int32_t n_types() const { return n_types_; }
  int32_t n_bond_types() const { return n_bond_types_; }

 private:
  void upload_parameters();
  void enter_frame_mode();
  void load_frame(const real_t* positions_dev, bool reuse_list);
  void upload_state();
  void set_up_params();
  // steps: steps advanced since the previous rebuild sequence (0 at set-up).
  void rebuild_sequence(bool forced, uint64_t steps);
  void force_pass(bool kick, bool observables);
  void step_once(bool observables);
  uint64_t graph_steps() const;
  void replay_step_graph();
  void materialize_velocities();
  Ctrl fetch_ctrl();
  void check_overflow(const Ctrl& c);
  bool output_due(uint64_t step) const;
  void stage_outputs(uint64_t step);
  void write_thermo_row(uint64_t step, const Ctrl& c);
  void write_trajectory_frame(uint64_t step);
  size_t trajectory_record_bytes() const;
  void write_trajectory_records(uint64_t step, const unsigned char* records);
  void write_outputs(uint64_t step, bool thermo, bool trajectory, const Ctrl& c,
                     const unsigned char* records);
  void write_final_state();
  void log(const std::string& level, const std::string& message);
  void log_run_start();
  void log_run_end(uint64_t runsteps, double wall_s, const Ctrl& c);
  void poll_interrupt();

  // configuration
  beads_spec_t spec_{};
  std::string thermo_path_, trajectory_path_, final_state_path_, log_path_;
  UnitSystem units_;
  BoxView box_{};
  real_t spec_box_hi_[3] = {0, 0, 0};  // upper bounds exactly as given
  NeighborPlan plan_;
  uint32_t n_ = 0;
  bool thermostat_ = false;
  bool half_list_ = false;
  bool use_graphs_ = true;
  hipGraphExec_t step_graph_exec_ = nullptr;
  bool has_bond_style_ = false, has_angle_style_ = false, has_dihedral_style_ = false;
  uint64_t step_ = 0;
  bool initialized_ = false;
  bool ke_from_scale_ = false;  // which reduction last measured the stored velocities
  uint64_t excluded_pairs_ = 0;

  hipStream_t stream_ = nullptr;      // where work is enqueued: own_stream_ or the caller's
  hipStream_t own_stream_ = nullptr;

  // device state
  DeviceBuffer<real4> pos_, vel_, force_, ref_pos_, s_pos_, s_vel_;
  DeviceBuffer<real4> cell_pos_;  // positions in cell order, .w = slot bits
  DeviceBuffer<uint4> cell_xbins_;  // per-cell x-bin prefix counts
  DeviceBuffer<int4> image_, ref_image_, s_image_;
  DeviceBuffer<uint32_t> tag_, molid_, s_tag_, s_molid_;
  DeviceBuffer<uint32_t> nbr_count_, nbr_;
  DeviceBuffer<uint32_t> cell_of_, s_cell_of_, cell_count_, cell_start_, cell_cursor_, cell_atoms_;
  DeviceBuffer<uint32_t> new_of_old_;
  DeviceBuffer<uint32_t> topo_count_, s_topo_count_, excl_, s_excl_;
  // records past the inline capacities (empty when nothing spills)
  DeviceBuffer<uint32_t> spill_index_, s_spill_index_, spill_excl_;
  DeviceBuffer<SpillEntry> spill_;
  DeviceBuffer<uint2> spill_bond_;
  DeviceBuffer<uint4> spill_angle_, spill_dih_;
  bool has_spill_ = false, spill_excl_any_ = false;
  DeviceBuffer<uint2> bond_, s_bond_;
  DeviceBuffer<uint4> angle_, s_angle_, dih_, s_dih_;
  DeviceBuffer<double> partials_;
  DeviceBuffer<double> pair_acc_;  // half list: partner-side LJ sums [3 N]

  DeviceBuffer<Ctrl> ctrl_;
  PinnedBuffer<Ctrl> ctrl_host_;
  uint32_t bcap_ = 0, acap_ = 0, dcap_ = 0, ecap_ = 0;
  bool has_topology_ = false;

  // Force field, integrator and thermostat, created from the spec's style
  // names (models.hpp); they own their parameter tables.
  ModelSet models_;
  int32_t n_types_ = 0, n_bond_types_ = 0;
  DeviceBuffer<real4> grad_force4_;
  bool frame_mode_ = false;   // positions supplied per frame; never reordered
  bool frame_loaded_ = false;
  bool frame_list_valid_ = false;  // the stored list and its reference positions are current
  float last_forward_ms_ = 0.f, last_backward_ms_ = 0.f;
  bool forward_time_pending_ = false, backward_time_pending_ = false;  // async calls
  hipEvent_t frame_events_[4] = {nullptr, nullptr, nullptr, nullptr};

  // kernel parameter blocks
  ForceParams force_params_{};
  IntegrateParams integrate_params_{};
  BuildParams build_params_{};
  BinParams bin_params_{};
  CellScanParams scan_params_{};
  ReorderParams reorder_params_{};
  ScaleParams scale_params_{};
  NeighborControl control_{};

  // outputs: staged through output_ (pinned slots, copy stream, writer
  // thread) unless BEADS_AMD_SYNC_OUTPUT=1 asks for host round trips
  bool async_output_ = true;
  std::unique_ptr<AsyncOutput> output_;
  std::ofstream thermo_stream_;
  std::ofstream trajectory_stream_;
  std::ofstream log_stream_;
  bool log_wrote_any_ = false;
  bool thermo_header_logged_ = false;
  uint64_t interrupt_interval_ = 1;
  std::chrono::steady_clock::time_point run_start_;
};

}  // namespace beads_amd
