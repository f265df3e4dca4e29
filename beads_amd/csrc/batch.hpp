// Batched replicas (host side). See batch.cpp.
#pragma once

#include <vector>

#include "common.hpp"
#include "replicas.hpp"

namespace beads_amd {

class Batch {
 public:
  Batch(const beads_spec_t* specs, uint32_t count);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // Execution configurations (persistent workgroup sizes, phased force team
  // shapes), by index; see beads_batch_config_t.
  static constexpr uint32_t kConfigs = 9;
  static beads_batch_config_t config_info(uint32_t index);
  int autotune(uint32_t trial_steps, double* seconds_out);
  void run(uint64_t steps, bool observables_on_last);
  void synchronize();
  double profile(uint64_t steps);
  void replica_stats(uint32_t r, beads_engine_stats_t* out);
  void read_state(uint32_t r, const beads_state_out_t& out);
  uint32_t count() const { return count_; }
  uint64_t total_atoms() const { return total_atoms_; }
  uint32_t config() const { return config_; }
  void set_config(uint32_t index);

 private:
  struct CachedGraph {
    uint32_t config;
    uint64_t steps;  // persistent: steps per launch; phased: 1, or 0 for the initial step
    bool obs;
    hipGraphExec_t exec;
  };
  void reset_state();
  hipGraphExec_t graph_for(uint32_t config, uint64_t steps, bool obs);

  uint32_t count_ = 0;
  uint64_t total_atoms_ = 0;
  uint64_t steps_done_ = 0;
  uint32_t config_ = 0;
  bool initialized_ = false;  // every replica has its first build and forces
  uint32_t total_chunks_ = 0;
  std::vector<uint32_t> order_, position_of_;
  std::vector<ReplicaDesc> desc_host_;
  std::vector<real4> init_pos_, init_vel_;
  std::vector<int4> init_img_;
  std::vector<uint32_t> init_tag_, init_cell_atoms_;
  std::vector<uint32_t> init_topo_count_, init_excl_;
  std::vector<uint2> init_bond_;
  std::vector<uint4> init_angle_, init_dih_;
  std::vector<CachedGraph> graphs_;
  hipStream_t stream_ = nullptr;

  DeviceBuffer<ReplicaDesc> desc_;
  DeviceBuffer<ReplicaState> state_;
  DeviceBuffer<LjCoeff> lj_;
  DeviceBuffer<real4> pos_, vel_, force_, ref_pos_, s_pos_, s_vel_, cell_pos_;
  DeviceBuffer<int4> image_, ref_image_, s_image_;
  DeviceBuffer<uint32_t> tag_, s_tag_, cell_of_, s_cell_of_, cell_atoms_, nbr_count_, nbr_;
  DeviceBuffer<uint32_t> cell_count_, cell_start_, cell_cursor_, cell_tmp_;
  DeviceBuffer<uint32_t> chunk_replica_, work_, work_reorder_, work_count_;
  DeviceBuffer<double> partials_;
  DeviceBuffer<uint32_t> topo_count_, s_topo_count_, excl_, s_excl_, new_of_old_;
  DeviceBuffer<uint2> bond_, s_bond_;
  DeviceBuffer<uint4> angle_, s_angle_, dih_, s_dih_;
  DeviceBuffer<real2> bond_coeff_, angle_coeff_;
  DeviceBuffer<real4> dih_coeff_;
  ReplicaParams params_{};
};

}  // namespace beads_amd
