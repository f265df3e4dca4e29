// Host side of the batched replicas: validates and packs many systems (LJ
// pairs, harmonic bonds, angles and dihedrals, bonded exclusions) into one
// set of device arrays, picks the execution configuration with a small
// autotuner, and advances every replica either with one persistent launch per
// run or with one graph of batched kernels per step (phased mode), each
// captured once into a hipGraph and replayed.
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

#include "engine.hpp"
#include "models.hpp"

namespace beads_amd {

Batch::Batch(const beads_spec_t* specs, uint32_t count) : count_(count) {
  require(specs != nullptr && count > 0, "batch needs at least one replica spec.");
  register_builtin_models();
  std::vector<NeighborPlan> plans(count);
  std::vector<ModelSet> models(count);
  for (uint32_t r = 0; r < count; ++r) {
    const beads_spec_t& s = specs[r];
    validate_spec(s);
    models[r] = create_models(s, ModelContext{unit_system_from(s.units).kb});
    require(s.n_types <= kMaxLdsTypes, "batched replicas support at most 16 particle types.");
    require(s.half_list == 0, "batched replicas use the full neighbour list.");
    plans[r] = make_neighbor_plan(make_box_view(s.box_bound), s.lj_cutoff, s.cutoff_buffer,
                                  s.n_particles, s.max_neighbors, s.rebuild_check_every,
                                  s.sort_every_rebuild);
  }
  // Dispatch order: largest replicas first, so the long-running workgroups
  // start in the first wave of the launch.
  order_.resize(count);
  std::iota(order_.begin(), order_.end(), 0u);
  std::stable_sort(order_.begin(), order_.end(), [&](uint32_t a, uint32_t b) {
    return specs[a].n_particles > specs[b].n_particles;
  });
  position_of_.resize(count);
  for (uint32_t k = 0; k < count; ++k) position_of_[order_[k]] = k;

  // Per-slot records of each replica (inline only) and the batch's
  // capacities: the largest of any replica, so every replica's records keep
  // their indices at cap * atom0 + k * n + i.
  std::vector<HostTopology> topo(count);
  HostTopology caps;
  for (uint32_t k = 0; k < count; ++k) {
    const beads_spec_t& s = specs[order_[k]];
    if (s.n_bonds + s.n_angles + s.n_dihedrals == 0) continue;
    topo[k] = build_host_topology(s, true);
    caps.bcap = std::max(caps.bcap, topo[k].bcap);
    caps.acap = std::max(caps.acap, topo[k].acap);
    caps.dcap = std::max(caps.dcap, topo[k].dcap);
    caps.ecap = std::max(caps.ecap, topo[k].ecap);
  }

  desc_host_.resize(count);
  uint64_t atoms = 0, cells = 0, nbr = 0, lj = 0, chunks = 0;
  std::vector<uint32_t> chunk_replica;
  std::vector<real2> bond_coeff, angle_coeff;
  std::vector<real4> dih_coeff;
  for (uint32_t k = 0; k < count; ++k) {
    const uint32_t r = order_[k];
    const beads_spec_t& s = specs[r];
    const NeighborPlan& p = plans[r];
    const UnitSystem units = unit_system_from(s.units);
    ReplicaDesc d{};
    d.n = s.n_particles;
    d.atom0 = static_cast<uint32_t>(atoms);
    d.cell0 = static_cast<uint32_t>(cells);
    d.max_neighbors = p.max_neighbors;
    d.nbr0 = nbr;
    d.box = make_box_view(s.box_bound);
    d.grid.nx = p.nx;
    d.grid.ny = p.ny;
    d.grid.nz = p.nz;
    d.grid.n_cells = p.n_cells;
    for (int a = 0; a < 3; ++a) d.grid.inv_cell[a] = p.inv_cell[a];
    d.grid.path = p.path;
    d.rs2 = p.search_cutoff_sq;
    d.rc2 = p.force_cutoff * p.force_cutoff;
    const real_t half = real_t(0.5) * p.cutoff_buffer;
    d.thr2 = half * half;
    d.dt = s.dt;
    d.cutoff_buffer = p.cutoff_buffer;
    d.berendsen = s.berendsen ? 1 : 0;
    d.n_types = s.n_types;
    d.lj0 = static_cast<uint32_t>(lj);
    d.target_thermal = units.kb * static_cast<double>(s.berendsen_temperature);
    d.dt_over_tau = s.berendsen ? static_cast<double>(s.dt) / s.berendsen_tau : 0.0;
    d.dof = std::max(0.0, 3.0 * s.n_particles - 3.0);
    d.check_every = p.check_every;
    d.sort_every = p.sort_every;
    d.sort_every_steps = s.sort_every_steps;
    d.chunk0 = static_cast<uint32_t>(chunks);
    d.n_chunks = (s.n_particles + kReplicaChunk - 1) / kReplicaChunk;
    d.has_listed = s.n_bonds + s.n_angles + s.n_dihedrals > 0 ? 1 : 0;
    d.has_excl = topo[k].excluded_pairs > 0 ? 1 : 0;
    d.bond_t0 = static_cast<uint32_t>(bond_coeff.size());
    d.angle_t0 = static_cast<uint32_t>(angle_coeff.size());
    d.dih_t0 = static_cast<uint32_t>(dih_coeff.size());
    const ModelSet& m = models[r];
    if (m.bond) {
      for (const real4& c : m.bond->table()) bond_coeff.push_back(make_real2(c.x, c.y));
    }
    if (m.angle) {
      for (const real4& c : m.angle->table()) angle_coeff.push_back(make_real2(c.x, c.y));
    }
    if (m.dihedral) {
      for (const real4& c : m.dihedral->table()) dih_coeff.push_back(c);
    }
    chunks += d.n_chunks;
    chunk_replica.insert(chunk_replica.end(), d.n_chunks, k);
    desc_host_[k] = d;
    atoms += s.n_particles;
    cells += static_cast<uint64_t>(p.n_cells) + 1;
    // Lists in chunks of kNbrChunk entries per atom.
    nbr += static_cast<uint64_t>((p.max_neighbors + kNbrChunk - 1) / kNbrChunk) * kNbrChunk *
           s.n_particles;
    lj += static_cast<uint64_t>(s.n_types) * s.n_types;
  }
  require(atoms <= 0xffffffffull && cells <= 0xffffffffull,
          "batch exceeds 2^32 atoms or cells.", BEADS_E_OVERFLOW);
  total_atoms_ = atoms;
  total_chunks_ = static_cast<uint32_t>(chunks);

  // Host images of the initial state, kept to restore after autotuning.
  init_pos_.resize(atoms);
  init_vel_.resize(atoms);
  init_img_.resize(atoms);
  init_tag_.resize(atoms);
  init_cell_atoms_.resize(atoms);
  std::vector<LjCoeff> lj_all;
  lj_all.reserve(lj);
  for (uint32_t k = 0; k < count; ++k) {
    const beads_spec_t& s = specs[order_[k]];
    const uint32_t a0 = desc_host_[k].atom0;
    for (uint32_t i = 0; i < s.n_particles; ++i) {
      const real_t tw = real_of_id(static_cast<uint32_t>(s.types[i] - 1));
      init_pos_[a0 + i] = make_real4(s.positions[3 * i], s.positions[3 * i + 1],
                                     s.positions[3 * i + 2], tw);
      init_vel_[a0 + i] = make_real4(s.velocities[3 * i], s.velocities[3 * i + 1],
                                     s.velocities[3 * i + 2], s.masses[i]);
      init_img_[a0 + i] = make_int4(s.images[3 * i], s.images[3 * i + 1], s.images[3 * i + 2], 0);
      init_tag_[a0 + i] = s.tags[i];
      init_cell_atoms_[a0 + i] = i;
    }
    const std::vector<LjCoeff> t = pack_lj_table(s);
    lj_all.insert(lj_all.end(), t.begin(), t.end());
  }
  // Records in the batch layout (zeros past a replica's own capacities).
  init_topo_count_.assign(atoms, 0u);
  init_bond_.assign(caps.bcap * atoms, uint2{0, 0});
  init_angle_.assign(caps.acap * atoms, uint4{0, 0, 0, 0});
  init_dih_.assign(caps.dcap * atoms, uint4{0, 0, 0, 0});
  init_excl_.assign(caps.ecap * atoms, 0u);
  for (uint32_t k = 0; k < count; ++k) {
    const HostTopology& h = topo[k];
    if (h.count.empty()) continue;
    const size_t a0 = desc_host_[k].atom0;
    std::copy(h.count.begin(), h.count.end(), init_topo_count_.begin() + a0);
    std::copy(h.bond.begin(), h.bond.end(), init_bond_.begin() + caps.bcap * a0);
    std::copy(h.angle.begin(), h.angle.end(), init_angle_.begin() + caps.acap * a0);
    std::copy(h.dih.begin(), h.dih.end(), init_dih_.begin() + caps.dcap * a0);
    std::copy(h.excl.begin(), h.excl.end(), init_excl_.begin() + caps.ecap * a0);
  }

  BEADS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  desc_.upload(desc_host_.data(), count, stream_);
  lj_.upload(lj_all.data(), lj_all.size(), stream_);
  bond_coeff_.upload(bond_coeff.data(), bond_coeff.size(), stream_);
  angle_coeff_.upload(angle_coeff.data(), angle_coeff.size(), stream_);
  dih_coeff_.upload(dih_coeff.data(), dih_coeff.size(), stream_);
  s_topo_count_.resize(atoms);
  s_bond_.resize(init_bond_.size());
  s_angle_.resize(init_angle_.size());
  s_dih_.resize(init_dih_.size());
  s_excl_.resize(init_excl_.size());
  new_of_old_.resize(atoms);
  const size_t n = atoms;
  force_.resize(n);
  ref_pos_.resize(n);
  ref_image_.resize(n);
  cell_of_.resize(n);
  nbr_count_.resize(n);
  nbr_.resize(nbr);
  cell_count_.resize(cells);
  cell_start_.resize(cells);
  cell_cursor_.resize(cells);
  s_pos_.resize(n);
  s_vel_.resize(n);
  s_image_.resize(n);
  s_tag_.resize(n);
  s_cell_of_.resize(n);
  cell_tmp_.resize(n);
  cell_pos_.resize(n);
  chunk_replica_.upload(chunk_replica.data(), chunk_replica.size(), stream_);
  work_.resize(chunks);
  work_reorder_.resize(chunks);
  work_count_.resize(2);
  partials_.resize(kReplicaSums * chunks);
  state_.resize(count);
  reset_state();

  ReplicaArrays& A = params_.arrays;
  A.pos = pos_.data();
  A.vel = vel_.data();
  A.force = force_.data();
  A.image = image_.data();
  A.tag = tag_.data();
  A.ref_pos = ref_pos_.data();
  A.ref_image = ref_image_.data();
  A.cell_of = cell_of_.data();
  A.cell_atoms = cell_atoms_.data();
  A.nbr_count = nbr_count_.data();
  A.nbr = nbr_.data();
  A.cell_count = cell_count_.data();
  A.cell_start = cell_start_.data();
  A.cell_cursor = cell_cursor_.data();
  A.s_pos = s_pos_.data();
  A.s_vel = s_vel_.data();
  A.s_image = s_image_.data();
  A.s_tag = s_tag_.data();
  A.s_cell_of = s_cell_of_.data();
  A.cell_tmp = cell_tmp_.data();
  A.cell_pos = cell_pos_.data();
  A.topo_count = topo_count_.data();
  A.bond = bond_.data();
  A.angle = angle_.data();
  A.dih = dih_.data();
  A.excl = excl_.data();
  A.s_topo_count = s_topo_count_.data();
  A.s_bond = s_bond_.data();
  A.s_angle = s_angle_.data();
  A.s_dih = s_dih_.data();
  A.s_excl = s_excl_.data();
  A.new_of_old = new_of_old_.data();
  params_.bcap = caps.bcap;
  params_.acap = caps.acap;
  params_.dcap = caps.dcap;
  params_.ecap = caps.ecap;
  params_.bond_coeff = bond_coeff_.data();
  params_.angle_coeff = angle_coeff_.data();
  params_.dih_coeff = dih_coeff_.data();
  params_.count = count;
  params_.total_chunks = total_chunks_;
  params_.chunk_replica = chunk_replica_.data();
  params_.partials = partials_.data();
  params_.work = work_.data();
  params_.work_reorder = work_reorder_.data();
  params_.work_count = work_count_.data();
  params_.desc = desc_.data();
  params_.state = state_.data();
  params_.lj = lj_.data();
  BEADS_HIP(hipStreamSynchronize(stream_));
}

Batch::~Batch() {
  for (auto& g : graphs_) {
    (void)hipGraphExecDestroy(g.exec);
  }
  if (stream_ != nullptr) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

void Batch::reset_state() {
  const size_t n = total_atoms_;
  pos_.upload(init_pos_.data(), n, stream_);
  vel_.upload(init_vel_.data(), n, stream_);
  image_.upload(init_img_.data(), n, stream_);
  tag_.upload(init_tag_.data(), n, stream_);
  cell_atoms_.upload(init_cell_atoms_.data(), n, stream_);
  topo_count_.upload(init_topo_count_.data(), init_topo_count_.size(), stream_);
  bond_.upload(init_bond_.data(), init_bond_.size(), stream_);
  angle_.upload(init_angle_.data(), init_angle_.size(), stream_);
  dih_.upload(init_dih_.data(), init_dih_.size(), stream_);
  excl_.upload(init_excl_.data(), init_excl_.size(), stream_);
  force_.zero(stream_);
  nbr_count_.zero(stream_);
  cell_count_.zero(stream_);
  cell_start_.zero(stream_);
  cell_of_.zero(stream_);
  work_count_.zero(stream_);
  initialized_ = false;
  std::vector<ReplicaState> st(count_);
  for (auto& s : st) {
    std::memset(&s, 0, sizeof(s));
    s.lambda = 1.0;
  }
  state_.upload(st.data(), count_, stream_);
}

static_assert(Batch::kConfigs == BEADS_BATCH_CONFIGS, "configuration table size");

beads_batch_config_t Batch::config_info(uint32_t index) {
  static const beads_batch_config_t table[kConfigs] = {
      {0, 256, 4, 4}, {0, 512, 4, 4}, {0, 1024, 4, 4},  // persistent
      {1, 256, 4, 4}, {1, 128, 2, 8}, {1, 512, 8, 2},   // phased
      {1, 256, 4, 8}, {1, 512, 8, 4}, {1, 128, 2, 4}};
  require(index < kConfigs, "batch configuration index out of range.", BEADS_E_INDEX);
  return table[index];
}

void Batch::set_config(uint32_t index) {
  (void)config_info(index);
  config_ = index;
}

hipGraphExec_t Batch::graph_for(uint32_t config, uint64_t steps, bool obs) {
  for (const auto& g : graphs_) {
    if (g.config == config && g.steps == steps && g.obs == obs) return g.exec;
  }
  const beads_batch_config_t c = config_info(config);
  hipGraph_t graph;
  BEADS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  if (c.mode == 0) {
    launch_replicas(params_, count_, c.block, steps, obs, stream_);
  } else {
    launch_phased_step(params_, PhasedShape{c.team, c.slots}, steps == 0, obs, stream_);
  }
  BEADS_HIP(hipStreamEndCapture(stream_, &graph));
  hipGraphExec_t exec;
  BEADS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  BEADS_HIP(hipGraphDestroy(graph));
  graphs_.push_back({config, steps, obs, exec});
  return exec;
}

void Batch::run(uint64_t steps, bool observables_on_last) {
  const beads_batch_config_t c = config_info(config_);
  if (c.mode == 0) {
    // steps == 0 still launches: the first launch performs every replica's
    // initial build and step-0 forces.
    BEADS_HIP(hipGraphLaunch(graph_for(config_, steps, observables_on_last), stream_));
  } else {
    if (!initialized_) {
      BEADS_HIP(hipGraphLaunch(graph_for(config_, 0, false), stream_));
    }
    for (uint64_t s = 0; s < steps; ++s) {
      const bool obs = observables_on_last && s + 1 == steps;
      BEADS_HIP(hipGraphLaunch(graph_for(config_, 1, obs), stream_));
    }
  }
  initialized_ = true;
  steps_done_ += steps;
}

void Batch::synchronize() { BEADS_HIP(hipStreamSynchronize(stream_)); }

double Batch::profile(uint64_t steps) {
  hipEvent_t a, b;
  BEADS_HIP(hipEventCreate(&a));
  BEADS_HIP(hipEventCreate(&b));
  BEADS_HIP(hipEventRecord(a, stream_));
  run(steps, false);
  BEADS_HIP(hipEventRecord(b, stream_));
  BEADS_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  BEADS_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms * 1e-3;
}

// Autotuner: time each configuration on the real batch from the initial
// state, keep the fastest, then restore the initial state.
int Batch::autotune(uint32_t trial_steps, double* seconds_out) {
  double best = 0.0;
  uint32_t best_config = config_;
  for (uint32_t k = 0; k < kConfigs; ++k) {
    reset_state();
    config_ = k;
    run(1, false);  // initial build + forces and the graphs' first replay, outside the timing
    const double t = profile(trial_steps);
    if (seconds_out != nullptr) seconds_out[k] = t;
    if (k == 0 || t < best) {
      best = t;
      best_config = k;
    }
  }
  config_ = best_config;
  reset_state();
  steps_done_ = 0;
  synchronize();
  return static_cast<int>(config_);
}

void Batch::replica_stats(uint32_t r, beads_engine_stats_t* out) {
  require(r < count_, "replica index out of range.", BEADS_E_INDEX);
  const uint32_t k = position_of_[r];
  ReplicaState st{};
  BEADS_HIP(hipMemcpyAsync(&st, state_.data() + k, sizeof(st), hipMemcpyDeviceToHost, stream_));
  synchronize();
  const ReplicaDesc& d = desc_host_[k];
  std::memset(out, 0, sizeof(*out));
  out->step = st.step;
  out->build_count = st.build_count;
  out->dangerous_rebuild_count = st.dangerous;
  out->reorder_count = st.reorder_count;
  out->overflow = static_cast<int32_t>(st.overflow);
  out->path = d.grid.path;
  out->nx = d.grid.nx;
  out->ny = d.grid.ny;
  out->nz = d.grid.nz;
  out->max_neighbors = d.max_neighbors;
  out->ke = st.ke;
  out->lambda = st.lambda;
  for (int q = 0; q < 4; ++q) {
    out->pe[q] = st.pe[q];
    out->virial[q] = st.virial[q];
  }
  out->temperature = d.dof > 0 ? 2.0 * st.ke / d.dof : 0.0;  // in kB units
}

void Batch::read_state(uint32_t r, const beads_state_out_t& o) {
  require(r < count_, "replica index out of range.", BEADS_E_INDEX);
  const ReplicaDesc& d = desc_host_[position_of_[r]];
  const size_t n = d.n, a0 = d.atom0;
  std::vector<real4> pos(n), vel(n), force(n);
  std::vector<int4> img(n);
  BEADS_HIP(hipMemcpyAsync(pos.data(), pos_.data() + a0, n * sizeof(real4), hipMemcpyDeviceToHost, stream_));
  BEADS_HIP(hipMemcpyAsync(vel.data(), vel_.data() + a0, n * sizeof(real4), hipMemcpyDeviceToHost, stream_));
  BEADS_HIP(hipMemcpyAsync(force.data(), force_.data() + a0, n * sizeof(real4), hipMemcpyDeviceToHost, stream_));
  BEADS_HIP(hipMemcpyAsync(img.data(), image_.data() + a0, n * sizeof(int4), hipMemcpyDeviceToHost, stream_));
  if (o.tags) {
    BEADS_HIP(hipMemcpyAsync(o.tags, tag_.data() + a0, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
  }
  synchronize();
  for (size_t i = 0; i < n; ++i) {
    if (o.positions) {
      o.positions[3 * i] = pos[i].x;
      o.positions[3 * i + 1] = pos[i].y;
      o.positions[3 * i + 2] = pos[i].z;
    }
    if (o.types) o.types[i] = static_cast<int32_t>(id_of(pos[i].w)) + 1;
    if (o.velocities) {
      o.velocities[3 * i] = vel[i].x;
      o.velocities[3 * i + 1] = vel[i].y;
      o.velocities[3 * i + 2] = vel[i].z;
    }
    if (o.masses) o.masses[i] = vel[i].w;
    if (o.forces) {
      o.forces[3 * i] = force[i].x;
      o.forces[3 * i + 1] = force[i].y;
      o.forces[3 * i + 2] = force[i].z;
    }
    if (o.images) {
      o.images[3 * i] = img[i].x;
      o.images[3 * i + 1] = img[i].y;
      o.images[3 * i + 2] = img[i].z;
    }
  }
}

}  // namespace beads_amd
