// extern "C" entry points declared in include/beads_amd.h. Every entry point
// catches all exceptions and converts them to a status plus a thread-local
// message (the reference lets pybind11 translate C++ exceptions instead,
// native/src/bindings/module.cpp:36-55).
#include <cmath>
#include <exception>
#include <new>
#include <string>

#include "beads_amd.h"
#include "common.hpp"
#include "batch.hpp"
#include "engine.hpp"

struct beads_engine {
  beads_amd::Engine* impl;
};

struct beads_batch {
  beads_amd::Batch* impl;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
beads_status_t guarded(F&& f) {
  try {
    f();
    return BEADS_OK;
  } catch (const beads_amd::Error& e) {
    g_last_error = e.what();
    return e.status();
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return BEADS_E_VALUE;
  } catch (const std::overflow_error& e) {
    g_last_error = e.what();
    return BEADS_E_OVERFLOW;
  } catch (const std::out_of_range& e) {
    g_last_error = e.what();
    return BEADS_E_INDEX;
  } catch (const std::bad_alloc&) {
    g_last_error = "host memory allocation failed";
    return BEADS_E_RUNTIME;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BEADS_E_RUNTIME;
  } catch (...) {
    g_last_error = "unknown native error";
    return BEADS_E_RUNTIME;
  }
}

beads_amd::Engine& engine_of(beads_engine_t* e) {
  if (e == nullptr || e->impl == nullptr) {
    beads_amd::fail(BEADS_E_VALUE, "engine handle is null");
  }
  return *e->impl;
}

beads_amd::Batch& batch_of(beads_batch_t* b) {
  if (b == nullptr || b->impl == nullptr) {
    beads_amd::fail(BEADS_E_VALUE, "batch handle is null");
  }
  return *b->impl;
}

}  // namespace

extern "C" {

const char* beads_last_error(void) { return g_last_error.c_str(); }

int beads_abi_version(void) { return BEADS_AMD_ABI_VERSION; }

beads_status_t beads_build_info(beads_build_info_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "build_info output is null");
    // module.cpp:19-21: derived from real_t itself.
    out->real_precision = beads_amd::kRealPrecision;
    out->real_dtype = beads_amd::kRealDtype;
    out->type_dtype = "int32";
    out->index_dtype = "uint32";
    out->image_dtype = "int32";
    out->runstep_dtype = "uint64";
    out->device_arch = "gfx950";
  });
}

beads_status_t beads_execute_simulation(const beads_spec_t* spec) {
  return guarded([&] {
    beads_amd::require(spec != nullptr, "spec is null");
    beads_amd::Engine engine(*spec);
    engine.run();
  });
}

beads_status_t beads_engine_create(const beads_spec_t* spec, beads_engine_t** out) {
  return guarded([&] {
    beads_amd::require(spec != nullptr && out != nullptr, "spec/out is null");
    *out = nullptr;
    auto* impl = new beads_amd::Engine(*spec);
    *out = new beads_engine{impl};
  });
}

beads_status_t beads_engine_initialize(beads_engine_t* e, int32_t with_observables) {
  return guarded([&] { engine_of(e).initialize(with_observables != 0); });
}

beads_status_t beads_engine_advance(beads_engine_t* e, uint64_t n, int32_t obs_last) {
  return guarded([&] { engine_of(e).advance(n, obs_last != 0); });
}

beads_status_t beads_engine_synchronize(beads_engine_t* e) {
  return guarded([&] { engine_of(e).synchronize(); });
}

beads_status_t beads_engine_stats(beads_engine_t* e, beads_engine_stats_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "stats output is null");
    engine_of(e).stats(out);
  });
}

beads_status_t beads_engine_read_state(beads_engine_t* e, const beads_state_out_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "state output is null");
    engine_of(e).read_state(*out);
  });
}

beads_status_t beads_engine_read_neighbors(beads_engine_t* e, uint32_t* counts, uint32_t* list,
                                           uint64_t capacity) {
  return guarded([&] {
    beads_amd::require(counts != nullptr && list != nullptr, "neighbor outputs are null");
    engine_of(e).read_neighbors(counts, list, capacity);
  });
}

beads_status_t beads_engine_read_cells(beads_engine_t* e, uint32_t* cell_of, uint32_t* cell_atoms,
                                       uint32_t* cell_start, uint32_t* cell_count,
                                       uint64_t capacity) {
  return guarded([&] {
    beads_amd::require(cell_of && cell_atoms && cell_start && cell_count, "cell outputs are null");
    engine_of(e).read_cells(cell_of, cell_atoms, cell_start, cell_count, capacity);
  });
}

beads_status_t beads_engine_profile_advance(beads_engine_t* e, uint64_t n, int32_t per_phase,
                                            beads_phase_times_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "profile output is null");
    engine_of(e).profile_advance(n, per_phase != 0, out);
  });
}

beads_status_t beads_engine_set_lj_params(beads_engine_t* e, const beads_real_t* epsilon,
                                          const beads_real_t* sigma) {
  return guarded([&] {
    beads_amd::require(epsilon && sigma, "lj parameter arrays are null", BEADS_E_TYPE);
    engine_of(e).set_lj_params(epsilon, sigma);
  });
}

beads_status_t beads_engine_set_bond_params(beads_engine_t* e, const beads_real_t* k, const beads_real_t* r0) {
  return guarded([&] {
    beads_amd::Engine& eng = engine_of(e);
    beads_amd::require(eng.n_bond_types() == 0 || (k && r0), "bond parameter arrays are null",
                       BEADS_E_TYPE);
    eng.set_bond_params(k, r0);
  });
}

beads_status_t beads_engine_set_angle_params(beads_engine_t* e, const beads_real_t* k,
                                             const beads_real_t* theta0_deg) {
  return guarded([&] {
    beads_amd::require(k && theta0_deg, "angle parameter arrays are null", BEADS_E_TYPE);
    engine_of(e).set_angle_params(k, theta0_deg);
  });
}

beads_status_t beads_engine_set_dihedral_params(beads_engine_t* e, const beads_real_t* k) {
  return guarded([&] {
    beads_amd::require(k != nullptr, "dihedral parameter array is null", BEADS_E_TYPE);
    engine_of(e).set_dihedral_params(k);
  });
}

beads_status_t beads_engine_adjoint_param_count(beads_engine_t* e, uint64_t* count) {
  return guarded([&] {
    beads_amd::require(count != nullptr, "count output is null", BEADS_E_TYPE);
    *count = engine_of(e).adjoint_param_count();
  });
}

beads_status_t beads_engine_frame_forces(beads_engine_t* e, const beads_real_t* positions, beads_real_t* forces,
                                         double* energies, int32_t reuse_list) {
  return guarded([&] {
    beads_amd::require(positions && forces, "frame arrays are null", BEADS_E_TYPE);
    engine_of(e).frame_forces(positions, forces, energies, reuse_list != 0);
  });
}

beads_status_t beads_engine_frame_load(beads_engine_t* e, const beads_real_t* positions,
                                       int32_t reuse_list) {
  return guarded([&] {
    beads_amd::require(positions != nullptr, "frame positions are null", BEADS_E_TYPE);
    engine_of(e).frame_load(positions, reuse_list != 0);
  });
}

beads_status_t beads_engine_frame_backward(beads_engine_t* e, const beads_real_t* grad_forces,
                                           double grad_energy, beads_real_t* grad_positions,
                                           double* grad_params) {
  return guarded([&] {
    beads_amd::require(grad_forces && grad_positions && grad_params, "adjoint arrays are null",
                       BEADS_E_TYPE);
    beads_amd::require(std::isfinite(grad_energy), "grad_energy must be finite");
    engine_of(e).frame_backward(grad_forces, grad_energy, grad_positions, grad_params);
  });
}

beads_status_t beads_engine_frame_timing(beads_engine_t* e, double* forward_s,
                                         double* backward_s) {
  return guarded([&] { engine_of(e).frame_timing(forward_s, backward_s); });
}

beads_status_t beads_engine_set_stream(beads_engine_t* e, void* stream) {
  return guarded([&] { engine_of(e).set_stream(static_cast<hipStream_t>(stream)); });
}

beads_status_t beads_engine_frame_forces_async(beads_engine_t* e, const beads_real_t* positions,
                                               beads_real_t* forces, double* energies,
                                               int32_t reuse_list) {
  return guarded([&] {
    beads_amd::require(positions && forces, "frame arrays are null", BEADS_E_TYPE);
    engine_of(e).frame_forces(positions, forces, energies, reuse_list != 0, true);
  });
}

beads_status_t beads_engine_frame_load_async(beads_engine_t* e, const beads_real_t* positions,
                                             int32_t reuse_list) {
  return guarded([&] {
    beads_amd::require(positions != nullptr, "frame positions are null", BEADS_E_TYPE);
    engine_of(e).frame_load(positions, reuse_list != 0, true);
  });
}

beads_status_t beads_engine_frame_backward_async(beads_engine_t* e, const beads_real_t* grad_forces,
                                                 double grad_energy, beads_real_t* grad_positions,
                                                 double* grad_params) {
  return guarded([&] {
    beads_amd::require(grad_forces && grad_positions && grad_params, "adjoint arrays are null",
                       BEADS_E_TYPE);
    beads_amd::require(std::isfinite(grad_energy), "grad_energy must be finite");
    engine_of(e).frame_backward(grad_forces, grad_energy, grad_positions, grad_params, true);
  });
}

beads_status_t beads_engine_frame_backward_noalloc(beads_engine_t* e, const beads_real_t* grad_forces,
                                                   double grad_energy, beads_real_t* grad_positions,
                                                   double* grad_params) {
  return guarded([&] {
    beads_amd::require(grad_forces && grad_positions && grad_params, "adjoint arrays are null",
                       BEADS_E_TYPE);
    beads_amd::require(std::isfinite(grad_energy), "grad_energy must be finite");
    engine_of(e).frame_backward(grad_forces, grad_energy, grad_positions, grad_params, true,
                                true);
  });
}

void beads_engine_destroy(beads_engine_t* e) {
  if (e != nullptr) {
    delete e->impl;
    delete e;
  }
}

beads_status_t beads_batch_create(const beads_spec_t* specs, uint32_t count, beads_batch_t** out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "batch output is null");
    *out = nullptr;
    auto* impl = new beads_amd::Batch(specs, count);
    *out = new beads_batch{impl};
  });
}

beads_status_t beads_batch_config_info(uint32_t index, beads_batch_config_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "config output is null");
    *out = beads_amd::Batch::config_info(index);
  });
}

beads_status_t beads_batch_autotune(beads_batch_t* b, uint32_t trial_steps, int32_t* chosen,
                                    double* seconds) {
  return guarded([&] {
    const int config = batch_of(b).autotune(trial_steps, seconds);
    if (chosen != nullptr) *chosen = config;
  });
}

beads_status_t beads_batch_set_config(beads_batch_t* b, uint32_t index) {
  return guarded([&] { batch_of(b).set_config(index); });
}

beads_status_t beads_batch_get_config(beads_batch_t* b, uint32_t* index) {
  return guarded([&] {
    beads_amd::require(index != nullptr, "config output is null");
    *index = batch_of(b).config();
  });
}

beads_status_t beads_batch_set_block(beads_batch_t* b, int32_t block) {
  return guarded([&] {
    beads_amd::require(block == 256 || block == 512 || block == 1024,
                       "batch workgroup size must be 256, 512 or 1024.");
    for (uint32_t k = 0; k < beads_amd::Batch::kConfigs; ++k) {
      const beads_batch_config_t c = beads_amd::Batch::config_info(k);
      if (c.mode == 0 && c.block == block) batch_of(b).set_config(k);
    }
  });
}

beads_status_t beads_batch_run(beads_batch_t* b, uint64_t n, int32_t obs_last) {
  return guarded([&] { batch_of(b).run(n, obs_last != 0); });
}

beads_status_t beads_batch_profile(beads_batch_t* b, uint64_t n, double* seconds) {
  return guarded([&] {
    beads_amd::require(seconds != nullptr, "profile output is null");
    *seconds = batch_of(b).profile(n);
  });
}

beads_status_t beads_batch_synchronize(beads_batch_t* b) {
  return guarded([&] { batch_of(b).synchronize(); });
}

beads_status_t beads_batch_replica_stats(beads_batch_t* b, uint32_t r, beads_engine_stats_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "stats output is null");
    batch_of(b).replica_stats(r, out);
  });
}

beads_status_t beads_batch_read_state(beads_batch_t* b, uint32_t r, const beads_state_out_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "state output is null");
    batch_of(b).read_state(r, *out);
  });
}

void beads_batch_destroy(beads_batch_t* b) {
  if (b != nullptr) {
    delete b->impl;
    delete b;
  }
}

}  // extern "C"
