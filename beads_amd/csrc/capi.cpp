// extern "C" entry points declared in include/beads_amd.h. Every entry point
// catches all exceptions and converts them to a status plus a thread-local
// message (the reference lets pybind11 translate C++ exceptions instead,
// native/src/bindings/module.cpp:36-55).
#include <exception>
#include <new>
#include <string>

#include "beads_amd.h"
#include "common.hpp"
#include "engine.hpp"

struct beads_engine {
  beads_amd::Engine* impl;
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

}  // namespace

extern "C" {

const char* beads_last_error(void) { return g_last_error.c_str(); }

int beads_abi_version(void) { return BEADS_AMD_ABI_VERSION; }

beads_status_t beads_build_info(beads_build_info_t* out) {
  return guarded([&] {
    beads_amd::require(out != nullptr, "build_info output is null");
    out->real_precision = "single";
    out->real_dtype = "float32";
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

void beads_engine_destroy(beads_engine_t* e) {
  if (e != nullptr) {
    delete e->impl;
    delete e;
  }
}

}  // extern "C"
