// Force-field, integrator and thermostat models behind style registries.
//
// The reference's seams (native/src/forcefield/pair/pair_model.hpp:42-61 with
// pair_factory.cpp:30-47; bond_model.hpp:41-57 and the bond/angle/dihedral
// factories; dynamics/integrator/integrator.cuh:14-30;
// dynamics/thermostat/thermostat.cuh:11-21) are re-created here for an engine
// whose per-step work is a few fused kernels: a model does not launch its own
// pass, it validates and packs its parameters, uploads its device tables and
// binds them into the parameter blocks of the fused passes (the force pass,
// the integrator pass, the adjoint). Each kind has a registry keyed by style
// name; the engine creates its models from the style names of the spec, so a
// new style is a model class, its registration, and its term in the fused
// kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "beads_amd.h"
#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace beads_amd {

// What a factory gets besides the spec.
struct ModelContext {
  double kb = 1.0;  // Boltzmann constant of the spec's unit system
};

// pair_model.hpp:42-61: the model that owns the base (pair) force write.
class PairModel {
 public:
  virtual ~PairModel() = default;
  virtual const char* style_name() const = 0;
  virtual real_t max_cutoff() const = 0;
  virtual int32_t type_count() const = 0;
  virtual bool shift_energy() const = 0;
  virtual void upload(hipStream_t stream) = 0;
  virtual void bind(ForceParams& f) const = 0;
  // Frame mode: new per-type-pair parameters (real_t, [T][T]).
  virtual void set_type_pair_parameters(const real_t* a, const real_t* b) = 0;
  // The adjoint's view of the parameters; throws if the adjoint cannot
  // differentiate this model in its current state.
  virtual void bind_adjoint(AdjointParams& p) const = 0;
};

// bond_model.hpp:41-57 and its angle/dihedral counterparts: per-type
// coefficient tables read by the listed-term section of the force pass.
class ListedModel {
 public:
  virtual ~ListedModel() = default;
  virtual const char* style_name() const = 0;
  virtual int32_t type_count() const = 0;
  virtual void upload(hipStream_t stream) = 0;
  virtual void bind(TopologyView& topo) const = 0;
  // The packed coefficients of each type as the force pass reads them
  // (bond {k, r0}, angle {k, theta0 rad}, dihedral {k, d, n}), zero-padded to
  // four components; the batched replicas concatenate these tables.
  virtual std::vector<real4> table() const = 0;
  // Frame mode / adjoint: new real_t parameters and the adjoint's FP64 view
  // of them.
  virtual void set_parameters(const real_t* a, const real_t* b);
  virtual void bind_adjoint(AdjointParams& p) const;
};

// integrator.cuh:14-30: pre-force (first half-kick, drift, wrap) runs in the
// integrate pass, post-force (second half-kick) in the force pass.
class Integrator {
 public:
  virtual ~Integrator() = default;
  virtual const char* style_name() const = 0;
  virtual real_t dt() const = 0;
  virtual void bind(IntegrateParams& in, ForceParams& f) const = 0;
};

// thermostat.cuh:11-21: the factor is formed in the force pass's reduction and
// applied at the start of the next integrate pass.
class Thermostat {
 public:
  virtual ~Thermostat() = default;
  virtual const char* style_name() const = 0;
  virtual void bind(ForceParams& f, real_t dt) const = 0;
};

template <class Model>
class ModelRegistry {
 public:
  using Factory = std::function<std::unique_ptr<Model>(const beads_spec_t&, const ModelContext&)>;

  static ModelRegistry& instance() {
    static ModelRegistry registry;
    return registry;
  }
  void add(const std::string& style, Factory factory) { factories_[style] = std::move(factory); }
  bool has(const std::string& style) const { return factories_.count(style) != 0; }
  std::vector<std::string> styles() const {
    std::vector<std::string> out;
    for (const auto& kv : factories_) out.push_back(kv.first);
    return out;
  }
  // pair_factory.cpp:30-47: an unknown style is an invalid argument.
  std::unique_ptr<Model> create(const std::string& style, const char* kind,
                                const beads_spec_t& spec, const ModelContext& ctx) const {
    const auto it = factories_.find(style);
    if (it == factories_.end()) {
      fail(BEADS_E_VALUE, std::string("unsupported ") + kind + " style \"" + style + "\".");
    }
    return it->second(spec, ctx);
  }

 private:
  std::map<std::string, Factory> factories_;
};

// Registers "lj"; "harmonic" bond, angle, dihedral; "velocity_verlet";
// "berendsen". Idempotent; the engine calls it before creating models.
void register_builtin_models();

// The models of one spec, created from its style names.
struct ModelSet {
  std::unique_ptr<PairModel> pair;
  std::unique_ptr<ListedModel> bond, angle, dihedral;  // null: style absent
  std::unique_ptr<Integrator> integrator;
  std::unique_ptr<Thermostat> thermostat;              // null: none
};
ModelSet create_models(const beads_spec_t& spec, const ModelContext& ctx);

// Style names a spec selects (defaults where the spec leaves a name NULL).
const char* pair_style_of(const beads_spec_t& spec);
const char* bond_style_of(const beads_spec_t& spec);      // nullptr: none
const char* angle_style_of(const beads_spec_t& spec);     // nullptr: none
const char* dihedral_style_of(const beads_spec_t& spec);  // nullptr: none
const char* integrator_style_of(const beads_spec_t& spec);
const char* thermostat_style_of(const beads_spec_t& spec);  // nullptr: none

// pair_lj.cu:171-192 coefficient packing, shared with the batched replicas.
std::vector<LjCoeff> pack_lj_table(int32_t n_types, const real_t* epsilon, const real_t* sigma,
                                   real_t cutoff, bool shift_energy);
std::vector<LjCoeff> pack_lj_table(const beads_spec_t& spec);

}  // namespace beads_amd
