// Built-in models and their registries (see models.hpp). Parameter checks
// keep the reference's messages (pair_lj.cu:196-232, bond_harmonic.cu:159-171,
// angle_harmonic.cu:238-255, dihedral_harmonic.cu:293-319,
// velocity_verlet.cu, berendsen_thermostat.cu:167-190).
#include "models.hpp"

#include <cmath>
#include <mutex>

namespace beads_amd {
namespace {

constexpr double kPiD = 3.14159265358979323846;

bool finite(double x) { return std::isfinite(x); }

// ------------------------------------------------------------------ pair ----

// pair_lj.cu: one global cutoff, per type pair {sigma^2, 24 eps, 4 eps, shift}
// in FP64 from the real_t parameters, plus FP64 copies of epsilon and sigma
// for the adjoint.
class LjPairModel final : public PairModel {
 public:
  LjPairModel(const beads_spec_t& s, const ModelContext&) {
    require(s.n_types >= 1 && s.lj_epsilon && s.lj_sigma,
            "forcefield requires pair_style \"lj\" with coefficients.");
    require(finite(s.lj_cutoff) && s.lj_cutoff > 0,
            "Pair(\"lj\") pair_style.cutoff must be finite and positive.");
    for (int32_t k = 0; k < s.n_types * s.n_types; ++k) {
      require(finite(s.lj_epsilon[k]) && s.lj_epsilon[k] > 0,
              "Pair(\"lj\") pair_coeff.epsilon must be finite and positive.");
      require(finite(s.lj_sigma[k]) && s.lj_sigma[k] > 0,
              "Pair(\"lj\") pair_coeff.sigma must be finite and positive.");
    }
    n_types_ = s.n_types;
    cutoff_ = s.lj_cutoff;
    shift_ = s.lj_shift != 0;
    const size_t tt = static_cast<size_t>(n_types_) * n_types_;
    eps_.assign(s.lj_epsilon, s.lj_epsilon + tt);
    sig_.assign(s.lj_sigma, s.lj_sigma + tt);
  }
  const char* style_name() const override { return "lj"; }
  real_t max_cutoff() const override { return cutoff_; }
  int32_t type_count() const override { return n_types_; }
  bool shift_energy() const override { return shift_; }
  void upload(hipStream_t stream) override {
    const std::vector<LjCoeff> t = pack_lj_table(n_types_, eps_.data(), sig_.data(), cutoff_, shift_);
    table_.upload(t.data(), t.size(), stream);
    const std::vector<double> e(eps_.begin(), eps_.end()), s(sig_.begin(), sig_.end());
    eps_d_.upload(e.data(), e.size(), stream);
    sig_d_.upload(s.data(), s.size(), stream);
  }
  void bind(ForceParams& f) const override {
    f.lj = table_.data();
    f.n_types = n_types_;
    f.rc2 = cutoff_ * cutoff_;
  }
  void set_type_pair_parameters(const real_t* eps, const real_t* sigma) override {
    const size_t tt = static_cast<size_t>(n_types_) * n_types_;
    for (size_t q = 0; q < tt; ++q) {
      require(std::isfinite(eps[q]) && eps[q] >= 0, "lj epsilon must be finite and >= 0.");
      require(std::isfinite(sigma[q]) && sigma[q] > 0, "lj sigma must be finite and > 0.");
    }
    eps_.assign(eps, eps + tt);
    sig_.assign(sigma, sigma + tt);
  }
  void bind_adjoint(AdjointParams& p) const override {
    require(n_types_ <= kMaxLdsTypes, "too many types for the adjoint kernels.", BEADS_E_VALUE);
    for (int a = 0; a < n_types_; ++a) {
      for (int b = 0; b < a; ++b) {
        require(eps_[a * n_types_ + b] == eps_[b * n_types_ + a] &&
                    sig_[a * n_types_ + b] == sig_[b * n_types_ + a],
                "the adjoint requires symmetric LJ tables.", BEADS_E_VALUE);
      }
    }
    p.rc2 = cutoff_ * cutoff_;
    p.rc2_f64 = static_cast<double>(cutoff_) * cutoff_;
    p.lj_shift = shift_ ? 1 : 0;
    p.n_types = n_types_;
    p.eps = eps_d_.data();
    p.sigma = sig_d_.data();
  }

 private:
  int32_t n_types_ = 0;
  real_t cutoff_ = 0;
  bool shift_ = false;
  std::vector<real_t> eps_, sig_;
  DeviceBuffer<LjCoeff> table_;
  DeviceBuffer<double> eps_d_, sig_d_;
};

// ---------------------------------------------------------------- listed ----

// bond_harmonic.cu: U = k/2 (r - r0)^2, table {k, r0} per type.
class HarmonicBondModel final : public ListedModel {
 public:
  HarmonicBondModel(const beads_spec_t& s, const ModelContext&) {
    for (int32_t t = 0; t < s.n_bond_types; ++t) {
      require(finite(s.bond_k[t]) && s.bond_k[t] >= 0,
              "Bond(\"harmonic\") bond_coeff.k must be non-negative.");
      require(finite(s.bond_r0[t]) && s.bond_r0[t] >= 0,
              "Bond(\"harmonic\") bond_coeff.r0 must be non-negative.");
    }
    n_ = s.n_bond_types;
    k_.assign(s.bond_k, s.bond_k + n_);
    r0_.assign(s.bond_r0, s.bond_r0 + n_);
  }
  const char* style_name() const override { return "harmonic"; }
  int32_t type_count() const override { return n_; }
  void upload(hipStream_t stream) override {
    std::vector<real2> c(std::max(1, n_), make_real2(0, 0));
    std::vector<double> k(std::max(1, n_), 0.0), r0(std::max(1, n_), 0.0);
    for (int t = 0; t < n_; ++t) {
      c[t] = make_real2(k_[t], r0_[t]);
      k[t] = k_[t];
      r0[t] = r0_[t];
    }
    coeff_.upload(c.data(), c.size(), stream);
    k_d_.upload(k.data(), k.size(), stream);
    r0_d_.upload(r0.data(), r0.size(), stream);
  }
  void bind(TopologyView& topo) const override { topo.bond_coeff = coeff_.data(); }
  std::vector<real4> table() const override {
    std::vector<real4> t(n_);
    for (int q = 0; q < n_; ++q) t[q] = make_real4(k_[q], r0_[q], 0, 0);
    return t;
  }
  void set_parameters(const real_t* k, const real_t* r0) override {
    for (int t = 0; t < n_; ++t) {
      require(std::isfinite(k[t]) && k[t] >= 0 && std::isfinite(r0[t]) && r0[t] >= 0,
              "bond k and r0 must be finite and non-negative.");
    }
    k_.assign(k, k + n_);
    r0_.assign(r0, r0 + n_);
  }
  void bind_adjoint(AdjointParams& p) const override {
    require(n_ <= kMaxBondTypesAdjoint, "too many types for the adjoint kernels.", BEADS_E_VALUE);
    p.n_bond_types = n_;
    p.bond_k = k_d_.data();
    p.bond_r0 = r0_d_.data();
  }

 private:
  int32_t n_ = 0;
  std::vector<real_t> k_, r0_;
  DeviceBuffer<real2> coeff_;
  DeviceBuffer<double> k_d_, r0_d_;
};

// angle_harmonic.cu: U = k/2 (theta - theta0)^2, theta0 given in degrees.
class HarmonicAngleModel final : public ListedModel {
 public:
  HarmonicAngleModel(const beads_spec_t& s, const ModelContext&) {
    for (int32_t t = 0; t < s.n_angle_types; ++t) {
      require(finite(s.angle_k[t]) && s.angle_k[t] >= 0,
              "Angle(\"harmonic\") angle_coeff.k must be non-negative.");
      require(finite(s.angle_theta0_deg[t]) && s.angle_theta0_deg[t] >= 0 &&
                  s.angle_theta0_deg[t] <= 180,
              "Angle(\"harmonic\") angle_coeff.theta0 must be between 0 and 180 degrees.");
    }
    n_ = s.n_angle_types;
    assign(s.angle_k, s.angle_theta0_deg);
  }
  const char* style_name() const override { return "harmonic"; }
  int32_t type_count() const override { return n_; }
  void upload(hipStream_t stream) override {
    coeff_.upload(c_.data(), c_.size(), stream);
    std::vector<double> k(std::max(1, n_), 0.0), th(std::max(1, n_), 0.0);
    for (int t = 0; t < n_; ++t) {
      k[t] = c_[t].x;
      th[t] = c_[t].y;
    }
    k_d_.upload(k.data(), k.size(), stream);
    theta0_d_.upload(th.data(), th.size(), stream);
  }
  void bind(TopologyView& topo) const override { topo.angle_coeff = coeff_.data(); }
  std::vector<real4> table() const override {
    std::vector<real4> t(n_);
    for (int q = 0; q < n_; ++q) t[q] = make_real4(c_[q].x, c_[q].y, 0, 0);
    return t;
  }
  // Frame mode: k and theta0 in degrees, as in the spec.
  void set_parameters(const real_t* k, const real_t* theta0_deg) override {
    for (int t = 0; t < n_; ++t) {
      require(std::isfinite(k[t]) && k[t] >= 0 && std::isfinite(theta0_deg[t]) &&
                  theta0_deg[t] >= 0 && theta0_deg[t] <= 180,
              "angle k must be finite and non-negative and theta0 between 0 and 180 degrees.");
    }
    assign(k, theta0_deg);
  }
  void bind_adjoint(AdjointParams& p) const override {
    require(n_ <= kMaxAngleTypesAdjoint, "too many angle types for the adjoint kernels.",
            BEADS_E_VALUE);
    p.n_angle_types = n_;
    p.angle_k = k_d_.data();
    p.angle_theta0 = theta0_d_.data();
  }

 private:
  void assign(const real_t* k, const real_t* theta0_deg) {
    c_.assign(std::max(1, n_), make_real2(0, 0));
    for (int t = 0; t < n_; ++t) {
      c_[t] = make_real2(k[t], theta0_deg[t] * static_cast<real_t>(kPiD) / real_t(180));
    }
  }
  int32_t n_ = 0;
  std::vector<real2> c_;
  DeviceBuffer<real2> coeff_;
  DeviceBuffer<double> k_d_, theta0_d_;
};

// dihedral_harmonic.cu: U = k [1 + d cos(n phi)], d = +-1, n >= 0.
class HarmonicDihedralModel final : public ListedModel {
 public:
  HarmonicDihedralModel(const beads_spec_t& s, const ModelContext&) {
    for (int32_t t = 0; t < s.n_dihedral_types; ++t) {
      require(finite(s.dihedral_k[t]) && s.dihedral_k[t] >= 0,
              "Dihedral(\"harmonic\") dihedral_coeff.k must be non-negative.");
      require(s.dihedral_d[t] == 1 || s.dihedral_d[t] == -1,
              "Dihedral(\"harmonic\") dihedral_coeff.d must be +1 or -1.");
      require(s.dihedral_n[t] >= 0, "Dihedral(\"harmonic\") dihedral_coeff.n must be non-negative.");
    }
    c_.resize(std::max(1, s.n_dihedral_types));
    for (int t = 0; t < s.n_dihedral_types; ++t) {
      c_[t] = make_real4(s.dihedral_k[t], static_cast<real_t>(s.dihedral_d[t]),
                         static_cast<real_t>(s.dihedral_n[t]), 0);
    }
    n_ = s.n_dihedral_types;
  }
  const char* style_name() const override { return "harmonic"; }
  int32_t type_count() const override { return n_; }
  void upload(hipStream_t stream) override {
    coeff_.upload(c_.data(), c_.size(), stream);
    std::vector<double> k(std::max(1, n_), 0.0), d(std::max(1, n_), 0.0),
        m(std::max(1, n_), 0.0);
    for (int t = 0; t < n_; ++t) {
      k[t] = c_[t].x;
      d[t] = c_[t].y;
      m[t] = c_[t].z;
    }
    k_d_.upload(k.data(), k.size(), stream);
    d_d_.upload(d.data(), d.size(), stream);
    n_d_.upload(m.data(), m.size(), stream);
  }
  void bind(TopologyView& topo) const override { topo.dih_coeff = coeff_.data(); }
  std::vector<real4> table() const override { return {c_.begin(), c_.begin() + n_}; }
  // Frame mode: k only (d and n are integers of the topology's style).
  void set_parameters(const real_t* k, const real_t*) override {
    for (int t = 0; t < n_; ++t) {
      require(std::isfinite(k[t]) && k[t] >= 0, "dihedral k must be finite and non-negative.");
    }
    for (int t = 0; t < n_; ++t) c_[t].x = k[t];
  }
  void bind_adjoint(AdjointParams& p) const override {
    require(n_ <= kMaxDihedralTypesAdjoint, "too many dihedral types for the adjoint kernels.",
            BEADS_E_VALUE);
    p.n_dihedral_types = n_;
    p.dih_k = k_d_.data();
    p.dih_d = d_d_.data();
    p.dih_n = n_d_.data();
  }

 private:
  int32_t n_ = 0;
  std::vector<real4> c_;
  DeviceBuffer<real4> coeff_;
  DeviceBuffer<double> k_d_, d_d_, n_d_;
};

// ------------------------------------------------------ dynamics ------------

// velocity_verlet.cu:59-111.
class VelocityVerletIntegrator final : public Integrator {
 public:
  VelocityVerletIntegrator(const beads_spec_t& s, const ModelContext&) : dt_(s.dt) {
    require(finite(s.dt) && s.dt > 0, "VelocityVerletIntegrator dt must be finite and positive.");
  }
  const char* style_name() const override { return "velocity_verlet"; }
  real_t dt() const override { return dt_; }
  void bind(IntegrateParams& in, ForceParams& f) const override {
    in.dt = dt_;
    f.dt = dt_;
  }

 private:
  real_t dt_;
};

// berendsen_thermostat.cu:104-251: lambda from the kinetic energy after the
// second half-kick, dof = 3N - 3.
class BerendsenThermostat final : public Thermostat {
 public:
  BerendsenThermostat(const beads_spec_t& s, const ModelContext& ctx) {
    require(finite(s.berendsen_temperature) && s.berendsen_temperature >= 0,
            "Thermostat(\"berendsen\").temperature must be non-negative.");
    require(finite(s.berendsen_tau) && s.berendsen_tau > 0,
            "Thermostat(\"berendsen\").tau must be positive.");
    target_thermal_ = ctx.kb * static_cast<double>(s.berendsen_temperature);
    tau_ = s.berendsen_tau;
  }
  const char* style_name() const override { return "berendsen"; }
  void bind(ForceParams& f, real_t dt) const override {
    f.berendsen = 1;
    f.target_thermal = target_thermal_;
    f.dt_over_tau = static_cast<double>(dt) / tau_;
  }

 private:
  double target_thermal_ = 0.0;
  real_t tau_ = 1;
};

template <class Model, class Impl>
typename ModelRegistry<Model>::Factory factory() {
  return [](const beads_spec_t& s, const ModelContext& ctx) -> std::unique_ptr<Model> {
    return std::make_unique<Impl>(s, ctx);
  };
}

// One registry per listed kind, so each keys its own style names.
struct BondKind {};
struct AngleKind {};
struct DihedralKind {};
template <class Kind>
class ListedRegistry : public ModelRegistry<ListedModel> {
 public:
  static ListedRegistry& instance() {
    static ListedRegistry registry;
    return registry;
  }
};

}  // namespace

void ListedModel::set_parameters(const real_t*, const real_t*) {
  fail(BEADS_E_VALUE, std::string("the ") + style_name() +
                          " style has no parameters that frame mode can update.");
}

void ListedModel::bind_adjoint(AdjointParams&) const {
  fail(BEADS_E_VALUE, std::string("the adjoint does not differentiate the ") + style_name() +
                          " style.");
}

void register_builtin_models() {
  static std::once_flag once;
  std::call_once(once, [] {
    ModelRegistry<PairModel>::instance().add("lj", factory<PairModel, LjPairModel>());
    ListedRegistry<BondKind>::instance().add("harmonic",
                                             factory<ListedModel, HarmonicBondModel>());
    ListedRegistry<AngleKind>::instance().add("harmonic",
                                              factory<ListedModel, HarmonicAngleModel>());
    ListedRegistry<DihedralKind>::instance().add("harmonic",
                                                 factory<ListedModel, HarmonicDihedralModel>());
    ModelRegistry<Integrator>::instance().add(
        "velocity_verlet", factory<Integrator, VelocityVerletIntegrator>());
    ModelRegistry<Thermostat>::instance().add("berendsen",
                                              factory<Thermostat, BerendsenThermostat>());
  });
}

const char* pair_style_of(const beads_spec_t& s) { return s.pair_style ? s.pair_style : "lj"; }
const char* bond_style_of(const beads_spec_t& s) {
  return s.bond_style ? s.bond_style : (s.n_bond_types > 0 ? "harmonic" : nullptr);
}
const char* angle_style_of(const beads_spec_t& s) {
  return s.angle_style ? s.angle_style : (s.n_angle_types > 0 ? "harmonic" : nullptr);
}
const char* dihedral_style_of(const beads_spec_t& s) {
  return s.dihedral_style ? s.dihedral_style : (s.n_dihedral_types > 0 ? "harmonic" : nullptr);
}
const char* integrator_style_of(const beads_spec_t& s) {
  return s.integrator_style ? s.integrator_style : "velocity_verlet";
}
const char* thermostat_style_of(const beads_spec_t& s) {
  return s.thermostat_style ? s.thermostat_style : (s.berendsen ? "berendsen" : nullptr);
}

ModelSet create_models(const beads_spec_t& s, const ModelContext& ctx) {
  register_builtin_models();
  ModelSet m;
  m.pair = ModelRegistry<PairModel>::instance().create(pair_style_of(s), "pair", s, ctx);
  if (const char* b = bond_style_of(s)) {
    m.bond = ListedRegistry<BondKind>::instance().create(b, "bond", s, ctx);
  }
  if (const char* a = angle_style_of(s)) {
    m.angle = ListedRegistry<AngleKind>::instance().create(a, "angle", s, ctx);
  }
  if (const char* d = dihedral_style_of(s)) {
    m.dihedral = ListedRegistry<DihedralKind>::instance().create(d, "dihedral", s, ctx);
  }
  m.integrator =
      ModelRegistry<Integrator>::instance().create(integrator_style_of(s), "integrator", s, ctx);
  if (const char* t = thermostat_style_of(s)) {
    m.thermostat = ModelRegistry<Thermostat>::instance().create(t, "thermostat", s, ctx);
  }
  return m;
}

// pair_lj.cu:171-192: {sigma^2, 24 eps, 4 eps, shift} per ordered type pair in
// FP64 from the real_t parameters; shift = 4 eps ((sigma/rc)^12 - (sigma/rc)^6)
// when the energy is shifted.
std::vector<LjCoeff> pack_lj_table(int32_t n_types, const real_t* epsilon, const real_t* sigma,
                                   real_t cutoff, bool shift_energy) {
  const int T = n_types;
  std::vector<LjCoeff> lj(static_cast<size_t>(T) * T);
  const double rc = cutoff;
  for (int a = 0; a < T; ++a) {
    for (int b = 0; b < T; ++b) {
      const double eps = epsilon[a * T + b], sig = sigma[a * T + b];
      LjCoeff c;
      c.sigma2 = sig * sig;
      c.eps4 = 4.0 * eps;
      c.eps24 = 24.0 * eps;
      c.shift = 0.0;
      if (shift_energy) {
        const double sr2 = c.sigma2 / (rc * rc);
        const double sr6 = sr2 * sr2 * sr2;
        c.shift = c.eps4 * (sr6 * sr6 - sr6);
      }
      lj[static_cast<size_t>(a) * T + b] = c;
    }
  }
  return lj;
}

std::vector<LjCoeff> pack_lj_table(const beads_spec_t& s) {
  return pack_lj_table(s.n_types, s.lj_epsilon, s.lj_sigma, s.lj_cutoff, s.lj_shift != 0);
}

}  // namespace beads_amd
