// pybind11 module beads_amd._native (beads_amd._native_real64 when compiled
// with BEADS_REAL64 against libbeads_amd_real64.so): build_info() and
// execute_simulation(spec). Real arrays must have the build's real dtype.
//
// Counterpart of native/src/bindings/module.cpp:15-55 and the dict parser
// native/src/bindings/native_spec_parser.cpp:163-791 (array checks as in
// ndarray_reader.hpp:38-77, scalar checks as in native_spec_value_reader.hpp).
// The parsed spec crosses into the engine as the plain C struct beads_spec_t
// through the C ABI in include/beads_amd.h; nothing here touches the device.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "beads_amd.h"

namespace py = pybind11;

namespace {

constexpr const char* kRealDtypeName = sizeof(beads_real_t) == 8 ? "float64" : "float32";

struct TypeErr : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ValueErr : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] void type_error(const std::string& m) { throw TypeErr(m); }
[[noreturn]] void value_error(const std::string& m) { throw ValueErr(m); }

py::dict require_dict(const py::handle& obj, const std::string& ctx) {
  if (!py::isinstance<py::dict>(obj)) type_error(ctx + " must be a dict.");
  return py::reinterpret_borrow<py::dict>(obj);
}

py::object require_item(const py::dict& d, const char* key, const std::string& ctx) {
  if (!d.contains(key)) value_error(ctx + " requires key \"" + key + "\".");
  return d[key];
}

void require_only_keys(const py::dict& d, const std::set<std::string>& allowed, const std::string& ctx) {
  for (auto item : d) {
    if (!py::isinstance<py::str>(item.first)) type_error(ctx + " keys must be strings.");
    const std::string k = py::cast<std::string>(item.first);
    if (!allowed.count(k)) value_error(ctx + " has unsupported key \"" + k + "\".");
  }
}

bool is_bool(const py::handle& v) { return PyBool_Check(v.ptr()); }

int64_t require_int(const py::handle& v, const std::string& ctx) {
  if (is_bool(v) || !PyLong_Check(v.ptr())) type_error(ctx + " must be an integer.");
  return py::cast<int64_t>(v);
}

uint64_t require_uint64(const py::handle& v, const std::string& ctx) {
  if (is_bool(v) || !PyLong_Check(v.ptr())) type_error(ctx + " must be an integer.");
  if (py::cast<py::int_>(v) < py::int_(0)) value_error(ctx + " must be non-negative.");
  return py::cast<uint64_t>(v);
}

double require_real(const py::handle& v, const std::string& ctx) {
  if (is_bool(v) || !(PyLong_Check(v.ptr()) || PyFloat_Check(v.ptr()))) {
    type_error(ctx + " must be a real number.");
  }
  const double x = py::cast<double>(v);
  if (!std::isfinite(x)) value_error(ctx + " must be finite.");
  if (!std::isfinite(static_cast<beads_real_t>(x))) {
    value_error(ctx + " must be representable as " + kRealDtypeName + ".");
  }
  return x;
}

std::string require_str(const py::handle& v, const std::string& ctx) {
  if (!py::isinstance<py::str>(v)) type_error(ctx + " must be a string.");
  return py::cast<std::string>(v);
}

// ndarray_reader.hpp:38-77: exact dtype, exact shape, C-contiguous.
template <typename T>
py::array_t<T> require_array(const py::handle& obj, const std::vector<py::ssize_t>& shape,
                             const std::string& ctx) {
  if (!py::isinstance<py::array>(obj)) type_error(ctx + " must be a numpy array.");
  py::array a = py::reinterpret_borrow<py::array>(obj);
  if (!a.dtype().is(py::dtype::of<T>())) {
    type_error(ctx + " must have dtype " + std::string(py::str(py::dtype::of<T>())) + ".");
  }
  if (a.ndim() != static_cast<py::ssize_t>(shape.size())) value_error(ctx + " has the wrong shape.");
  for (size_t k = 0; k < shape.size(); ++k) {
    if (a.shape(k) != shape[k]) value_error(ctx + " has the wrong shape.");
  }
  if (!(a.flags() & py::array::c_style)) value_error(ctx + " must be C-contiguous.");
  return py::reinterpret_borrow<py::array_t<T>>(a);
}

struct Params {
  std::map<std::string, py::object> values;
};

Params read_params(const py::dict& d, const std::string& ctx) {
  Params p;
  for (auto item : d) {
    if (!py::isinstance<py::str>(item.first)) type_error(ctx + " parameter names must be strings.");
    const py::handle v = item.second;
    if (!(is_bool(v) || PyLong_Check(v.ptr()) || PyFloat_Check(v.ptr()) || py::isinstance<py::str>(v))) {
      type_error(ctx + " parameters must be bool, int, float or string.");
    }
    p.values[py::cast<std::string>(item.first)] = py::reinterpret_borrow<py::object>(v);
  }
  return p;
}

void require_param_keys(const Params& p, const std::set<std::string>& required,
                        const std::set<std::string>& optional, const std::string& ctx) {
  for (const auto& kv : p.values) {
    if (!required.count(kv.first) && !optional.count(kv.first)) {
      value_error(ctx + " does not accept parameter \"" + kv.first + "\".");
    }
  }
  for (const auto& k : required) {
    if (!p.values.count(k)) value_error(ctx + " requires parameter \"" + k + "\".");
  }
}

double positive_param(const Params& p, const char* key, const std::string& ctx) {
  const double v = require_real(p.values.at(key), ctx + "." + key);
  if (!(v > 0.0)) value_error(ctx + "." + key + " must be finite and positive.");
  return v;
}

double nonnegative_param(const Params& p, const char* key, const std::string& ctx) {
  const double v = require_real(p.values.at(key), ctx + "." + key);
  if (!(v >= 0.0)) value_error(ctx + "." + key + " must be finite and non-negative.");
  return v;
}

// Owns every buffer the C struct points into while the run executes.
struct ParsedSpec {
  beads_spec_t c{};
  std::vector<py::object> keep;  // numpy arrays borrowed for the call
  std::vector<uint32_t> bonds, angles, dihedrals;
  std::vector<beads_real_t> lj_eps, lj_sig, bond_k, bond_r0, angle_k, angle_theta0, dih_k;
  std::vector<int32_t> dih_d, dih_n;
  std::string thermo_path, trajectory_path, final_state_path, log_path;
};

void parse_system(const py::dict& sys, ParsedSpec& out) {
  require_only_keys(sys, {"units", "n_particles", "box_bound", "positions", "velocities", "masses",
                          "types", "tags", "molecule_ids", "images", "topology"},
                    "system");
  const std::string units = require_str(require_item(sys, "units", "system"), "system.units");
  if (units == "reduced") out.c.units = BEADS_UNITS_REDUCED;
  else if (units == "nm_kjmol") out.c.units = BEADS_UNITS_NM_KJMOL;
  else value_error("system.units must be \"reduced\" or \"nm_kjmol\".");
  const uint64_t n = require_uint64(require_item(sys, "n_particles", "system"), "system.n_particles");
  if (n == 0 || n > std::numeric_limits<uint32_t>::max()) value_error("system.n_particles must be in 1..2^32-1.");
  out.c.n_particles = static_cast<uint32_t>(n);
  const py::ssize_t N = static_cast<py::ssize_t>(n);
  auto bb = require_array<beads_real_t>(require_item(sys, "box_bound", "system"), {2, 3}, "system.box_bound");
  auto pos = require_array<beads_real_t>(require_item(sys, "positions", "system"), {N, 3}, "system.positions");
  auto vel = require_array<beads_real_t>(require_item(sys, "velocities", "system"), {N, 3}, "system.velocities");
  auto mass = require_array<beads_real_t>(require_item(sys, "masses", "system"), {N}, "system.masses");
  auto types = require_array<int32_t>(require_item(sys, "types", "system"), {N}, "system.types");
  auto tags = require_array<uint32_t>(require_item(sys, "tags", "system"), {N}, "system.tags");
  auto mol = require_array<uint32_t>(require_item(sys, "molecule_ids", "system"), {N}, "system.molecule_ids");
  auto img = require_array<int32_t>(require_item(sys, "images", "system"), {N, 3}, "system.images");
  out.keep.insert(out.keep.end(), {bb, pos, vel, mass, types, tags, mol, img});
  out.c.box_bound = bb.data();
  out.c.positions = pos.data();
  out.c.velocities = vel.data();
  out.c.masses = mass.data();
  out.c.types = types.data();
  out.c.tags = tags.data();
  out.c.molecule_ids = mol.data();
  out.c.images = img.data();

  const py::dict topo = require_dict(require_item(sys, "topology", "system"), "system.topology");
  require_only_keys(topo, {"bonds", "angles", "dihedrals"}, "system.topology");
  auto records = [&](const char* key, std::vector<const char*> fields, std::vector<uint32_t>& dst) {
    if (!topo.contains(key)) return;
    const py::handle list = topo[key];
    if (!py::isinstance<py::list>(list) && !py::isinstance<py::tuple>(list)) {
      type_error(std::string("system.topology.") + key + " must be a list.");
    }
    for (const py::handle rec : list) {
      const py::dict r = require_dict(rec, std::string("system.topology.") + key + " entry");
      for (const char* f : fields) {
        const int64_t v = require_int(require_item(r, f, key), std::string(key) + "." + f);
        if (v < 1 || v > std::numeric_limits<int32_t>::max()) {
          value_error(std::string("system.topology.") + key + "." + f + " must be positive.");
        }
        dst.push_back(static_cast<uint32_t>(v));
      }
    }
  };
  records("bonds", {"tag_i", "tag_j", "type"}, out.bonds);
  records("angles", {"tag_i", "tag_j", "tag_k", "type"}, out.angles);
  records("dihedrals", {"tag_i", "tag_j", "tag_k", "tag_l", "type"}, out.dihedrals);
  out.c.n_bonds = static_cast<uint32_t>(out.bonds.size() / 3);
  out.c.bonds = out.bonds.data();
  out.c.n_angles = static_cast<uint32_t>(out.angles.size() / 4);
  out.c.angles = out.angles.data();
  out.c.n_dihedrals = static_cast<uint32_t>(out.dihedrals.size() / 5);
  out.c.dihedrals = out.dihedrals.data();
}

// Dense 1..max coefficient list keyed by type (bond/angle/dihedral_coeffs).
std::map<int64_t, Params> typed_coeffs(const py::handle& list, const std::string& style,
                                       const std::string& ctx) {
  std::map<int64_t, Params> out;
  if (!py::isinstance<py::list>(list) && !py::isinstance<py::tuple>(list)) type_error(ctx + " must be a list.");
  for (const py::handle e : list) {
    const py::dict d = require_dict(e, ctx + " entry");
    if (require_str(require_item(d, "style", ctx), ctx + ".style") != style) {
      value_error(ctx + " style must match the active style.");
    }
    const int64_t t = require_int(require_item(d, "type", ctx), ctx + ".type");
    if (t < 1) value_error(ctx + ".type must be positive.");
    if (out.count(t)) value_error(ctx + " must not contain duplicates.");
    out[t] = read_params(require_dict(require_item(d, "params", ctx), ctx + ".params"), ctx);
  }
  int64_t expect = 1;
  for (const auto& kv : out) {
    if (kv.first != expect++) value_error(ctx + " must cover types 1..max without gaps.");
  }
  return out;
}

std::string optional_style(const py::dict& ff, const char* key, const std::string& expected) {
  const py::object v = require_item(ff, key, "forcefield");
  if (v.is_none()) return "";
  const py::dict d = require_dict(v, std::string("forcefield.") + key);
  const std::string style = require_str(require_item(d, "style", key), std::string(key) + ".style");
  if (style != expected) value_error(std::string("Unknown ") + key + " \"" + style + "\".");
  const Params p = read_params(require_dict(require_item(d, "params", key), std::string(key) + ".params"), key);
  require_param_keys(p, {}, {}, key);
  return style;
}

void parse_forcefield(const py::dict& ff, ParsedSpec& out, int32_t n_types) {
  const py::dict ps = require_dict(require_item(ff, "pair_style", "forcefield"), "forcefield.pair_style");
  const std::string pstyle = require_str(require_item(ps, "style", "pair_style"), "pair_style.style");
  if (pstyle != "lj") value_error("Unknown pair_style \"" + pstyle + "\".");
  out.c.pair_style = "lj";  // the style whose parameter block this parser fills
  const Params pp = read_params(require_dict(require_item(ps, "params", "pair_style"), "pair_style.params"),
                                "pair_style");
  require_param_keys(pp, {"cutoff"}, {"shift"}, "pair_style");
  out.c.lj_cutoff = static_cast<beads_real_t>(positive_param(pp, "cutoff", "pair_style"));
  out.c.lj_shift = 0;
  if (pp.values.count("shift")) {
    if (!is_bool(pp.values.at("shift"))) type_error("pair_style.shift must be a bool.");
    out.c.lj_shift = py::cast<bool>(pp.values.at("shift")) ? 1 : 0;
  }
  out.c.n_types = n_types;
  const size_t T = static_cast<size_t>(n_types);
  out.lj_eps.assign(T * T, 0.f);
  out.lj_sig.assign(T * T, 0.f);
  std::vector<uint8_t> filled(T * T, 0);
  const py::handle coeffs = require_item(ff, "pair_coeffs", "forcefield");
  for (const py::handle e : coeffs) {
    const py::dict d = require_dict(e, "pair_coeff");
    if (require_str(require_item(d, "style", "pair_coeff"), "pair_coeff.style") != "lj") {
      value_error("pair_coeff style must match active pair_style.");
    }
    const int64_t a = require_int(require_item(d, "type_i", "pair_coeff"), "pair_coeff.type_i");
    const int64_t b = require_int(require_item(d, "type_j", "pair_coeff"), "pair_coeff.type_j");
    if (a < 1 || b < 1 || a > n_types || b > n_types) {
      value_error("Pair(\"lj\") pair_coeff type ids must be active System types.");
    }
    const Params p = read_params(require_dict(require_item(d, "params", "pair_coeff"), "pair_coeff.params"),
                                 "pair_coeff");
    require_param_keys(p, {"epsilon", "sigma"}, {}, "pair_coeff");
    const beads_real_t eps = static_cast<beads_real_t>(positive_param(p, "epsilon", "pair_coeff"));
    const beads_real_t sig = static_cast<beads_real_t>(positive_param(p, "sigma", "pair_coeff"));
    const size_t ij = (a - 1) * T + (b - 1), ji = (b - 1) * T + (a - 1);
    if (filled[ij]) value_error("Pair(\"lj\") pair_coeff duplicate coeff for type pair.");
    filled[ij] = filled[ji] = 1;
    out.lj_eps[ij] = out.lj_eps[ji] = eps;
    out.lj_sig[ij] = out.lj_sig[ji] = sig;
  }
  for (uint8_t f : filled) {
    if (!f) value_error("Pair(\"lj\") pair_coeff requires complete type-pair coeff coverage.");
  }
  out.c.lj_epsilon = out.lj_eps.data();
  out.c.lj_sigma = out.lj_sig.data();

  if (!optional_style(ff, "bond_style", "harmonic").empty()) {
    out.c.bond_style = "harmonic";
    const auto m = typed_coeffs(require_item(ff, "bond_coeffs", "forcefield"), "harmonic", "bond_coeffs");
    for (const auto& kv : m) {
      require_param_keys(kv.second, {"k", "r0"}, {}, "bond_coeff");
      out.bond_k.push_back(static_cast<beads_real_t>(nonnegative_param(kv.second, "k", "bond_coeff")));
      out.bond_r0.push_back(static_cast<beads_real_t>(nonnegative_param(kv.second, "r0", "bond_coeff")));
    }
  }
  if (!optional_style(ff, "angle_style", "harmonic").empty()) {
    out.c.angle_style = "harmonic";
    const auto m = typed_coeffs(require_item(ff, "angle_coeffs", "forcefield"), "harmonic", "angle_coeffs");
    for (const auto& kv : m) {
      require_param_keys(kv.second, {"k", "theta0"}, {}, "angle_coeff");
      out.angle_k.push_back(static_cast<beads_real_t>(nonnegative_param(kv.second, "k", "angle_coeff")));
      const double th = nonnegative_param(kv.second, "theta0", "angle_coeff");
      if (th > 180.0) value_error("Angle(\"harmonic\") angle_coeff.theta0 must be between 0 and 180 degrees.");
      out.angle_theta0.push_back(static_cast<beads_real_t>(th));
    }
  }
  if (!optional_style(ff, "dihedral_style", "harmonic").empty()) {
    out.c.dihedral_style = "harmonic";
    const auto m = typed_coeffs(require_item(ff, "dihedral_coeffs", "forcefield"), "harmonic",
                                "dihedral_coeffs");
    for (const auto& kv : m) {
      require_param_keys(kv.second, {"k", "d", "n"}, {}, "dihedral_coeff");
      out.dih_k.push_back(static_cast<beads_real_t>(nonnegative_param(kv.second, "k", "dihedral_coeff")));
      const int64_t d = require_int(kv.second.values.at("d"), "dihedral_coeff.d");
      if (d != 1 && d != -1) value_error("Dihedral(\"harmonic\") dihedral_coeff.d must be +1 or -1.");
      const int64_t nn = require_int(kv.second.values.at("n"), "dihedral_coeff.n");
      if (nn < 0 || nn > std::numeric_limits<int32_t>::max()) {
        value_error("Dihedral(\"harmonic\") dihedral_coeff.n must be non-negative.");
      }
      out.dih_d.push_back(static_cast<int32_t>(d));
      out.dih_n.push_back(static_cast<int32_t>(nn));
    }
  }
  out.c.n_bond_types = static_cast<int32_t>(out.bond_k.size());
  out.c.bond_k = out.bond_k.data();
  out.c.bond_r0 = out.bond_r0.data();
  out.c.n_angle_types = static_cast<int32_t>(out.angle_k.size());
  out.c.angle_k = out.angle_k.data();
  out.c.angle_theta0_deg = out.angle_theta0.data();
  out.c.n_dihedral_types = static_cast<int32_t>(out.dih_k.size());
  out.c.dihedral_k = out.dih_k.data();
  out.c.dihedral_d = out.dih_d.data();
  out.c.dihedral_n = out.dih_n.data();

  // native_spec_parser.cpp:730-753 default exclusion policy.
  py::object dist = py::none();
  if (ff.contains("bonded_exclusion_distance")) dist = ff["bonded_exclusion_distance"];
  std::string policy = ff.contains("bonded_exclusion_policy")
                           ? require_str(ff["bonded_exclusion_policy"], "bonded_exclusion_policy")
                           : "auto";
  int32_t distance = 0;
  if (!dist.is_none()) {
    const int64_t d = require_int(dist, "bonded_exclusion_distance");
    if (d < 1 || d > 3) value_error("bond graph exclusion distance must be 1, 2, or 3.");
    distance = static_cast<int32_t>(d);
  }
  const bool has_bonds = out.c.n_bonds > 0;
  int32_t resolved = BEADS_EXCLUSION_NONE;
  if (policy == "auto") {
    if (distance != 0) {
      resolved = BEADS_EXCLUSION_EXPLICIT;
    } else if (has_bonds) {
      distance = 2;
      resolved = BEADS_EXCLUSION_DEFAULT;
    }
  } else if (policy == "default") {
    if (distance == 0 && has_bonds) distance = 2;
    resolved = BEADS_EXCLUSION_DEFAULT;
  } else if (policy == "none") {
    distance = 0;
  } else if (policy == "explicit") {
    resolved = BEADS_EXCLUSION_EXPLICIT;
  } else {
    value_error("forcefield.bonded_exclusion_policy is not supported.");
  }
  out.c.exclusion_distance = distance;
  out.c.exclusion_policy = resolved;
}

void parse_dynamics(const py::dict& dyn, ParsedSpec& out) {
  require_only_keys(dyn, {"style", "params", "thermostat"}, "dynamics");
  const std::string style = require_str(require_item(dyn, "style", "dynamics"), "dynamics.style");
  if (style != "velocity_verlet") value_error("Unknown Dynamics style \"" + style + "\".");
  const Params p = read_params(require_dict(require_item(dyn, "params", "dynamics"), "dynamics.params"),
                               "dynamics");
  require_param_keys(p, {"dt"}, {}, "Dynamics(\"velocity_verlet\")");
  out.c.dt = static_cast<beads_real_t>(positive_param(p, "dt", "dynamics"));
  out.c.integrator_style = "velocity_verlet";
  out.c.berendsen = 0;
  if (dyn.contains("thermostat") && !dyn["thermostat"].is_none()) {
    const py::dict th = require_dict(dyn["thermostat"], "dynamics.thermostat");
    const std::string ts = require_str(require_item(th, "style", "thermostat"), "thermostat.style");
    if (ts != "berendsen") value_error("Unknown Thermostat style \"" + ts + "\".");
    const Params tp = read_params(require_dict(require_item(th, "params", "thermostat"), "thermostat.params"),
                                  "thermostat");
    require_param_keys(tp, {"temperature", "tau"}, {}, "Thermostat(\"berendsen\")");
    out.c.berendsen = 1;
    out.c.thermostat_style = "berendsen";
    out.c.berendsen_temperature =
        static_cast<beads_real_t>(nonnegative_param(tp, "temperature", "Thermostat(\"berendsen\")"));
    out.c.berendsen_tau = static_cast<beads_real_t>(positive_param(tp, "tau", "Thermostat(\"berendsen\")"));
  }
}

void parse_neighbor(const py::dict& nb, ParsedSpec& out) {
  require_only_keys(nb, {"cutoff_buffer", "rebuild_check_every", "sort_every_rebuild", "max_neighbors",
                         "half_list", "sort_every_steps"},
                    "neighbor");
  // Optional, no reference counterpart: a step-based sort cadence.
  out.c.sort_every_steps = 0;
  if (nb.contains("sort_every_steps")) {
    out.c.sort_every_steps =
        require_uint64(nb["sort_every_steps"], "neighbor.sort_every_steps");
    if (out.c.sort_every_steps == 0) value_error("neighbor.sort_every_steps must be positive.");
  }
  // Optional, no reference counterpart: store each LJ pair once.
  out.c.half_list = 0;
  if (nb.contains("half_list")) {
    const py::handle h = nb["half_list"];
    if (!py::isinstance<py::bool_>(h)) type_error("neighbor.half_list must be a bool.");
    out.c.half_list = h.cast<bool>() ? 1 : 0;
  }
  const double buf = require_real(require_item(nb, "cutoff_buffer", "neighbor"), "neighbor.cutoff_buffer");
  if (buf < 0.0) value_error("neighbor cutoff_buffer must be finite and non-negative.");
  out.c.cutoff_buffer = static_cast<beads_real_t>(buf);
  out.c.rebuild_check_every =
      require_uint64(require_item(nb, "rebuild_check_every", "neighbor"), "neighbor.rebuild_check_every");
  out.c.sort_every_rebuild =
      require_uint64(require_item(nb, "sort_every_rebuild", "neighbor"), "neighbor.sort_every_rebuild");
  const uint64_t maxn = require_uint64(require_item(nb, "max_neighbors", "neighbor"), "neighbor.max_neighbors");
  if (out.c.rebuild_check_every == 0) value_error("neighbor rebuild_check_every must be positive.");
  if (out.c.sort_every_rebuild == 0) value_error("neighbor.sort_every_rebuild must be positive.");
  if (maxn == 0 || maxn > std::numeric_limits<uint32_t>::max()) {
    value_error("neighbor max_neighbors must be positive.");
  }
  out.c.max_neighbors = static_cast<uint32_t>(maxn);
}

std::string prefix_of(const py::dict& d, const std::string& ctx) {
  const std::string p = require_str(require_item(d, "prefix", ctx), ctx + ".prefix");
  if (p.empty()) value_error(ctx + ".prefix must not be empty.");
  return p;
}

void parse_output(const py::dict& o, ParsedSpec& out) {
  require_only_keys(o, {"log", "thermo", "final_state", "trajectory"}, "output");
  out.c.log_echo = BEADS_LOG_NONE;  // no output.log: no log output
  if (o.contains("log")) {
    const py::dict lg = require_dict(o["log"], "output.log");
    const std::string echo = require_str(require_item(lg, "echo", "output.log"), "output.log.echo");
    if (echo == "screen") out.c.log_echo = BEADS_LOG_SCREEN;
    else if (echo == "log") out.c.log_echo = BEADS_LOG_FILE;
    else if (echo == "both") out.c.log_echo = BEADS_LOG_BOTH;
    else if (echo == "none") out.c.log_echo = BEADS_LOG_NONE;
    else value_error("output.log.echo must be one of screen, log, both, none.");
    if (lg.contains("prefix")) {
      out.log_path = prefix_of(lg, "output.log") + ".log";
      out.c.log_path = out.log_path.c_str();
    }
  }
  if (o.contains("thermo")) {
    const py::dict th = require_dict(o["thermo"], "output.thermo");
    out.c.thermo_every = require_uint64(require_item(th, "every", "output.thermo"), "output.thermo.every");
    if (out.c.thermo_every == 0) value_error("output.thermo.every must be positive.");
    out.thermo_path = prefix_of(th, "output.thermo") + ".thermo.csv";
    out.c.thermo_path = out.thermo_path.c_str();
  }
  if (o.contains("final_state")) {
    const py::dict fs = require_dict(o["final_state"], "output.final_state");
    out.final_state_path = prefix_of(fs, "output.final_state") + ".state.beadsbin";
    out.c.final_state_path = out.final_state_path.c_str();
  }
  if (o.contains("trajectory")) {
    const py::dict tr = require_dict(o["trajectory"], "output.trajectory");
    out.c.trajectory_every =
        require_uint64(require_item(tr, "every", "output.trajectory"), "output.trajectory.every");
    if (out.c.trajectory_every == 0) value_error("output.trajectory.every must be positive.");
    out.trajectory_path = prefix_of(tr, "output.trajectory") + ".trajectory.beadsbin";
    out.c.trajectory_path = out.trajectory_path.c_str();
    static const char* kFields[] = {"tag", "type", "position", "image", "velocity", "force"};
    uint32_t mask = 0;
    int last = -1;
    for (const py::handle f : require_item(tr, "fields", "output.trajectory")) {
      const std::string name = require_str(f, "output.trajectory.fields entry");
      int idx = -1;
      for (int k = 0; k < 6; ++k) {
        if (name == kFields[k]) idx = k;
      }
      if (idx < 0) value_error("trajectory field \"" + name + "\" is not supported.");
      if (idx <= last) value_error("trajectory fields must follow canonical order and not repeat.");
      last = idx;
      mask |= 1u << idx;
    }
    if (!(mask & BEADS_TRAJ_TAG)) value_error("trajectory fields must include tag as the first field.");
    out.c.trajectory_fields = mask;
  }
}

// Interrupt hook: reacquire the GIL and run Python's signal handlers
// (module.cpp:29-34). A raised exception is parked until the run unwinds.
struct InterruptState {
  PyObject* type = nullptr;
  PyObject* value = nullptr;
  PyObject* trace = nullptr;
};

int check_signals(void* ctx) {
  py::gil_scoped_acquire gil;
  if (PyErr_CheckSignals() != 0) {
    auto* st = static_cast<InterruptState*>(ctx);
    PyErr_Fetch(&st->type, &st->value, &st->trace);
    return 1;
  }
  return 0;
}

[[noreturn]] void raise_status(beads_status_t st) {
  const std::string msg = beads_last_error();
  switch (st) {
    case BEADS_E_VALUE:
      throw py::value_error(msg);
    case BEADS_E_TYPE:
      throw py::type_error(msg);
    case BEADS_E_OVERFLOW:
      PyErr_SetString(PyExc_OverflowError, msg.c_str());
      throw py::error_already_set();
    case BEADS_E_INDEX:
      throw py::index_error(msg);
    default:
      throw std::runtime_error(msg);
  }
}

py::dict build_info() {
  beads_build_info_t info{};
  const beads_status_t st = beads_build_info(&info);
  if (st != BEADS_OK) raise_status(st);
  py::dict d;
  d["real_precision"] = info.real_precision;
  d["real_dtype"] = info.real_dtype;
  d["type_dtype"] = info.type_dtype;
  d["index_dtype"] = info.index_dtype;
  d["image_dtype"] = info.image_dtype;
  d["runstep_dtype"] = info.runstep_dtype;
  return d;
}

void execute_simulation(const py::object& spec_obj) {
  ParsedSpec ps;
  try {
    const py::dict spec = require_dict(spec_obj, "native spec");
    for (const char* k : {"system", "forcefield", "dynamics", "neighbor", "runsteps"}) {
      (void)require_item(spec, k, "native spec");
    }
    require_only_keys(spec, {"system", "forcefield", "dynamics", "neighbor", "output", "runsteps"},
                      "native spec");
    parse_system(require_dict(spec["system"], "system"), ps);
    // Number of types = largest type id present (types are dense from 1).
    int32_t n_types = 0;
    for (uint32_t i = 0; i < ps.c.n_particles; ++i) n_types = std::max(n_types, ps.c.types[i]);
    if (n_types < 1) value_error("system.types must be one-based positive integers.");
    parse_forcefield(require_dict(spec["forcefield"], "forcefield"), ps, n_types);
    parse_dynamics(require_dict(spec["dynamics"], "dynamics"), ps);
    parse_neighbor(require_dict(spec["neighbor"], "neighbor"), ps);
    if (spec.contains("output")) parse_output(require_dict(spec["output"], "output"), ps);
    else ps.c.log_echo = BEADS_LOG_NONE;
    ps.c.runsteps = require_uint64(spec["runsteps"], "runsteps");
  } catch (const TypeErr& e) {
    throw py::type_error(e.what());
  } catch (const ValueErr& e) {
    throw py::value_error(e.what());
  }
  InterruptState st;
  ps.c.interrupt = &check_signals;
  ps.c.interrupt_context = &st;
  beads_status_t status;
  {
    py::gil_scoped_release release;
    status = beads_execute_simulation(&ps.c);
  }
  if (status == BEADS_E_INTERRUPTED && st.type != nullptr) {
    PyErr_Restore(st.type, st.value, st.trace);
    throw py::error_already_set();
  }
  if (status != BEADS_OK) raise_status(status);
}

}  // namespace

#ifdef BEADS_REAL64
PYBIND11_MODULE(_native_real64, m) {
#else
PYBIND11_MODULE(_native, m) {
#endif
  m.doc() = "beads_amd native extension (MI355X engine)";
  m.def("build_info", &build_info);
  m.def("execute_simulation", &execute_simulation);
}
