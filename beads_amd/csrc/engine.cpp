// Host engine implementation. See engine.hpp.
#include "engine.hpp"

#include "adjoint.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <limits>
#include <numeric>
#include <sstream>

namespace beads_amd {

namespace {

constexpr uint64_t kOverflowPollSteps = 256;

bool finite(double v) { return std::isfinite(v); }

std::string fmt9(double v) {
  std::ostringstream s;
  s << std::setprecision(9) << static_cast<real_t>(v);
  return s.str();
}

// Log layout (log_output.cpp:19-21, 72-90): "  label", padded so that ": "
// starts at column 36, then the value.
constexpr int kLogValueColumn = 2 + 34 + 3;
constexpr int kThermoColumnWidth = 16;

template <typename T>
std::string label_value(const std::string& label, const T& value) {
  std::ostringstream line;
  const std::string head = "  " + label;
  const int pad = kLogValueColumn - static_cast<int>(head.size()) - 3;
  line << head << (pad > 0 ? std::string(static_cast<size_t>(pad), ' ') : std::string(" "))
       << ": " << value;
  return line.str();
}

// '(' x ',' y ',' z ')' at the stream's default precision (log_output.cpp:65-69).
std::string triple(const real_t* v) {
  std::ostringstream s;
  s << '(' << v[0] << ',' << v[1] << ',' << v[2] << ')';
  return s.str();
}

std::string prec9(double v) {
  std::ostringstream s;
  s << std::setprecision(9) << v;
  return s.str();
}

template <typename T>
void write_raw(std::ofstream& out, const T* data, size_t count, const std::string& what) {
  out.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(count * sizeof(T)));
  if (!out) {
    fail(BEADS_E_RUNTIME, "failed to write " + what + ".");
  }
}

template <typename T>
void write_scalar(std::ofstream& out, T value, const std::string& what) {
  write_raw(out, &value, 1, what);
}

}  // namespace

UnitSystem unit_system_from(int32_t units) {
  if (units == BEADS_UNITS_REDUCED) {
    return {"reduced", 1.0, 1.0};
  }
  if (units == BEADS_UNITS_NM_KJMOL) {
    return {"nm_kjmol", static_cast<double>(static_cast<real_t>(0.00831446261815324)),
            static_cast<double>(static_cast<real_t>(16.60539067))};
  }
  fail(BEADS_E_VALUE, "system.units must be \"reduced\" or \"nm_kjmol\".");
}

BoxView make_box_view(const real_t* bb) {
  BoxView b{};
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = bb[a];
    b.L[a] = bb[3 + a] - bb[a];
    b.hL[a] = real_t(0.5) * b.L[a];
    b.invL[a] = real_t(1) / b.L[a];
  }
  return b;
}

NeighborPlan make_neighbor_plan(const BoxView& box, real_t rc, real_t buffer, uint32_t n,
                                uint32_t max_neighbors, uint64_t check_every,
                                uint64_t sort_every) {
  NeighborPlan p;
  require(n > 0, "neighbor n_particles must be positive.");
  require(finite(rc) && rc > 0, "neighbor force_cutoff must be finite and positive.");
  require(finite(buffer) && buffer >= 0,
          "neighbor cutoff_buffer must be finite and non-negative.");
  require(max_neighbors > 0, "neighbor max_neighbors must be positive.");
  require(check_every > 0, "neighbor rebuild_check_every must be positive.");
  require(sort_every > 0, "neighbor.sort_every_rebuild must be positive.");
  p.force_cutoff = rc;
  p.cutoff_buffer = buffer;
  p.search_cutoff = rc + buffer;
  p.search_cutoff_sq = p.search_cutoff * p.search_cutoff;
  p.max_neighbors = max_neighbors;
  p.check_every = check_every;
  p.sort_every = sort_every;
  const real_t half_shortest = real_t(0.5) * std::min(box.L[0], std::min(box.L[1], box.L[2]));
  require(p.search_cutoff <= half_shortest,
          "neighbor search_cutoff must not exceed half shortest box length.");
  uint32_t counts[3];
  for (int a = 0; a < 3; ++a) {
    const double raw =
        std::floor(static_cast<double>(box.L[a]) / static_cast<double>(p.search_cutoff));
    require(std::isfinite(raw), "neighbor box length cell count must be finite.");
    require(raw <= static_cast<double>(std::numeric_limits<uint32_t>::max()),
            "neighbor cell count exceeds index_t range.");
    counts[a] = raw < 1.0 ? 1u : static_cast<uint32_t>(raw);
    p.cell_size[a] = box.L[a] / static_cast<real_t>(counts[a]);
    p.inv_cell[a] = real_t(1) / p.cell_size[a];
    if (counts[a] >= 3) {
      const real_t tol = std::max(real_t(1.0e-6), p.search_cutoff * real_t(1.0e-5));
      require(p.cell_size[a] + tol >= p.search_cutoff,
              "neighbor cell size must be at least neighbor search_cutoff.");
    }
  }
  const uint64_t cells = static_cast<uint64_t>(counts[0]) * counts[1] * counts[2];
  require(cells <= std::numeric_limits<uint32_t>::max() - 1,
          "neighbor n_cells exceeds index_t range.");
  p.nx = counts[0];
  p.ny = counts[1];
  p.nz = counts[2];
  p.n_cells = static_cast<uint32_t>(cells);
  if (p.nx < 3 && p.ny < 3 && p.nz < 3) {
    p.path = 0;
  } else if (p.nx >= 3 && p.ny >= 3 && p.nz >= 3) {
    p.path = 1;
  } else {
    p.path = 2;
  }
  return p;
}

std::vector<std::pair<uint32_t, uint32_t>> bond_graph_exclusions(uint32_t n, uint32_t n_bonds,
                                                                 const uint32_t* bonds,
                                                                 int32_t distance) {
  std::vector<std::pair<uint32_t, uint32_t>> out;
  if (distance <= 0 || n_bonds == 0) {
    return out;
  }
  // Tag-indexed adjacency in CSR form.
  std::vector<uint32_t> deg(static_cast<size_t>(n) + 2, 0);
  for (uint32_t b = 0; b < n_bonds; ++b) {
    ++deg[bonds[3 * b]];
    ++deg[bonds[3 * b + 1]];
  }
  std::vector<uint32_t> off(static_cast<size_t>(n) + 2, 0);
  for (uint32_t t = 1; t <= n + 1; ++t) {
    off[t] = off[t - 1] + deg[t - 1];
  }
  std::vector<uint32_t> adj(off[n + 1]);
  std::vector<uint32_t> fill(off.begin(), off.end() - 1);
  for (uint32_t b = 0; b < n_bonds; ++b) {
    const uint32_t a = bonds[3 * b], c = bonds[3 * b + 1];
    adj[fill[a]++] = c;
    adj[fill[c]++] = a;
  }
  std::vector<uint32_t> mark(static_cast<size_t>(n) + 1, 0);
  std::vector<std::pair<uint32_t, int>> frontier;
  for (uint32_t start = 1; start <= n; ++start) {
    if (deg[start] == 0) {
      continue;
    }
    frontier.clear();
    frontier.emplace_back(start, 0);
    mark[start] = start;
    for (size_t r = 0; r < frontier.size(); ++r) {
      const auto [t, depth] = frontier[r];
      if (depth == distance) {
        continue;
      }
      for (uint32_t e = off[t]; e < off[t + 1]; ++e) {
        const uint32_t q = adj[e];
        if (mark[q] == start) {
          continue;
        }
        mark[q] = start;
        if (q > start) {
          out.emplace_back(start, q);
        }
        frontier.emplace_back(q, depth + 1);
      }
    }
  }
  std::sort(out.begin(), out.end());
  return out;
}

uint32_t inline_capacity(const std::vector<uint32_t>& counts) {
  constexpr uint32_t kSmall = 16;
  const uint32_t mx = counts.empty() ? 0u : *std::max_element(counts.begin(), counts.end());
  if (mx <= kSmall) return mx;
  const size_t allow = std::max<size_t>(16, counts.size() / 256);
  std::vector<size_t> above(static_cast<size_t>(mx) + 2, 0);  // above[c] = #(count > c)
  for (const uint32_t c : counts) ++above[c];
  size_t run = 0;
  for (int64_t c = mx; c >= 0; --c) {
    const size_t here = above[static_cast<size_t>(c)];
    above[static_cast<size_t>(c)] = run;
    run += here;
  }
  uint32_t cap = mx;
  for (uint32_t c = kSmall; c <= mx; ++c) {
    if (above[c] <= allow) {
      cap = c;
      break;
    }
  }
  return std::min<uint32_t>(cap, 255u);
}

// Per-slot records of the listed terms (every role of every term, so each
// atom evaluates its own terms) and of the excluded partners. The reference
// keeps bonded terms as term lists (bond_harmonic.cu:24-99) and per-slot
// exclusions as up to 6 inline partners plus an overflow region
// (exclusions.cu:131-210, 337-445); here every kind has an inline capacity
// (inline_capacity) and records past it go to the particle's spill segment,
// so a particle may take part in any number of terms and exclusions.
HostTopology build_host_topology(const beads_spec_t& s, bool inline_only) {
  HostTopology h;
  const uint32_t n = s.n_particles;
  std::vector<uint32_t> slot_of_tag(static_cast<size_t>(n) + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    slot_of_tag[s.tags[i]] = i;
  }
  std::vector<uint32_t> nb(n, 0), na(n, 0), nd(n, 0), ne(n, 0);
  for (uint32_t b = 0; b < s.n_bonds; ++b) {
    ++nb[slot_of_tag[s.bonds[3 * b]]];
    ++nb[slot_of_tag[s.bonds[3 * b + 1]]];
  }
  for (uint32_t a = 0; a < s.n_angles; ++a) {
    for (int r = 0; r < 3; ++r) ++na[slot_of_tag[s.angles[4 * a + r]]];
  }
  for (uint32_t d = 0; d < s.n_dihedrals; ++d) {
    for (int r = 0; r < 4; ++r) ++nd[slot_of_tag[s.dihedrals[5 * d + r]]];
  }
  const auto pairs = bond_graph_exclusions(n, s.n_bonds, s.bonds, s.exclusion_distance);
  h.excluded_pairs = pairs.size();
  for (const auto& p : pairs) {
    ++ne[slot_of_tag[p.first]];
    ++ne[slot_of_tag[p.second]];
  }
  auto capacity = [&](const std::vector<uint32_t>& c, const char* kind) {
    if (!inline_only) return inline_capacity(c);
    const uint32_t mx = c.empty() ? 0u : *std::max_element(c.begin(), c.end());
    require(mx <= 255u, std::string("batched replicas keep at most 255 ") + kind +
                            " per particle; a particle has " + std::to_string(mx) + ".");
    return mx;
  };
  h.bcap = capacity(nb, "bonds");
  h.acap = capacity(na, "angles");
  h.dcap = capacity(nd, "dihedrals");
  h.ecap = capacity(ne, "exclusions");
  const uint32_t caps[4] = {h.bcap, h.acap, h.dcap, h.ecap};
  const std::vector<uint32_t>* all[4] = {&nb, &na, &nd, &ne};
  // Spill entries and segment offsets.
  std::vector<uint32_t> entry(n, kNoSpill);
  size_t spill_total[4] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < n; ++i) {
    bool spills = false;
    for (int q = 0; q < 4; ++q) spills |= (*all[q])[i] > caps[q];
    if (!spills) continue;
    SpillEntry e{};
    for (int q = 0; q < 4; ++q) {
      const uint32_t extra = (*all[q])[i] > caps[q] ? (*all[q])[i] - caps[q] : 0u;
      require(spill_total[q] + extra <= 0xffffffffull, "topology records exceed 2^32 per kind.",
              BEADS_E_OVERFLOW);
      e.off[q] = static_cast<uint32_t>(spill_total[q]);
      e.cnt[q] = extra;
      spill_total[q] += extra;
    }
    entry[i] = static_cast<uint32_t>(h.spill.size());
    h.spill.push_back(e);
  }
  if (!h.spill.empty()) {
    h.spill_index = entry;
    h.spill_bond.assign(spill_total[0], uint2{0, 0});
    h.spill_angle.assign(spill_total[1], uint4{0, 0, 0, 0});
    h.spill_dih.assign(spill_total[2], uint4{0, 0, 0, 0});
    h.spill_excl.assign(spill_total[3], 0u);
    h.spills_exclusions = spill_total[3] > 0;
  }
  h.count.assign(n, 0);
  h.bond.assign(static_cast<size_t>(h.bcap) * n, uint2{0, 0});
  h.angle.assign(static_cast<size_t>(h.acap) * n, uint4{0, 0, 0, 0});
  h.dih.assign(static_cast<size_t>(h.dcap) * n, uint4{0, 0, 0, 0});
  h.excl.assign(static_cast<size_t>(h.ecap) * n, 0u);
  std::vector<uint32_t> cb(n, 0), ca(n, 0), cd(n, 0), ce(n, 0);
  const size_t N = n;
  // k-th record of slot i of kind q: inline below the capacity, else spilled.
  auto put_bond = [&](uint32_t i, uint2 r) {
    const uint32_t k = cb[i]++;
    if (k < h.bcap) h.bond[k * N + i] = r;
    else h.spill_bond[h.spill[entry[i]].off[0] + (k - h.bcap)] = r;
  };
  auto put_angle = [&](uint32_t i, uint4 r) {
    const uint32_t k = ca[i]++;
    if (k < h.acap) h.angle[k * N + i] = r;
    else h.spill_angle[h.spill[entry[i]].off[1] + (k - h.acap)] = r;
  };
  auto put_dih = [&](uint32_t i, uint4 r) {
    const uint32_t k = cd[i]++;
    if (k < h.dcap) h.dih[k * N + i] = r;
    else h.spill_dih[h.spill[entry[i]].off[2] + (k - h.dcap)] = r;
  };
  auto put_excl = [&](uint32_t i, uint32_t partner) {
    const uint32_t k = ce[i]++;
    if (k < h.ecap) h.excl[k * N + i] = partner;
    else h.spill_excl[h.spill[entry[i]].off[3] + (k - h.ecap)] = partner;
  };
  for (uint32_t b = 0; b < s.n_bonds; ++b) {
    const uint32_t si = slot_of_tag[s.bonds[3 * b]], sj = slot_of_tag[s.bonds[3 * b + 1]];
    const uint32_t ty = s.bonds[3 * b + 2] - 1;
    put_bond(si, uint2{sj, ty | 0x80000000u});
    put_bond(sj, uint2{si, ty});
  }
  for (uint32_t a = 0; a < s.n_angles; ++a) {
    const uint32_t si = slot_of_tag[s.angles[4 * a]], sj = slot_of_tag[s.angles[4 * a + 1]],
                   sk = slot_of_tag[s.angles[4 * a + 2]];
    const uint32_t ty = s.angles[4 * a + 3] - 1;
    put_angle(si, uint4{sj, sk, ty, 0});
    put_angle(sj, uint4{si, sk, ty, 1});
    put_angle(sk, uint4{si, sj, ty, 2});
  }
  for (uint32_t d = 0; d < s.n_dihedrals; ++d) {
    uint32_t sl[4];
    for (int r = 0; r < 4; ++r) sl[r] = slot_of_tag[s.dihedrals[5 * d + r]];
    const uint32_t ty = s.dihedrals[5 * d + 4] - 1;
    for (uint32_t r = 0; r < 4; ++r) {
      uint32_t others[3];
      int m = 0;
      for (int q = 0; q < 4; ++q) {
        if (q != static_cast<int>(r)) others[m++] = sl[q];
      }
      put_dih(sl[r], uint4{others[0], others[1], others[2], ty | (r << 30)});
    }
  }
  for (const auto& p : pairs) {
    const uint32_t a = slot_of_tag[p.first], b = slot_of_tag[p.second];
    put_excl(a, b);
    put_excl(b, a);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t ib = std::min(nb[i], h.bcap), ia = std::min(na[i], h.acap);
    const uint32_t id = std::min(nd[i], h.dcap), ie = std::min(ne[i], h.ecap);
    h.count[i] = ib | (ia << 8) | (id << 16) | (ie << 24);
  }
  return h;
}

// Shell validation of the typed spec (native_spec_validation.hpp:23-716,
// restated for the fields this engine consumes).
void validate_spec(const beads_spec_t& s) {
  (void)unit_system_from(s.units);
  require(s.n_particles > 0, "system.n_particles must be positive.");
  // The force pass gathers neighbour positions by a 32-bit byte offset.
  constexpr uint64_t kMaxParticles = (uint64_t{1} << 32) / sizeof(real4) - 1;
  require(s.n_particles <= kMaxParticles,
          "system.n_particles must be at most " + std::to_string(kMaxParticles) +
              " with " + kRealPrecision + "-precision storage.",
          BEADS_E_OVERFLOW);
  require(s.box_bound && s.positions && s.velocities && s.masses && s.types && s.tags &&
              s.molecule_ids && s.images,
          "system arrays must not be null.", BEADS_E_TYPE);
  const uint32_t n = s.n_particles;
  for (int a = 0; a < 3; ++a) {
    require(finite(s.box_bound[a]) && finite(s.box_bound[3 + a]), "system.box_bound must be finite.");
    require(s.box_bound[3 + a] > s.box_bound[a],
            "system.box_bound upper row must be greater than lower row.");
  }
  std::vector<uint8_t> seen(static_cast<size_t>(n) + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      const real_t x = s.positions[3 * i + a];
      require(finite(x), "system.positions must be finite.");
      require(x >= s.box_bound[a] && x < s.box_bound[3 + a],
              "system.positions must be inside box_bound.");
      require(finite(s.velocities[3 * i + a]), "system.velocities must be finite.");
    }
    require(finite(s.masses[i]) && s.masses[i] > 0, "system.masses must be finite and positive.");
    require(s.types[i] >= 1 && s.types[i] <= s.n_types,
            "system.types must be covered by the pair coefficient table.");
    const uint32_t t = s.tags[i];
    require(t >= 1 && t <= n && !seen[t], "system.tags must be a permutation of 1..N.");
    seen[t] = 1;
    require(s.molecule_ids[i] >= 1, "system.molecule_ids must be one-based.");
  }
  require(s.n_types >= 1, "system.types must be covered by the pair coefficient table.");
  auto check_tag = [&](uint32_t t, const char* what) {
    require(t >= 1 && t <= n, std::string(what) + " references a tag outside the system.");
  };
  require(s.n_bonds == 0 || s.n_bond_types >= 1, "System topology bonds require an active bond_style.");
  for (uint32_t b = 0; b < s.n_bonds; ++b) {
    check_tag(s.bonds[3 * b], "bond");
    check_tag(s.bonds[3 * b + 1], "bond");
    require(s.bonds[3 * b] != s.bonds[3 * b + 1], "bond entries must reference distinct tags.");
    require(s.bonds[3 * b + 2] >= 1 && s.bonds[3 * b + 2] <= static_cast<uint32_t>(s.n_bond_types),
            "bond type is not covered by bond_coeffs.");
  }
  require(s.n_angles == 0 || s.n_angle_types >= 1, "System topology angles require an active angle_style.");
  for (uint32_t a = 0; a < s.n_angles; ++a) {
    for (int r = 0; r < 3; ++r) check_tag(s.angles[4 * a + r], "angle");
    require(s.angles[4 * a + 3] >= 1 && s.angles[4 * a + 3] <= static_cast<uint32_t>(s.n_angle_types),
            "angle type is not covered by angle_coeffs.");
  }
  require(s.n_dihedrals == 0 || s.n_dihedral_types >= 1,
          "System topology dihedrals require an active dihedral_style.");
  for (uint32_t d = 0; d < s.n_dihedrals; ++d) {
    for (int r = 0; r < 4; ++r) check_tag(s.dihedrals[5 * d + r], "dihedral");
    require(s.dihedrals[5 * d + 4] >= 1 &&
                s.dihedrals[5 * d + 4] <= static_cast<uint32_t>(s.n_dihedral_types),
            "dihedral type is not covered by dihedral_coeffs.");
  }
  require(s.exclusion_distance >= 0 && s.exclusion_distance <= 3,
          "bond graph exclusion distance must be 1, 2, or 3.");
  require(s.exclusion_distance == 0 || s.n_bonds > 0,
          "ForceField.exclude_bonded(...) requires System topology bonds.");
  // Style parameters are checked by their models (models.cpp).
  ModelContext ctx;
  ctx.kb = unit_system_from(s.units).kb;
  (void)create_models(s, ctx);
  require(!(s.cutoff_buffer == 0.f && s.rebuild_check_every != 1),
          "neighbor cutoff_buffer=0 requires rebuild_check_every=1");
  require(s.half_list == 0 || s.half_list == 1, "neighbor half_list must be 0 or 1.");
  // neighbor_list.cu:13-28: slot-major indices are 32-bit.
  if (static_cast<uint64_t>(s.n_particles) * s.max_neighbors > 0xffffffffull) {
    fail(BEADS_E_OVERFLOW, "NeighborList n_particles * max_neighbors exceeds index_t range.");
  }
  require(s.thermo_path == nullptr || s.thermo_every > 0, "output.thermo.every must be positive.");
  require(s.trajectory_path == nullptr || s.trajectory_every > 0,
          "output.trajectory.every must be positive.");
  require(s.log_echo >= 0 && s.log_echo <= 3, "output.log.echo is not supported.");
  require(!((s.log_echo == BEADS_LOG_FILE || s.log_echo == BEADS_LOG_BOTH) && s.log_path == nullptr),
          "output.log.prefix is required for echo=log/both.");
}

Engine::Engine(const beads_spec_t& spec) : spec_(spec) {
  validate_spec(spec);
  if (spec.thermo_path) thermo_path_ = spec.thermo_path;
  if (spec.trajectory_path) trajectory_path_ = spec.trajectory_path;
  if (spec.final_state_path) final_state_path_ = spec.final_state_path;
  if (spec.log_path) log_path_ = spec.log_path;
  // The spec only borrows caller memory for the duration of construction.
  spec_.thermo_path = spec_.trajectory_path = spec_.final_state_path = spec_.log_path = nullptr;
  units_ = unit_system_from(spec.units);
  box_ = make_box_view(spec.box_bound);
  for (int a = 0; a < 3; ++a) spec_box_hi_[a] = spec.box_bound[3 + a];
  n_ = spec.n_particles;
  {
    ModelContext ctx;
    ctx.kb = units_.kb;
    models_ = create_models(spec, ctx);
  }
  thermostat_ = models_.thermostat != nullptr;
  half_list_ = spec.half_list != 0;
  {
    // Diagnostics switch: BEADS_AMD_GRAPHS=0 launches every step's kernels
    // one by one instead of replaying captured step graphs.
    const char* env = std::getenv("BEADS_AMD_GRAPHS");
    use_graphs_ = !(env != nullptr && env[0] == '0');
  }
  has_bond_style_ = models_.bond != nullptr;
  has_angle_style_ = models_.angle != nullptr;
  has_dihedral_style_ = models_.dihedral != nullptr;
  plan_ = make_neighbor_plan(box_, models_.pair->max_cutoff(), spec.cutoff_buffer, n_,
                             spec.max_neighbors,
                             spec.rebuild_check_every, spec.sort_every_rebuild);
  interrupt_interval_ = std::max<uint64_t>(1, 100000 / std::max<uint32_t>(1, n_));

  BEADS_HIP(hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking));
  stream_ = own_stream_;
  upload_state();
  set_up_params();
  // The spec's arrays belong to the caller; drop them.
  spec_.box_bound = nullptr;
  spec_.positions = spec_.velocities = spec_.masses = nullptr;
  spec_.types = nullptr;
  spec_.tags = spec_.molecule_ids = nullptr;
  spec_.images = nullptr;
  spec_.bonds = spec_.angles = spec_.dihedrals = nullptr;
  spec_.lj_epsilon = spec_.lj_sigma = nullptr;
  spec_.bond_k = spec_.bond_r0 = spec_.angle_k = spec_.angle_theta0_deg = nullptr;
  spec_.dihedral_k = nullptr;
  spec_.dihedral_d = spec_.dihedral_n = nullptr;
}

Engine::~Engine() {
  output_.reset();  // the writer thread writes through the streams below
  if (step_graph_exec_ != nullptr) {
    if (stream_ != nullptr) (void)hipStreamSynchronize(stream_);
    (void)hipGraphExecDestroy(step_graph_exec_);
  }
  if (stream_ != own_stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : frame_events_) {
    if (e != nullptr) (void)hipEventDestroy(e);
  }
  if (own_stream_ != nullptr) {
    (void)hipStreamSynchronize(own_stream_);
    (void)hipStreamDestroy(own_stream_);
  }
}

void Engine::upload_state() {
  const beads_spec_t& s = spec_;
  const size_t n = n_;
  std::vector<real4> pos(n), vel(n);
  std::vector<int4> img(n);
  for (size_t i = 0; i < n; ++i) {
    const real_t tw = real_of_id(static_cast<uint32_t>(s.types[i] - 1));
    pos[i] = make_real4(s.positions[3 * i], s.positions[3 * i + 1], s.positions[3 * i + 2], tw);
    vel[i] = make_real4(s.velocities[3 * i], s.velocities[3 * i + 1], s.velocities[3 * i + 2],
                        s.masses[i]);
    img[i] = make_int4(s.images[3 * i], s.images[3 * i + 1], s.images[3 * i + 2], 0);
  }
  pos_.resize(static_cast<size_t>(n) + kCandidatePad);  // the direct build's candidate stream
  pos_.upload(pos.data(), n, stream_);
  vel_.upload(vel.data(), n, stream_);
  image_.upload(img.data(), n, stream_);
  tag_.upload(s.tags, n, stream_);
  molid_.upload(s.molecule_ids, n, stream_);
  force_.resize(n);
  force_.zero(stream_);
  ref_pos_.resize(n);
  ref_image_.resize(n);
  nbr_count_.resize(n);
  nbr_count_.zero(stream_);
  nbr_.resize(static_cast<size_t>(nbr_padded_capacity(plan_.max_neighbors)) * n);
  cell_of_.resize(n);
  cell_atoms_.resize(n);
  cell_pos_.resize(static_cast<size_t>(n) + kCandidatePad);
  cell_count_.resize(plan_.n_cells);
  cell_count_.zero(stream_);
  cell_xbins_.resize(plan_.n_cells);
  cell_xbins_.zero(stream_);
  cell_start_.resize(static_cast<size_t>(plan_.n_cells) + 1);
  cell_start_.zero(stream_);
  cell_cursor_.resize(plan_.n_cells);
  // Until the first binning every atom sits in cell order = slot order.
  std::vector<uint32_t> ident(n);
  std::iota(ident.begin(), ident.end(), 0u);
  cell_atoms_.upload(ident.data(), n, stream_);
  cell_of_.zero(stream_);

  // Listed terms and exclusions, per slot.
  const HostTopology topo = build_host_topology(s);
  excluded_pairs_ = topo.excluded_pairs;
  bcap_ = topo.bcap;
  acap_ = topo.acap;
  dcap_ = topo.dcap;
  ecap_ = topo.ecap;
  has_topology_ = (bcap_ | acap_ | dcap_ | ecap_) != 0;
  topo_count_.upload(topo.count.data(), n, stream_);
  bond_.upload(topo.bond.data(), topo.bond.size(), stream_);
  angle_.upload(topo.angle.data(), topo.angle.size(), stream_);
  dih_.upload(topo.dih.data(), topo.dih.size(), stream_);
  excl_.upload(topo.excl.data(), topo.excl.size(), stream_);
  has_spill_ = !topo.spill.empty();
  spill_excl_any_ = topo.spills_exclusions;
  if (has_spill_) {
    spill_index_.upload(topo.spill_index.data(), n, stream_);
    spill_.upload(topo.spill.data(), topo.spill.size(), stream_);
    spill_bond_.upload(topo.spill_bond.data(), topo.spill_bond.size(), stream_);
    spill_angle_.upload(topo.spill_angle.data(), topo.spill_angle.size(), stream_);
    spill_dih_.upload(topo.spill_dih.data(), topo.spill_dih.size(), stream_);
    spill_excl_.upload(topo.spill_excl.data(), topo.spill_excl.size(), stream_);
  }

  if (plan_.path != 0) {
    s_pos_.resize(n);
    s_vel_.resize(n);
    s_image_.resize(n);
    s_tag_.resize(n);
    s_molid_.resize(n);
    s_cell_of_.resize(n);
    s_topo_count_.resize(n);
    s_bond_.resize(topo.bond.size());
    s_angle_.resize(topo.angle.size());
    s_dih_.resize(topo.dih.size());
    s_excl_.resize(topo.excl.size());
    if (has_spill_) s_spill_index_.resize(n);
    new_of_old_.resize(n);
  }

  n_types_ = models_.pair->type_count();
  n_bond_types_ = models_.bond ? models_.bond->type_count() : 0;
  upload_parameters();

  // Per-workgroup partial sums of the force pass (the widest grid) and of scale.
  partials_.resize(9 * static_cast<size_t>(std::max(force_grid(n_), grid_for(n_))));
  if (half_list_) {
    pair_acc_.resize(3 * n);
    pair_acc_.zero(stream_);
  }
  Ctrl c{};
  c.lambda = 1.0;
  ctrl_.upload(&c, 1, stream_);
  ctrl_host_.resize(1);
  BEADS_HIP(hipStreamSynchronize(stream_));
}

void Engine::set_up_params() {
  const size_t n = n_;
  (void)n;
  TopologyView topo{};
  topo.count = topo_count_.data();
  topo.bond = bond_.data();
  topo.angle = angle_.data();
  topo.dih = dih_.data();
  topo.excl = excl_.data();
  topo.bcap = bcap_;
  topo.acap = acap_;
  topo.dcap = dcap_;
  topo.ecap = ecap_;
  if (models_.bond) models_.bond->bind(topo);
  if (models_.angle) models_.angle->bind(topo);
  if (models_.dihedral) models_.dihedral->bind(topo);
  topo.has_listed = (bcap_ | acap_ | dcap_) != 0;
  topo.spill_index = has_spill_ ? spill_index_.data() : nullptr;
  topo.spill = spill_.data();
  topo.spill_bond = spill_bond_.data();
  topo.spill_angle = spill_angle_.data();
  topo.spill_dih = spill_dih_.data();

  ForceParams& f = force_params_;
  f.n = n_;
  f.pos = pos_.data();
  f.vel = vel_.data();
  f.force = force_.data();
  f.nbr_count = nbr_count_.data();
  f.nbr = nbr_.data();
  f.box = box_;
  models_.pair->bind(f);
  f.topo = topo;
  f.berendsen = 0;
  f.target_thermal = 0.0;
  f.dt_over_tau = 0.0;
  models_.integrator->bind(integrate_params_, f);
  if (models_.thermostat) models_.thermostat->bind(f, models_.integrator->dt());
  f.dof = std::max(0.0, 3.0 * static_cast<double>(n_) - 3.0);
  f.partials = partials_.data();
  f.ctrl = ctrl_.data();
  f.cell_count = cell_count_.data();
  f.n_cells = plan_.n_cells;
  f.half = half_list_ ? 1 : 0;
  f.pair_acc = pair_acc_.data();

  IntegrateParams& in = integrate_params_;
  in.n = n_;
  in.pos = pos_.data();
  in.vel = vel_.data();
  in.force = force_.data();
  in.image = image_.data();
  in.ref_pos = ref_pos_.data();
  in.ref_image = ref_image_.data();
  in.box = box_;
  in.dt = models_.integrator->dt();
  const real_t half_buffer = real_t(0.5) * plan_.cutoff_buffer;
  in.threshold_sq = half_buffer * half_buffer;
  in.ctrl = ctrl_.data();

  CellGridView grid{};
  grid.nx = plan_.nx;
  grid.ny = plan_.ny;
  grid.nz = plan_.nz;
  grid.n_cells = plan_.n_cells;
  for (int a = 0; a < 3; ++a) grid.inv_cell[a] = plan_.inv_cell[a];
  grid.path = plan_.path;

  BuildParams& b = build_params_;
  b.n = n_;
  b.pos = pos_.data();
  b.image = image_.data();
  b.ref_pos = ref_pos_.data();
  b.ref_image = ref_image_.data();
  b.box = box_;
  b.grid = grid;
  b.rs2 = plan_.search_cutoff_sq;
  b.cell_of = cell_of_.data();
  b.cell_start = cell_start_.data();
  b.cell_count = cell_count_.data();
  b.cell_atoms = cell_atoms_.data();
  b.cell_pos = cell_pos_.data();
  b.cell_xbins = cell_xbins_.data();
  b.topo_count = topo_count_.data();
  b.excl = excl_.data();
  b.ecap = ecap_;
  b.spill_index = spill_excl_any_ ? spill_index_.data() : nullptr;
  b.spill = spill_.data();
  b.spill_excl = spill_excl_.data();
  b.max_neighbors = plan_.max_neighbors;
  b.nbr_count = nbr_count_.data();
  b.nbr = nbr_.data();
  b.ctrl = ctrl_.data();
  b.half = half_list_ ? 1 : 0;

  BinParams& bp = bin_params_;
  bp.n = n_;
  bp.pos = pos_.data();
  bp.box = box_;
  bp.grid = grid;
  bp.cell_of = cell_of_.data();
  bp.cell_count = cell_count_.data();
  bp.ctrl = ctrl_.data();

  CellScanParams& sp = scan_params_;
  sp.n = n_;
  sp.n_cells = plan_.n_cells;
  sp.cell_of = cell_of_.data();
  sp.cell_count = cell_count_.data();
  sp.cell_start = cell_start_.data();
  sp.cell_cursor = cell_cursor_.data();
  sp.cell_atoms = cell_atoms_.data();
  sp.pos = pos_.data();
  sp.cell_pos = cell_pos_.data();
  sp.cell_xbins = cell_xbins_.data();
  sp.box = box_;
  sp.grid = grid;
  sp.ctrl = ctrl_.data();

  ReorderParams& r = reorder_params_;
  r.n = n_;
  r.pos = pos_.data();
  r.vel = vel_.data();
  r.image = image_.data();
  r.tag = tag_.data();
  r.molid = molid_.data();
  r.cell_of = cell_of_.data();
  r.topo_count = topo_count_.data();
  r.bond = bond_.data();
  r.angle = angle_.data();
  r.dih = dih_.data();
  r.excl = excl_.data();
  r.s_pos = s_pos_.data();
  r.s_vel = s_vel_.data();
  r.s_image = s_image_.data();
  r.s_tag = s_tag_.data();
  r.s_molid = s_molid_.data();
  r.s_cell_of = s_cell_of_.data();
  r.s_topo_count = s_topo_count_.data();
  r.s_bond = s_bond_.data();
  r.s_angle = s_angle_.data();
  r.s_dih = s_dih_.data();
  r.s_excl = s_excl_.data();
  r.bcap = bcap_;
  r.acap = acap_;
  r.dcap = dcap_;
  r.ecap = ecap_;
  r.spill_index = has_spill_ ? spill_index_.data() : nullptr;
  r.s_spill_index = s_spill_index_.data();
  r.spill = spill_.data();
  r.spill_bond = spill_bond_.data();
  r.spill_angle = spill_angle_.data();
  r.spill_dih = spill_dih_.data();
  r.spill_excl = spill_excl_.data();
  r.cell_atoms = cell_atoms_.data();
  r.cell_pos = cell_pos_.data();
  r.new_of_old = new_of_old_.data();
  r.ctrl = ctrl_.data();

  ScaleParams& sc = scale_params_;
  sc.n = n_;
  sc.vel = vel_.data();
  sc.partials = partials_.data();
  sc.ctrl = ctrl_.data();

  control_.check_every = plan_.check_every;
  control_.sort_every = plan_.sort_every;
  control_.sort_every_steps = spec_.sort_every_steps;
  control_.path = plan_.path;
}

// NeighborSystem::rebuild_now (neighbor_system.cpp:107-146) as a predicated
// launch sequence.
void Engine::rebuild_sequence(bool forced, uint64_t steps) {
  launch_control(ctrl_.data(), control_, forced, steps, stream_);
  if (plan_.path != 0) {
    launch_bin(bin_params_, stream_);
    launch_cell_scan(scan_params_, stream_);
    launch_cell_scatter(scan_params_, stream_);
    launch_cell_sort(scan_params_, stream_);
    launch_reorder(reorder_params_, stream_);
  }
  launch_build(build_params_, stream_);
}

void Engine::force_pass(bool kick, bool observables) {
  launch_force(force_params_, kick, observables, stream_);
}

void Engine::initialize(bool observables) {
  require(!frame_mode_, "an engine used in frame mode cannot run dynamics.", BEADS_E_RUNTIME);
  rebuild_sequence(true, 0);
  force_pass(false, observables);
  BEADS_HIP(hipGetLastError());
  ke_from_scale_ = false;
  initialized_ = true;
}

// SimulationRun::advance_dynamics_step (simulation_run.cpp:212-235): pending
// thermostat factor + pre-force, neighbour update, forces + post-force +
// Berendsen factor.
void Engine::step_once(bool observables) {
  const uint64_t physical = step_ + 1;
  const bool every = plan_.cutoff_buffer == 0.f;
  const bool check = !every && physical % plan_.check_every == 0;
  launch_integrate(integrate_params_, thermostat_, check, false, stream_);
  if (every || check) {
    rebuild_sequence(every, every ? uint64_t{1} : plan_.check_every);
  }
  force_pass(true, observables);
  ke_from_scale_ = false;
  step_ = physical;
}

// Steps per captured graph: whole check periods, at least 16 steps, so a
// replay always starts on a step whose check pattern is the captured one.
uint64_t Engine::graph_steps() const {
  const uint64_t period = plan_.cutoff_buffer == 0.f ? 1 : plan_.check_every;
  return period * ((16 + period - 1) / period);
}

// One hipGraph of graph_steps() consecutive steps without observables,
// captured from the same launch sequence step_once issues (no memsets, no
// host synchronisation inside), instantiated once and replayed.
void Engine::replay_step_graph() {
  const uint64_t g = graph_steps();
  if (step_graph_exec_ == nullptr) {
    const uint64_t saved = step_;
    BEADS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (uint64_t k = 0; k < g; ++k) step_once(false);
    hipGraph_t graph = nullptr;
    BEADS_HIP(hipStreamEndCapture(stream_, &graph));
    BEADS_HIP(hipGraphInstantiate(&step_graph_exec_, graph, nullptr, nullptr, 0));
    BEADS_HIP(hipGraphDestroy(graph));
    step_ = saved;  // capture recorded the launches without running them
  }
  BEADS_HIP(hipGraphLaunch(step_graph_exec_, stream_));
  step_ += g;
  ke_from_scale_ = false;
}

void Engine::advance(uint64_t n_steps, bool observables_on_last) {
  require(initialized_, "engine must be initialized before advancing.", BEADS_E_RUNTIME);
  const uint64_t g = graph_steps();
  uint64_t left = n_steps;
  while (left > 0) {
    // Whole graphs while they fit and leave the observable step to step_once.
    const uint64_t reserve = observables_on_last ? 1 : 0;
    if (use_graphs_ && step_ % g == 0 && left >= g + reserve) {
      replay_step_graph();
      left -= g;
      continue;
    }
    step_once(observables_on_last && left == 1);
    --left;
  }
  BEADS_HIP(hipGetLastError());
}

void Engine::materialize_velocities() {
  launch_scale(scale_params_, thermostat_, stream_);
  ke_from_scale_ = true;
}

// Waits for the enqueued work; in frame mode also reports a neighbour-list
// overflow that async frame calls left unchecked.
void Engine::synchronize() {
  BEADS_HIP(hipStreamSynchronize(stream_));
  if (frame_mode_) check_overflow(fetch_ctrl());
}

Ctrl Engine::fetch_ctrl() {
  BEADS_HIP(hipMemcpyAsync(ctrl_host_.data(), ctrl_.data(), sizeof(Ctrl), hipMemcpyDeviceToHost,
                           stream_));
  BEADS_HIP(hipStreamSynchronize(stream_));
  return *ctrl_host_.data();
}

void Engine::check_overflow(const Ctrl& c) {
  if (c.overflow != 0) {
    std::ostringstream m;
    m << "NeighborList overflow"
      << "\npath=" << (plan_.path == 0 ? "direct" : (plan_.path == 1 ? "regular_cell" : "small_dimension"))
      << "\nn_particles=" << n_ << "\nmax_neighbors=" << plan_.max_neighbors
      << "\nforce_cutoff=" << plan_.force_cutoff << "\ncutoff_buffer=" << plan_.cutoff_buffer
      << "\nsearch_cutoff=" << plan_.search_cutoff
      << ". Increase max_neighbors before consuming this neighbor list.";
    fail(BEADS_E_RUNTIME, m.str());
  }
}

void Engine::stats(beads_engine_stats_t* out) {
  const Ctrl c = fetch_ctrl();
  std::memset(out, 0, sizeof(*out));
  out->step = step_;
  out->build_count = c.build_count;
  out->dangerous_rebuild_count = c.dangerous;
  out->reorder_count = c.reorder_count;
  out->overflow = static_cast<int32_t>(c.overflow);
  out->path = plan_.path;
  out->nx = plan_.nx;
  out->ny = plan_.ny;
  out->nz = plan_.nz;
  out->max_neighbors = plan_.max_neighbors;
  out->mean_neighbors = static_cast<double>(c.neighbor_total) / n_;
  // Kinetic energy of the velocities as stored (before any pending factor).
  out->ke = ke_from_scale_ ? c.ke : c.ke_pre;
  out->lambda = c.lambda;
  for (int q = 0; q < 4; ++q) {
    out->pe[q] = c.pe[q];
    out->virial[q] = c.virial[q];
  }
  const double dof = std::max(0.0, 3.0 * n_ - 3.0);
  out->temperature = dof > 0 ? 2.0 * out->ke / (dof * units_.kb) : 0.0;
  const double vol = static_cast<double>(box_.L[0]) * box_.L[1] * box_.L[2];
  const double w = c.virial[0] + c.virial[1] + c.virial[2] + c.virial[3];
  out->pressure = units_.pressure_scale * (2.0 * out->ke + w) / (3.0 * vol);
}

void Engine::read_state(const beads_state_out_t& o) {
  const size_t n = n_;
  std::vector<real4> pos(n), vel(n), force(n);
  std::vector<int4> img(n);
  pos_.download(pos.data(), n, stream_);
  vel_.download(vel.data(), n, stream_);
  force_.download(force.data(), n, stream_);
  image_.download(img.data(), n, stream_);
  if (o.tags) tag_.download(o.tags, n, stream_);
  if (o.molecule_ids) molid_.download(o.molecule_ids, n, stream_);
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

void Engine::read_neighbors(uint32_t* counts, uint32_t* list, uint64_t capacity) {
  require(capacity >= static_cast<uint64_t>(plan_.max_neighbors) * n_,
          "neighbor list capacity must be max_neighbors * n_particles.");
  // The device keeps 16-entry chunks per atom; hand out the reference's
  // slot-major layout (entry k of atom i at k * N + i, neighbor_list.cuh:12-19).
  const size_t n = n_;
  const uint32_t padded = nbr_padded_capacity(plan_.max_neighbors);
  std::vector<uint32_t> chunked(static_cast<size_t>(padded) * n);
  nbr_count_.download(counts, n_, stream_);
  nbr_.download(chunked.data(), chunked.size(), stream_);
  synchronize();
  for (size_t i = 0; i < n; ++i) {
    for (uint32_t k = 0; k < counts[i] && k < plan_.max_neighbors; ++k) {
      list[static_cast<size_t>(k) * n + i] =
          chunked[((k / kNbrChunk) * n + i) * kNbrChunk + (k % kNbrChunk)];
    }
  }
}

void Engine::read_cells(uint32_t* cell_of, uint32_t* cell_atoms, uint32_t* cell_start,
                        uint32_t* cell_count, uint64_t capacity) {
  require(capacity >= plan_.n_cells, "cell capacity must be at least n_cells.");
  std::vector<uint32_t> start(static_cast<size_t>(plan_.n_cells) + 1);
  cell_of_.download(cell_of, n_, stream_);
  cell_atoms_.download(cell_atoms, n_, stream_);
  cell_start_.download(start.data(), start.size(), stream_);
  synchronize();
  for (uint32_t c = 0; c < plan_.n_cells; ++c) {
    cell_start[c] = start[c];
    cell_count[c] = start[c + 1] - start[c];
  }
}

void Engine::profile_advance(uint64_t n_steps, bool per_phase, beads_phase_times_t* out) {
  require(initialized_, "engine must be initialized before advancing.", BEADS_E_RUNTIME);
  std::memset(out, 0, sizeof(*out));
  hipEvent_t t0, t1;
  BEADS_HIP(hipEventCreate(&t0));
  BEADS_HIP(hipEventCreate(&t1));
  std::vector<hipEvent_t> ev;
  if (per_phase) {
    ev.resize(4 * n_steps);
    for (auto& e : ev) BEADS_HIP(hipEventCreate(&e));
  }
  BEADS_HIP(hipEventRecord(t0, stream_));
  if (!per_phase) advance(n_steps, false);
  for (uint64_t k = 0; k < n_steps && per_phase; ++k) {
    const uint64_t physical = step_ + 1;
    const bool every = plan_.cutoff_buffer == 0.f;
    const bool check = !every && physical % plan_.check_every == 0;
    BEADS_HIP(hipEventRecord(ev[4 * k], stream_));
    launch_integrate(integrate_params_, thermostat_, check, false, stream_);
    BEADS_HIP(hipEventRecord(ev[4 * k + 1], stream_));
    if (every || check) {
      rebuild_sequence(every, every ? uint64_t{1} : plan_.check_every);
    }
    BEADS_HIP(hipEventRecord(ev[4 * k + 2], stream_));
    force_pass(true, false);
    BEADS_HIP(hipEventRecord(ev[4 * k + 3], stream_));
    step_ = physical;
  }
  BEADS_HIP(hipEventRecord(t1, stream_));
  BEADS_HIP(hipEventSynchronize(t1));
  BEADS_HIP(hipGetLastError());
  float ms = 0.f;
  BEADS_HIP(hipEventElapsedTime(&ms, t0, t1));
  out->steps = n_steps;
  out->total_s = ms * 1e-3;
  for (uint64_t k = 0; k < n_steps && per_phase; ++k) {
    float a = 0, b = 0, c = 0;
    BEADS_HIP(hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 1]));
    BEADS_HIP(hipEventElapsedTime(&b, ev[4 * k + 1], ev[4 * k + 2]));
    BEADS_HIP(hipEventElapsedTime(&c, ev[4 * k + 2], ev[4 * k + 3]));
    out->integrate_s += a * 1e-3;
    out->rebuild_s += b * 1e-3;
    out->force_s += c * 1e-3;
    out->force_launches += 1;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
}

// -------------------------------------------------------------- frame mode --

// Binary32 tables as the forward uses them ({sigma^2, 24 eps, 4 eps, shift},
// {k, r0}) plus FP64 copies of the same real_t values for the adjoint.
void Engine::upload_parameters() {
  models_.pair->upload(stream_);
  if (models_.bond) models_.bond->upload(stream_);
  if (models_.angle) models_.angle->upload(stream_);
  if (models_.dihedral) models_.dihedral->upload(stream_);
  BEADS_HIP(hipStreamSynchronize(stream_));
}

void Engine::set_lj_params(const real_t* eps, const real_t* sigma) {
  models_.pair->set_type_pair_parameters(eps, sigma);
  upload_parameters();
}

void Engine::set_bond_params(const real_t* k, const real_t* r0) {
  require(models_.bond != nullptr, "this engine has no bond style.", BEADS_E_VALUE);
  models_.bond->set_parameters(k, r0);
  upload_parameters();
}

void Engine::set_angle_params(const real_t* k, const real_t* theta0_deg) {
  require(models_.angle != nullptr, "this engine has no angle style.", BEADS_E_VALUE);
  models_.angle->set_parameters(k, theta0_deg);
  upload_parameters();
}

void Engine::set_dihedral_params(const real_t* k) {
  require(models_.dihedral != nullptr, "this engine has no dihedral style.", BEADS_E_VALUE);
  models_.dihedral->set_parameters(k, nullptr);
  upload_parameters();
}

size_t Engine::adjoint_param_count() const {
  const size_t na = models_.angle ? static_cast<size_t>(models_.angle->type_count()) : 0;
  const size_t nd = models_.dihedral ? static_cast<size_t>(models_.dihedral->type_count()) : 0;
  return 2 * static_cast<size_t>(n_types_) * n_types_ + 2 * static_cast<size_t>(n_bond_types_) +
         2 * na + nd;
}

// One force evaluation at caller-supplied positions ([n][3] real_t in slot
// order, device memory): wrap into the box, neighbour rebuild without
// reordering (forced, or with reuse_list only when some atom moved more than
// half the skin since the last build, the integrator's test), forces +
// energies. Frame mode and dynamics are exclusive.
void Engine::enter_frame_mode() {
  require(!initialized_, "frame mode is not available on an engine that runs dynamics.",
          BEADS_E_RUNTIME);
  if (!frame_mode_) {
    for (auto& e : frame_events_) BEADS_HIP(hipEventCreate(&e));
    frame_mode_ = true;
    control_.path = 0;  // never reorder: slots stay the caller's order
  }
}

void Engine::load_frame(const real_t* positions_dev, bool reuse_list) {
  enter_frame_mode();
  // The list may be kept only if one was built in frame mode and the last
  // call completed (so its reference positions are the list's).
  const bool check = reuse_list && frame_list_valid_ && plan_.cutoff_buffer > 0.f;
  frame_list_valid_ = false;
  FrameLoadParams lp{};
  lp.n = n_;
  lp.xyz = positions_dev;
  lp.pos = pos_.data();
  lp.image = image_.data();
  lp.ref_pos = ref_pos_.data();
  lp.ref_image = ref_image_.data();
  lp.box = box_;
  lp.threshold_sq = integrate_params_.threshold_sq;
  lp.ctrl = ctrl_.data();
  launch_frame_load(lp, check, stream_);
  rebuild_sequence(!check, 0);
}

void Engine::frame_forces(const real_t* positions_dev, real_t* forces_dev, double* energies_out,
                          bool reuse_list, bool async) {
  enter_frame_mode();
  frame_loaded_ = false;
  BEADS_HIP(hipEventRecord(frame_events_[0], stream_));
  load_frame(positions_dev, reuse_list);
  force_pass(false, true);
  launch_frame_store(n_, force_.data(), forces_dev, stream_);
  BEADS_HIP(hipEventRecord(frame_events_[1], stream_));
  BEADS_HIP(hipGetLastError());
  if (async) {
    if (energies_out != nullptr) {
      const char* pe = reinterpret_cast<const char*>(ctrl_.data()) + offsetof(Ctrl, pe);
      BEADS_HIP(hipMemcpyAsync(energies_out, pe, 4 * sizeof(double), hipMemcpyDeviceToDevice,
                               stream_));
    }
    forward_time_pending_ = true;
  } else {
    const Ctrl c = fetch_ctrl();
    BEADS_HIP(hipEventElapsedTime(&last_forward_ms_, frame_events_[0], frame_events_[1]));
    forward_time_pending_ = false;
    check_overflow(c);
    if (energies_out != nullptr) {
      for (int q = 0; q < 4; ++q) energies_out[q] = c.pe[q];
    }
  }
  frame_loaded_ = true;
  frame_list_valid_ = true;
}

// Later frame calls enqueue on the caller's stream, taken as given: NULL is
// HIP's null stream (torch's default stream), not the engine's own. What is
// already enqueued on the current stream completes first, so the switch keeps
// every ordering.
void Engine::set_stream(hipStream_t stream) {
  enter_frame_mode();
  if (stream == stream_) return;
  BEADS_HIP(hipStreamSynchronize(stream_));
  stream_ = stream;
}

// Positions and neighbour list of a frame without evaluating forces: what
// frame_backward needs when another frame was evaluated in between.
void Engine::frame_load(const real_t* positions_dev, bool reuse_list, bool async) {
  frame_loaded_ = false;
  load_frame(positions_dev, reuse_list);
  // The force pass normally clears the rebuild flags; no force pass runs here.
  BEADS_HIP(hipMemsetAsync(ctrl_.data(), 0, 2 * sizeof(uint32_t), stream_));
  BEADS_HIP(hipGetLastError());
  if (!async) check_overflow(fetch_ctrl());
  frame_loaded_ = true;
  frame_list_valid_ = true;
}

// Adjoint of the last frame_forces call: grad_forces [n][3] real_t (device),
// grad_energy scales dE/d(.), outputs dL/dx [n][3] real_t and the FP64
// parameter gradients [eps T*T | sigma T*T | k nb | r0 nb] (device).
void Engine::frame_backward(const real_t* grad_forces_dev, double grad_energy,
                            real_t* grad_positions_dev, double* grad_params_dev, bool async,
                            bool noalloc) {
  require(frame_loaded_, "frame_backward needs a preceding frame_forces call.", BEADS_E_RUNTIME);
  require(!half_list_, "the adjoint reads the full neighbour list; create the engine with "
          "half_list = 0.", BEADS_E_VALUE);
  AdjointParams p{};
  models_.pair->bind_adjoint(p);
  if (models_.bond) models_.bond->bind_adjoint(p);
  if (models_.angle) models_.angle->bind_adjoint(p);
  if (models_.dihedral) models_.dihedral->bind_adjoint(p);
  const size_t n_params = adjoint_param_count();
  real4* staging = nullptr;
  if (noalloc) {
    require(n_ <= kNoAllocMaxAtoms,
            "frame_backward_noalloc serves at most " + std::to_string(kNoAllocMaxAtoms) +
                " particles (its static scratch); this engine has " + std::to_string(n_) + ".",
            BEADS_E_VALUE);
    staging = scratch_address();
    async = true;
  } else {
    grad_force4_.resize(n_);
    staging = grad_force4_.data();
  }
  if (noalloc) {
    // no timing events either: nothing but kernel launches on the stream
    launch_fill_f64(grad_params_dev, static_cast<uint32_t>(n_params), stream_);
  } else {
    BEADS_HIP(hipEventRecord(frame_events_[2], stream_));
    BEADS_HIP(hipMemsetAsync(grad_params_dev, 0, n_params * sizeof(double), stream_));
  }
  launch_pack_xyz(n_, grad_forces_dev, staging, stream_);
  p.n = n_;
  p.pos = pos_.data();
  p.g = staging;
  p.nbr_count = nbr_count_.data();
  p.nbr = nbr_.data();
  p.box = box_;
  p.topo_count = topo_count_.data();
  p.bond = bond_.data();
  p.bcap = bcap_;
  p.spill_index = has_spill_ ? spill_index_.data() : nullptr;
  p.spill = spill_.data();
  p.spill_bond = spill_bond_.data();
  p.angle = angle_.data();
  p.dih = dih_.data();
  p.acap = acap_;
  p.dcap = dcap_;
  p.spill_angle = spill_angle_.data();
  p.spill_dih = spill_dih_.data();
  p.grad_energy = grad_energy;
  p.grad_pos = grad_positions_dev;
  p.grad_params = grad_params_dev;
  launch_fm_backward(p, stream_);
  BEADS_HIP(hipGetLastError());
  if (noalloc) return;
  BEADS_HIP(hipEventRecord(frame_events_[3], stream_));
  BEADS_HIP(hipGetLastError());
  if (async) {
    backward_time_pending_ = true;
  } else {
    BEADS_HIP(hipEventSynchronize(frame_events_[3]));
    BEADS_HIP(hipEventElapsedTime(&last_backward_ms_, frame_events_[2], frame_events_[3]));
    backward_time_pending_ = false;
  }
}

// Device time of the last forward and backward frame calls; after async calls
// this waits for their end events.
void Engine::frame_timing(double* forward_s, double* backward_s) {
  if (forward_time_pending_) {
    BEADS_HIP(hipEventSynchronize(frame_events_[1]));
    BEADS_HIP(hipEventElapsedTime(&last_forward_ms_, frame_events_[0], frame_events_[1]));
    forward_time_pending_ = false;
  }
  if (backward_time_pending_) {
    BEADS_HIP(hipEventSynchronize(frame_events_[3]));
    BEADS_HIP(hipEventElapsedTime(&last_backward_ms_, frame_events_[2], frame_events_[3]));
    backward_time_pending_ = false;
  }
  if (forward_s != nullptr) *forward_s = last_forward_ms_ * 1e-3;
  if (backward_s != nullptr) *backward_s = last_backward_ms_ * 1e-3;
}

// ----------------------------------------------------------------- outputs --

bool Engine::output_due(uint64_t step) const {
  const bool thermo = !thermo_path_.empty() && step % spec_.thermo_every == 0;
  const bool traj = !trajectory_path_.empty() && step % spec_.trajectory_every == 0;
  return thermo || traj;
}

// Synchronous: read back and write at once. Asynchronous: pack the records on
// the device, stage them with the control block in a pinned slot and let the
// writer thread write; the compute stream never waits on the host.
void Engine::stage_outputs(uint64_t step) {
  const bool thermo = !thermo_path_.empty() && step % spec_.thermo_every == 0;
  const bool traj = !trajectory_path_.empty() && step % spec_.trajectory_every == 0;
  materialize_velocities();
  if (!async_output_) {
    const Ctrl c = fetch_ctrl();
    check_overflow(c);
    if (thermo) write_thermo_row(step, c);
    if (traj) write_trajectory_frame(step);
    return;
  }
  if (!output_) {
    const size_t bytes = trajectory_path_.empty() ? 0 : trajectory_record_bytes() * n_;
    output_ = std::make_unique<AsyncOutput>(
        bytes, [this](uint64_t s, bool t, bool j, const Ctrl& c, const unsigned char* r) {
          write_outputs(s, t, j, c, r);
        });
  }
  const int slot = output_->acquire();
  try {
    if (traj) {
      PackParams pp{};
      pp.n = n_;
      pp.mask = spec_.trajectory_fields | BEADS_TRAJ_TAG;
      pp.words = static_cast<uint32_t>(trajectory_record_bytes() / sizeof(uint32_t));
      pp.pos = pos_.data();
      pp.vel = vel_.data();
      pp.force = force_.data();
      pp.image = image_.data();
      pp.tag = tag_.data();
      pp.out = reinterpret_cast<uint32_t*>(output_->device_records(slot));
      launch_pack_records(pp, stream_);
      BEADS_HIP(hipGetLastError());
    }
    output_->submit(slot, ctrl_.data(), stream_, step, thermo, traj);
  } catch (...) {
    output_->cancel(slot);
    throw;
  }
}

// Writer side of an asynchronous output step (also the order of the
// synchronous one): an overflowed list stops the run before anything of the
// step is written.
void Engine::write_outputs(uint64_t step, bool thermo, bool trajectory, const Ctrl& c,
                           const unsigned char* records) {
  check_overflow(c);
  if (thermo) write_thermo_row(step, c);
  if (trajectory) write_trajectory_records(step, records);
}

// thermo_output.cu:437-509.
void Engine::write_thermo_row(uint64_t step, const Ctrl& c) {
  if (!thermo_stream_.is_open()) {
    thermo_stream_.open(thermo_path_);
    if (!thermo_stream_) {
      fail(BEADS_E_RUNTIME, "failed to open thermo output file " + thermo_path_ + ".");
    }
    thermo_stream_ << "step,pe,pair_pe";
    if (has_bond_style_) thermo_stream_ << ",bond_pe";
    if (has_angle_style_) thermo_stream_ << ",angle_pe";
    if (has_dihedral_style_) thermo_stream_ << ",dihedral_pe";
    thermo_stream_ << ",ke,temp,press\n";
  }
  const real_t pair = static_cast<real_t>(c.pe[0]);
  const real_t bond = has_bond_style_ ? static_cast<real_t>(c.pe[1]) : real_t(0);
  const real_t angle = has_angle_style_ ? static_cast<real_t>(c.pe[2]) : real_t(0);
  const real_t dih = has_dihedral_style_ ? static_cast<real_t>(c.pe[3]) : real_t(0);
  const real_t pe = pair + bond + angle + dih;
  const double dof = std::max(0.0, 3.0 * n_ - 3.0);
  const double temp = dof > 0 ? 2.0 * c.ke / (dof * units_.kb) : 0.0;
  const double vol = static_cast<double>(box_.L[0]) * box_.L[1] * box_.L[2];
  const double w = c.virial[0] + c.virial[1] + c.virial[2] + c.virial[3];
  const double press = vol > 0 ? units_.pressure_scale * (2.0 * c.ke + w) / (3.0 * vol) : 0.0;
  std::ostringstream row;
  row << step << ',' << fmt9(pe) << ',' << fmt9(pair);
  if (has_bond_style_) row << ',' << fmt9(bond);
  if (has_angle_style_) row << ',' << fmt9(angle);
  if (has_dihedral_style_) row << ',' << fmt9(dih);
  row << ',' << fmt9(c.ke) << ',' << fmt9(temp) << ',' << fmt9(press);
  thermo_stream_ << row.str() << '\n';
  if (!thermo_stream_) {
    fail(BEADS_E_RUNTIME, "failed to write thermo output file " + thermo_path_ + ".");
  }
  if (spec_.log_echo != BEADS_LOG_NONE) {
    const int w = kThermoColumnWidth;
    if (!thermo_header_logged_) {
      std::ostringstream h;
      h << std::setw(w) << "Step" << std::setw(w) << "PotEng" << std::setw(w) << "PairEng";
      if (has_bond_style_) h << std::setw(w) << "BondEng";
      if (has_angle_style_) h << std::setw(w) << "AngleEng";
      if (has_dihedral_style_) h << std::setw(w) << "DihedEng";
      h << std::setw(w) << "KinEng" << std::setw(w) << "Temp" << std::setw(w) << "Press";
      log("THERMO", h.str());
      thermo_header_logged_ = true;
    }
    std::ostringstream t;
    t << std::setw(w) << step << std::setw(w) << fmt9(pe) << std::setw(w) << fmt9(pair);
    if (has_bond_style_) t << std::setw(w) << fmt9(bond);
    if (has_angle_style_) t << std::setw(w) << fmt9(angle);
    if (has_dihedral_style_) t << std::setw(w) << fmt9(dih);
    t << std::setw(w) << fmt9(c.ke) << std::setw(w) << fmt9(temp) << std::setw(w) << fmt9(press);
    log("THERMO", t.str());
  }
}

// Bytes of one trajectory record (trajectory_output.cu:33-46).
size_t Engine::trajectory_record_bytes() const {
  const uint32_t mask = spec_.trajectory_fields | BEADS_TRAJ_TAG;
  size_t rec = sizeof(uint32_t);
  if (mask & BEADS_TRAJ_TYPE) rec += sizeof(int32_t);
  if (mask & BEADS_TRAJ_POSITION) rec += 3 * sizeof(real_t);
  if (mask & BEADS_TRAJ_IMAGE) rec += 3 * sizeof(int32_t);
  if (mask & BEADS_TRAJ_VELOCITY) rec += 3 * sizeof(real_t);
  if (mask & BEADS_TRAJ_FORCE) rec += 3 * sizeof(real_t);
  return rec;
}

// trajectory_output.cu:148-250, 555-595 (BEADS_TRAJECTORY v1): the header on
// first use, then one frame of n records already in file layout.
void Engine::write_trajectory_records(uint64_t step, const unsigned char* records) {
  const uint32_t mask = spec_.trajectory_fields | BEADS_TRAJ_TAG;
  const size_t rec = trajectory_record_bytes();
  if (!trajectory_stream_.is_open()) {
    trajectory_stream_.open(trajectory_path_, std::ios::binary | std::ios::trunc);
    if (!trajectory_stream_) {
      fail(BEADS_E_RUNTIME, "failed to open trajectory output file " + trajectory_path_ + ".");
    }
    auto& o = trajectory_stream_;
    write_raw(o, "BEADS_TRAJECTORY", 16, "trajectory magic");
    write_scalar<uint32_t>(o, 1, "trajectory version");
    write_scalar<uint32_t>(o, 0x01020304u, "trajectory endian marker");
    write_scalar<uint32_t>(o, sizeof(real_t), "trajectory real size");
    write_scalar<uint32_t>(o, sizeof(uint32_t), "trajectory index size");
    write_scalar<uint32_t>(o, sizeof(int32_t), "trajectory image size");
    write_scalar<uint32_t>(o, sizeof(int32_t), "trajectory type size");
    write_scalar<uint32_t>(o, sizeof(uint32_t), "trajectory tag size");
    write_scalar<uint64_t>(o, units_.name.size(), "trajectory units length");
    write_raw(o, units_.name.data(), units_.name.size(), "trajectory units");
    write_scalar<uint32_t>(o, mask, "trajectory field mask");
    write_scalar<uint64_t>(o, n_, "trajectory particle count");
    write_scalar<uint64_t>(o, rec, "trajectory record bytes");
  }
  auto& o = trajectory_stream_;
  write_scalar<uint64_t>(o, step, "trajectory frame step");
  const real_t lo[3] = {box_.lo[0], box_.lo[1], box_.lo[2]};
  const real_t hi[3] = {spec_box_hi_[0], spec_box_hi_[1], spec_box_hi_[2]};
  write_raw(o, lo, 3, "trajectory box lower");
  write_raw(o, hi, 3, "trajectory box upper");
  write_raw(o, records, rec * n_, "trajectory records");
}

// Synchronous frame: read the state back and lay the records out on the host.
void Engine::write_trajectory_frame(uint64_t step) {
  const uint32_t mask = spec_.trajectory_fields | BEADS_TRAJ_TAG;
  const size_t rec = trajectory_record_bytes();
  const size_t n = n_;
  std::vector<real_t> pos(3 * n), vel(3 * n), force(3 * n);
  std::vector<int32_t> img(3 * n), type(n);
  std::vector<uint32_t> tag(n);
  beads_state_out_t out{};
  out.positions = pos.data();
  out.velocities = vel.data();
  out.forces = force.data();
  out.images = img.data();
  out.tags = tag.data();
  out.types = type.data();
  read_state(out);
  std::vector<unsigned char> buf(rec * n);
  for (size_t i = 0; i < n; ++i) {
    unsigned char* r = buf.data() + i * rec;
    size_t off = 0;
    auto put = [&](const void* src, size_t bytes) {
      std::memcpy(r + off, src, bytes);
      off += bytes;
    };
    put(&tag[i], 4);
    if (mask & BEADS_TRAJ_TYPE) put(&type[i], 4);
    if (mask & BEADS_TRAJ_POSITION) put(&pos[3 * i], 3 * sizeof(real_t));
    if (mask & BEADS_TRAJ_IMAGE) put(&img[3 * i], 12);
    if (mask & BEADS_TRAJ_VELOCITY) put(&vel[3 * i], 3 * sizeof(real_t));
    if (mask & BEADS_TRAJ_FORCE) put(&force[3 * i], 3 * sizeof(real_t));
  }
  write_trajectory_records(step, buf.data());
}

// final_state_output.cu:103-185 (BEADS_SNAPSHOT v1, SoA in slot order).
void Engine::write_final_state() {
  const size_t n = n_;
  std::vector<real_t> pos(3 * n), vel(3 * n), mass(n);
  std::vector<int32_t> img(3 * n), type(n);
  std::vector<uint32_t> tag(n), molid(n);
  beads_state_out_t out{};
  out.positions = pos.data();
  out.velocities = vel.data();
  out.images = img.data();
  out.tags = tag.data();
  out.types = type.data();
  out.molecule_ids = molid.data();
  out.masses = mass.data();
  read_state(out);
  std::ofstream o(final_state_path_, std::ios::binary | std::ios::trunc);
  if (!o) {
    fail(BEADS_E_RUNTIME, "failed to open final-state output file " + final_state_path_ + ".");
  }
  write_raw(o, "BEADS_SNAPSHOT", 14, "snapshot magic");
  write_scalar<uint32_t>(o, 1, "snapshot version");
  write_scalar<uint32_t>(o, sizeof(real_t), "snapshot real size");
  write_scalar<uint32_t>(o, sizeof(uint32_t), "snapshot index size");
  write_scalar<uint32_t>(o, sizeof(int32_t), "snapshot image size");
  write_scalar<uint32_t>(o, sizeof(int32_t), "snapshot type size");
  write_scalar<uint32_t>(o, sizeof(uint32_t), "snapshot tag size");
  write_scalar<uint32_t>(o, 0x01020304u, "snapshot endian marker");
  write_scalar<uint64_t>(o, units_.name.size(), "snapshot units length");
  write_raw(o, units_.name.data(), units_.name.size(), "snapshot units");
  write_scalar<uint64_t>(o, step_, "snapshot step");
  write_scalar<uint64_t>(o, n, "snapshot particle count");
  const real_t lo[3] = {box_.lo[0], box_.lo[1], box_.lo[2]};
  const real_t hi[3] = {spec_box_hi_[0], spec_box_hi_[1], spec_box_hi_[2]};
  write_raw(o, lo, 3, "snapshot box lower");
  write_raw(o, hi, 3, "snapshot box upper");
  write_raw(o, tag.data(), n, "snapshot tag");
  write_raw(o, type.data(), n, "snapshot type");
  write_raw(o, molid.data(), n, "snapshot molecule_id");
  write_raw(o, mass.data(), n, "snapshot mass");
  std::vector<real_t> col(n);
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < n; ++i) col[i] = pos[3 * i + a];
    write_raw(o, col.data(), n, "snapshot position");
  }
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < n; ++i) col[i] = vel[3 * i + a];
    write_raw(o, col.data(), n, "snapshot velocity");
  }
  std::vector<int32_t> icol(n);
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < n; ++i) icol[i] = img[3 * i + a];
    write_raw(o, icol.data(), n, "snapshot image");
  }
  o.flush();
  if (!o) {
    fail(BEADS_E_RUNTIME, "failed to write final-state output file " + final_state_path_ + ".");
  }
}

// log_output.cpp:233-258: "[BEADS] LEVEL message", blank line between blocks.
void Engine::log(const std::string& level, const std::string& message) {
  const int echo = spec_.log_echo;
  if (echo == BEADS_LOG_NONE) return;
  const bool screen = echo == BEADS_LOG_SCREEN || echo == BEADS_LOG_BOTH;
  const bool file = echo == BEADS_LOG_FILE || echo == BEADS_LOG_BOTH;
  const bool separate = log_wrote_any_ && level != "THERMO";
  const std::string line = "[BEADS] " + level + " " + message;
  if (screen) {
    if (separate) std::cout << '\n';
    std::cout << line << '\n';
  }
  if (file) {
    if (!log_stream_.is_open()) {
      log_stream_.open(log_path_);
      if (!log_stream_) fail(BEADS_E_RUNTIME, "failed to open log output file " + log_path_ + ".");
    }
    if (separate) log_stream_ << '\n';
    log_stream_ << line << '\n';
  }
  log_wrote_any_ = true;
}

void Engine::poll_interrupt() {
  if (spec_.interrupt != nullptr && spec_.interrupt(spec_.interrupt_context) != 0) {
    (void)hipStreamSynchronize(stream_);
    fail(BEADS_E_INTERRUPTED, "interrupted");
  }
}

// Run-start blocks (log_output.cpp:260-348; values as simulation_run.cpp:12-55
// gathers them).
void Engine::log_run_start() {
  if (spec_.log_echo == BEADS_LOG_NONE) return;
  const real_t lo[3] = {box_.lo[0], box_.lo[1], box_.lo[2]};
  const real_t len[3] = {box_.L[0], box_.L[1], box_.L[2]};
  {
    std::ostringstream m;
    m << "execute start" << '\n' << label_value("units", units_.name) << '\n'
      << label_value("particles", n_) << '\n' << label_value("box_lo", triple(lo)) << '\n'
      << label_value("box_hi", triple(spec_box_hi_)) << '\n'
      << label_value("box_lengths", triple(len)) << '\n'
      << label_value("runsteps", spec_.runsteps);
    log("INFO", m.str());
  }
  {
    std::ostringstream m;
    m << "dynamics summary" << '\n'
      << label_value("style", models_.integrator->style_name()) << '\n'
      << label_value("dt", prec9(models_.integrator->dt())) << '\n'
      << label_value("thermostat",
                     models_.thermostat ? models_.thermostat->style_name() : "none");
    log("INFO", m.str());
  }
  {
    std::ostringstream m;
    m << "neighbor summary" << '\n' << label_value("cutoff_buffer", spec_.cutoff_buffer) << '\n'
      << label_value("path", plan_.path == 0 ? "direct"
                                            : (plan_.path == 1 ? "regular_cell" : "small_dimension"))
      << '\n' << label_value("rebuild_check_every", spec_.rebuild_check_every) << '\n'
      << label_value("sort_every_rebuild", spec_.sort_every_rebuild) << '\n';
    if (spec_.sort_every_steps != 0) {
      m << label_value("sort_every_steps", spec_.sort_every_steps) << '\n';
    }
    m << label_value("max_neighbors", spec_.max_neighbors);
    log("INFO", m.str());
  }
  {
    const char* policy = spec_.exclusion_policy == BEADS_EXCLUSION_DEFAULT
                             ? "default"
                             : (spec_.exclusion_policy == BEADS_EXCLUSION_EXPLICIT ? "explicit"
                                                                                   : "none");
    const std::string distance =
        spec_.exclusion_distance > 0 ? std::to_string(spec_.exclusion_distance) : "none";
    std::ostringstream m;
    auto style = [](const std::unique_ptr<ListedModel>& m) {
      return m ? m->style_name() : "none";
    };
    m << "forcefield summary" << '\n' << label_value("pair", models_.pair->style_name()) << '\n'
      << label_value("bond", style(models_.bond)) << '\n'
      << label_value("angle", style(models_.angle)) << '\n'
      << label_value("dihedral", style(models_.dihedral)) << '\n'
      << label_value("bonded_exclusion_policy", policy) << '\n'
      << label_value("bonded_exclusion_distance", distance) << '\n'
      << label_value("bonded_excluded_pairs", excluded_pairs_);
    log("INFO", m.str());
  }
  {
    std::ostringstream m;
    m << "topology summary" << '\n' << label_value("bonds", spec_.n_bonds) << '\n'
      << label_value("angles", spec_.n_angles) << '\n'
      << label_value("dihedrals", spec_.n_dihedrals);
    log("INFO", m.str());
  }
  {
    const int echo = spec_.log_echo;
    const char* echo_name =
        echo == BEADS_LOG_SCREEN ? "screen" : (echo == BEADS_LOG_FILE ? "log" : "both");
    auto on = [](bool b) { return b ? "enabled" : "disabled"; };
    std::ostringstream m;
    m << "output summary" << '\n' << label_value("thermo", on(!thermo_path_.empty())) << '\n'
      << label_value("final_state", on(!final_state_path_.empty())) << '\n'
      << label_value("trajectory", on(!trajectory_path_.empty())) << '\n'
      << label_value("log", echo_name);
    log("INFO", m.str());
  }
}

// Run-end block (log_output.cpp:359-402).
void Engine::log_run_end(uint64_t runsteps, double wall_s, const Ctrl& c) {
  if (spec_.log_echo == BEADS_LOG_NONE) return;
  std::ostringstream m;
  m << "execute end" << '\n' << label_value("runsteps", runsteps) << '\n'
    << label_value("run_wall_s", prec9(wall_s));
  if (wall_s >= 60.0) {
    const int whole = static_cast<int>(wall_s);
    const int h = whole / 3600, mins = (whole % 3600) / 60;
    std::ostringstream t;
    if (h > 0) t << h << "h ";
    t << mins << "m " << std::fixed << std::setprecision(2) << (wall_s - h * 3600.0 - mins * 60.0)
      << "s";
    m << '\n' << label_value("run_wall_time", t.str());
  }
  m << '\n'
    << label_value("steps_per_second",
                   runsteps ? prec9(static_cast<double>(runsteps) / wall_s) : "not_applicable")
    << '\n' << label_value("neighbor_build_count", c.build_count) << '\n'
    << label_value("dangerous_rebuild_count", c.dangerous);
  log("INFO", m.str());
}

// SimulationRun::execute (simulation_run.cpp:158-244).
void Engine::run() {
  const uint64_t runsteps = spec_.runsteps;
  const bool any_output = !thermo_path_.empty() || !trajectory_path_.empty() ||
                          !final_state_path_.empty() || spec_.log_echo != BEADS_LOG_NONE;
  if (runsteps == 0 && !any_output) {
    return;
  }
  run_start_ = std::chrono::steady_clock::now();
  const char* sync_output = std::getenv("BEADS_AMD_SYNC_OUTPUT");
  async_output_ = !(sync_output != nullptr && sync_output[0] == '1');
  log_run_start();
  const bool thermo0 = !thermo_path_.empty();
  const bool traj_force0 = !trajectory_path_.empty() && (spec_.trajectory_fields & BEADS_TRAJ_FORCE);
  if (runsteps > 0 || thermo0 || traj_force0) {
    initialize(thermo0);
  }
  if (output_due(0)) {
    stage_outputs(0);
  }
  const uint64_t g = graph_steps();
  // A multiple of `every` in (from, to]?
  auto hits = [](uint64_t every, uint64_t from, uint64_t to) {
    return every > 0 && to / every > from / every;
  };
  for (uint64_t s = 0; s < runsteps;) {
    // Stretches of g steps with no thermo row (which needs observables) and
    // no trajectory frame before their last step replay a captured graph;
    // the interrupt callback then runs once per stretch.
    const bool thermo_inside =
        !thermo_path_.empty() && hits(spec_.thermo_every, step_, step_ + g);
    const bool traj_inside =
        !trajectory_path_.empty() && hits(spec_.trajectory_every, step_, step_ + g - 1);
    if (use_graphs_ && step_ % g == 0 && s + g <= runsteps && !thermo_inside && !traj_inside) {
      poll_interrupt();
      const uint64_t before = step_;
      replay_step_graph();
      s += g;
      if (output_due(step_)) {
        stage_outputs(step_);
      } else if (hits(kOverflowPollSteps, before, step_)) {
        check_overflow(fetch_ctrl());
      }
      continue;
    }
    if (s % interrupt_interval_ == 0) {
      poll_interrupt();
    }
    ++s;
    const uint64_t physical = step_ + 1;
    const bool thermo_due = !thermo_path_.empty() && physical % spec_.thermo_every == 0;
    step_once(thermo_due);
    if (output_due(step_)) {
      stage_outputs(step_);
    } else if (step_ % kOverflowPollSteps == 0) {
      check_overflow(fetch_ctrl());
    }
  }
  BEADS_HIP(hipGetLastError());
  if (output_) output_->drain();  // rethrows what the writer hit
  materialize_velocities();
  const Ctrl c = fetch_ctrl();
  if (initialized_) check_overflow(c);
  if (!final_state_path_.empty()) {
    write_final_state();
  }
  if (thermo_stream_.is_open()) thermo_stream_.flush();
  if (trajectory_stream_.is_open()) trajectory_stream_.flush();
  const double wall =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - run_start_).count();
  log_run_end(runsteps, wall, c);
  if (log_stream_.is_open()) log_stream_.flush();
  std::cout.flush();
}

}  // namespace beads_amd
