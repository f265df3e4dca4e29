// FP64 host oracle for the BEADS per-step MD path.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker. The product path never links or
// calls it.
//
// Every function restates the reference's algorithm; the citation next to it
// names the reference file:line (paths relative to the reference's native/src).
// Build with -ffp-contract=off: the binary32 decision expressions below must be
// evaluated exactly in the written order, like the reference's source text.
//
// Parity pins (what checks each function; all in tests/test_oracle_*.py):
//   plan            closed-form cell counts / path choice / error on r_s > L/2
//   cell ids, sort  brute-force recount, permutation + stability invariants
//   exclusions      brute-force all-pairs shortest path on small graphs
//   neighbour build three paths agree with each other and with an O(N^2) scan
//   forces          known answers (LJ minimum, U(sigma)=0, shift), central FD of
//                   the energy, sum of forces = 0, f64-vs-f32-input agreement
//   virials         two-body LJ and one-bond closed forms, W = -dU/dlambda under
//                   uniform scaling of positions and box (central difference),
//                   zero angle/dihedral virial (scale invariance)
//   vv / berendsen  closed forms, lambda = 1 at T = T0, NVE energy drift
//   displacement    binary32 strict threshold at the last representable value,
//   test, cadence   image unwrapping, closed-form build/dangerous counts of a
//                   drifting bead for check intervals 1, 3, 10 and zero skin
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

// angle_harmonic.cu:25: real_t{0.001} of the binary32 and binary64 builds.
constexpr double kSinFloor32 = static_cast<double>(0.001f);
constexpr double kSinFloor64 = 0.001;

// box_geometry.cpp:9-17 in the reference's real_t: binary32 (BoxF), or
// binary64 (BoxD) for the BEADS_REAL64 build.
template <typename Real>
struct BoxT {
  Real lo[3];
  Real L[3];
  Real hL[3];
  Real invL[3];
};
using BoxF = BoxT<float>;
using BoxD = BoxT<double>;

template <typename Real>
BoxT<Real> make_box_t(const float lo[3], const float hi[3]) {
  BoxT<Real> b;
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = lo[a];
    b.L[a] = static_cast<Real>(hi[a]) - static_cast<Real>(lo[a]);
    b.hL[a] = Real(0.5) * b.L[a];
    b.invL[a] = Real(1) / b.L[a];
  }
  return b;
}

BoxF make_box(const float lo[3], const float hi[3]) { return make_box_t<float>(lo, hi); }

// The neighbour plan in binary64 (same fields as orc_plan_t).
using PlanD = orc_plan_f64_t;

// box_geometry_view.cuh:50-62, minimum image for wrapped coordinates.
inline float mic_f(float d, float L, float hL) {
  if (d > hL) {
    d -= L;
  } else if (d < -hL) {
    d += L;
  }
  return d;
}

inline double mic_d(double d, double L, double hL) {
  if (d > hL) {
    d -= L;
  } else if (d < -hL) {
    d += L;
  }
  return d;
}

template <typename Real>
inline Real mic_any(Real d, Real L, Real hL) {
  if (d > hL) {
    d -= L;
  } else if (d < -hL) {
    d += L;
  }
  return d;
}

template <typename Real>
inline Real mic_diff_any(Real xi, Real xj, Real L, Real hL) {
  const Real d = xi - xj;
  if (d > hL) {
    return (xi - L) - xj;
  }
  if (d < -hL) {
    return xi - (xj - L);
  }
  return d;
}

// neighbor_plan.cpp:29-49 (count computed in double), :133-166, :107-115,
// :90-105 and the constructor checks :182-236, in real_t = Real. Plan is
// orc_plan_t (Real = float) or PlanD (Real = double).
template <typename Real, typename Plan>
bool make_plan_t(const float lo[3], const float hi[3], Real rc, Real buffer, uint32_t n,
                 Plan* p) {
  const BoxT<Real> b = make_box_t<Real>(lo, hi);
  if (n == 0) {
    g_error = "neighbor n_particles must be positive.";
    return false;
  }
  if (!std::isfinite(rc) || rc <= Real(0)) {
    g_error = "neighbor force_cutoff must be finite and positive.";
    return false;
  }
  if (!std::isfinite(buffer) || buffer < Real(0)) {
    g_error = "neighbor cutoff_buffer must be finite and non-negative.";
    return false;
  }
  const Real rs = rc + buffer;
  const Real rs2 = rs * rs;
  const Real half_shortest = Real(0.5) * std::min(b.L[0], std::min(b.L[1], b.L[2]));
  if (rs > half_shortest) {
    g_error = "neighbor search_cutoff must not exceed half shortest box length.";
    return false;
  }
  uint32_t counts[3];
  for (int a = 0; a < 3; ++a) {
    const double raw = std::floor(static_cast<double>(b.L[a]) / static_cast<double>(rs));
    counts[a] = raw < 1.0 ? 1u : static_cast<uint32_t>(raw);
    p->cell_size[a] = b.L[a] / static_cast<Real>(counts[a]);
    p->inv_cell[a] = Real(1) / p->cell_size[a];
    if (counts[a] >= 3) {
      const Real tol = std::max(Real(1.0e-6), rs * Real(1.0e-5));
      if (p->cell_size[a] + tol < rs) {
        g_error = "neighbor cell size must be at least neighbor search_cutoff.";
        return false;
      }
    }
  }
  p->nx = counts[0];
  p->ny = counts[1];
  p->nz = counts[2];
  p->n_cells = counts[0] * counts[1] * counts[2];
  p->force_cutoff = rc;
  p->search_cutoff = rs;
  p->search_cutoff_sq = rs2;
  if (counts[0] < 3 && counts[1] < 3 && counts[2] < 3) {
    p->path = ORC_PATH_DIRECT;
  } else if (counts[0] >= 3 && counts[1] >= 3 && counts[2] >= 3) {
    p->path = ORC_PATH_REGULAR;
  } else {
    p->path = ORC_PATH_SMALL;
  }
  return true;
}

bool make_plan(const float lo[3], const float hi[3], float rc, float buffer,
               uint32_t n, orc_plan_t* p) {
  return make_plan_t<float>(lo, hi, rc, buffer, n, p);
}

// cell_list_build.cu:110-123 in real_t = Real.
template <typename Real>
inline uint32_t axis_cell(Real x, Real lo, Real inv, uint32_t count) {
  const Real s = (x - lo) * inv;
  uint32_t i = s <= Real(0) ? 0u : static_cast<uint32_t>(s);
  if (i >= count) {
    i = count - 1;
  }
  return i;
}

// regular_cell_neighbor_builder.cu:40-53 (wrapped neighbour cell).
inline uint32_t wrapped_axis(uint32_t c, int delta, uint32_t count) {
  if (delta < 0) {
    return c == 0 ? count - 1 : c - 1;
  }
  if (delta > 0) {
    const uint32_t next = c + 1;
    return next == count ? 0 : next;
  }
  return c;
}

// small_dimension_neighbor_builder.cu:56-76 (per-axis unique cells).
inline int unique_axis(uint32_t c, uint32_t count, uint32_t out[3]) {
  if (count <= 1) {
    out[0] = 0;
    return 1;
  }
  if (count == 2) {
    out[0] = c == 0 ? 1 : 0;
    out[1] = c;
    return 2;
  }
  out[0] = wrapped_axis(c, -1, count);
  out[1] = c;
  out[2] = wrapped_axis(c, 1, count);
  return 3;
}

// Counting sort by cell; inside a cell by (x, slot) when x is given, else by
// slot. The reference's scatter (cell_list_build.cu:168-181) is atomic and
// leaves a cell's members in arbitrary order; the engine under test orders
// them by (x, slot), which this restates. Any order inside a cell gives the
// same neighbour sets; it fixes the candidate order of a build.
template <typename Key = float>
void stable_order(uint32_t n, const uint32_t* cell, uint32_t n_cells,
                  std::vector<uint32_t>& order, std::vector<uint32_t>& counts,
                  std::vector<uint32_t>& offsets, const Key* x = nullptr,
                  size_t x_stride = 1) {
  counts.assign(n_cells, 0);
  for (uint32_t i = 0; i < n; ++i) {
    ++counts[cell[i]];
  }
  offsets.assign(n_cells, 0);
  uint32_t run = 0;
  for (uint32_t c = 0; c < n_cells; ++c) {
    offsets[c] = run;
    run += counts[c];
  }
  std::vector<uint32_t> cursor = offsets;
  order.assign(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    order[cursor[cell[i]]++] = i;
  }
  if (x != nullptr) {
    for (uint32_t c = 0; c < n_cells; ++c) {
      std::stable_sort(order.begin() + offsets[c], order.begin() + offsets[c] + counts[c],
                       [&](uint32_t a, uint32_t b) { return x[a * x_stride] < x[b * x_stride]; });
    }
  }
}

// exclusions.cu:514-598: breadth-first search on the bond graph from every tag;
// every partner first reached at depth <= distance forms an excluded pair.
std::vector<std::pair<uint32_t, uint32_t>> exclusion_pairs(
    uint32_t n, uint32_t n_bonds, const uint32_t* bonds, int distance) {
  std::vector<std::vector<uint32_t>> adj(static_cast<size_t>(n) + 1);
  for (uint32_t b = 0; b < n_bonds; ++b) {
    const uint32_t a = bonds[3 * b];
    const uint32_t c = bonds[3 * b + 1];
    adj[a].push_back(c);
    adj[c].push_back(a);
  }
  for (auto& v : adj) {
    std::sort(v.begin(), v.end());
  }
  std::vector<std::pair<uint32_t, uint32_t>> out;
  if (distance <= 0) {
    return out;
  }
  std::vector<uint32_t> seen(static_cast<size_t>(n) + 1, 0);
  std::vector<std::pair<uint32_t, int>> queue;
  uint32_t gen = 0;
  for (uint32_t start = 1; start <= n; ++start) {
    ++gen;
    queue.clear();
    queue.emplace_back(start, 0);
    seen[start] = gen;
    for (size_t r = 0; r < queue.size(); ++r) {
      const uint32_t t = queue[r].first;
      const int depth = queue[r].second;
      if (depth == distance) {
        continue;
      }
      for (uint32_t p : adj[t]) {
        if (seen[p] == gen) {
          continue;
        }
        seen[p] = gen;
        out.emplace_back(std::min(start, p), std::max(start, p));
        queue.emplace_back(p, depth + 1);
      }
    }
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

// Slot-level exclusion lookup (exclusions.cuh:151-192 restated with a
// per-slot sorted partner vector).
struct SlotExclusions {
  std::vector<std::vector<uint32_t>> partners;  // by slot, sorted slots
  bool any = false;
  bool excluded(uint32_t i, uint32_t j) const {
    if (!any) {
      return false;
    }
    const auto& v = partners[i];
    return std::binary_search(v.begin(), v.end(), j);
  }
};

SlotExclusions slot_exclusions(const orc_system_t* s, const uint32_t* tags) {
  SlotExclusions e;
  if (s->exclusion_distance <= 0 || s->n_bonds == 0) {
    return e;
  }
  std::vector<uint32_t> slot_of_tag(static_cast<size_t>(s->n) + 1, 0);
  for (uint32_t i = 0; i < s->n; ++i) {
    slot_of_tag[tags[i]] = i;
  }
  const auto pairs = exclusion_pairs(s->n, s->n_bonds, s->bonds, s->exclusion_distance);
  e.partners.assign(s->n, {});
  for (const auto& p : pairs) {
    const uint32_t a = slot_of_tag[p.first];
    const uint32_t b = slot_of_tag[p.second];
    e.partners[a].push_back(b);
    e.partners[b].push_back(a);
  }
  for (auto& v : e.partners) {
    std::sort(v.begin(), v.end());
  }
  e.any = true;
  return e;
}

// Candidate enumeration of the three full-list builders
// (regular_cell_neighbor_builder.cu:80-107, small_dimension_neighbor_builder.cu:
// 100-131, direct_neighbor_builder.cu:33-47) with the append rule of
// full_neighbor_path_common.cuh:49-94. Pos is any accessor giving x,y,z of a
// slot; Real is the decision arithmetic; cells and their (x, slot) order are
// evaluated in the plan's real type (binary32 for orc_plan_t, binary64 for
// PlanD) on the box of the same type.
template <typename Real, typename Pos, typename Box, typename Plan>
void build_lists(uint32_t n, const Pos& pos, const Box& box, const Plan& plan,
                 int path, const SlotExclusions& excl, Real rs2,
                 std::vector<std::vector<uint32_t>>& lists) {
  using CR = std::remove_cv_t<std::remove_reference_t<decltype(plan.inv_cell[0])>>;
  lists.assign(n, {});
  std::vector<uint32_t> cell(n), order, counts, offsets;
  if (path != ORC_PATH_DIRECT) {
    for (uint32_t i = 0; i < n; ++i) {
      const uint32_t ix = axis_cell<CR>(static_cast<CR>(pos.x(i)), box.lo[0], plan.inv_cell[0], plan.nx);
      const uint32_t iy = axis_cell<CR>(static_cast<CR>(pos.y(i)), box.lo[1], plan.inv_cell[1], plan.ny);
      const uint32_t iz = axis_cell<CR>(static_cast<CR>(pos.z(i)), box.lo[2], plan.inv_cell[2], plan.nz);
      cell[i] = ix + plan.nx * (iy + plan.ny * iz);
    }
    std::vector<CR> xs(n);
    for (uint32_t i = 0; i < n; ++i) xs[i] = static_cast<CR>(pos.x(i));
    stable_order<CR>(n, cell.data(), plan.n_cells, order, counts, offsets, xs.data());
  }
  const Real L[3] = {static_cast<Real>(box.L[0]), static_cast<Real>(box.L[1]), static_cast<Real>(box.L[2])};
  const Real hL[3] = {static_cast<Real>(box.hL[0]), static_cast<Real>(box.hL[1]), static_cast<Real>(box.hL[2])};

#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
    const uint32_t i = static_cast<uint32_t>(ii);
    const Real xi = pos.x(i), yi = pos.y(i), zi = pos.z(i);
    auto& out = lists[i];
    auto visit = [&](uint32_t j) {
      if (i == j || excl.excluded(i, j)) {
        return;
      }
      const Real dx = mic_any<Real>(xi - pos.x(j), L[0], hL[0]);
      const Real dy = mic_any<Real>(yi - pos.y(j), L[1], hL[1]);
      const Real dz = mic_any<Real>(zi - pos.z(j), L[2], hL[2]);
      const Real r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > rs2) {
        return;
      }
      out.push_back(j);
    };
    if (path == ORC_PATH_DIRECT) {
      for (uint32_t j = 0; j < n; ++j) {
        visit(j);
      }
      continue;
    }
    const uint32_t c = cell[i];
    const uint32_t cx = c % plan.nx;
    const uint32_t cy = (c / plan.nx) % plan.ny;
    const uint32_t cz = (c / plan.nx) / plan.ny;
    uint32_t xs[3], ys[3], zs[3];
    int nxs, nys, nzs;
    if (path == ORC_PATH_REGULAR) {
      for (int d = -1; d <= 1; ++d) {
        xs[d + 1] = wrapped_axis(cx, d, plan.nx);
        ys[d + 1] = wrapped_axis(cy, d, plan.ny);
        zs[d + 1] = wrapped_axis(cz, d, plan.nz);
      }
      nxs = nys = nzs = 3;
    } else {
      nxs = unique_axis(cx, plan.nx, xs);
      nys = unique_axis(cy, plan.ny, ys);
      nzs = unique_axis(cz, plan.nz, zs);
    }
    for (int a = 0; a < nzs; ++a) {
      for (int b = 0; b < nys; ++b) {
        for (int d = 0; d < nxs; ++d) {
          const uint32_t nc = xs[d] + plan.nx * (ys[b] + plan.ny * zs[a]);
          for (uint32_t k = 0; k < counts[nc]; ++k) {
            visit(order[offsets[nc] + k]);
          }
        }
      }
    }
  }
}


// Position accessors: binary32 AoS input, or binary64 AoS.
struct PosF {
  const float* p;
  float x(uint32_t i) const { return p[3 * i]; }
  float y(uint32_t i) const { return p[3 * i + 1]; }
  float z(uint32_t i) const { return p[3 * i + 2]; }
};
struct PosD {
  const double* p;
  double x(uint32_t i) const { return p[3 * i]; }
  double y(uint32_t i) const { return p[3 * i + 1]; }
  double z(uint32_t i) const { return p[3 * i + 2]; }
};

// pair_lj.cu:171-192 (packing) evaluated in binary64 from the binary32
// parameters the engine receives.
struct LjCoeff {
  double sigma2, eps4, eps24, shift;
};

std::vector<LjCoeff> lj_table(const orc_system_t* s) {
  const int T = s->n_types;
  std::vector<LjCoeff> t(static_cast<size_t>(T) * T);
  const double rc = s->lj_cutoff;
  for (int a = 0; a < T; ++a) {
    for (int b = 0; b < T; ++b) {
      const double eps = s->lj_eps[a * T + b];
      const double sig = s->lj_sig[a * T + b];
      LjCoeff c;
      c.sigma2 = sig * sig;
      c.eps4 = 4.0 * eps;
      c.eps24 = 24.0 * eps;
      c.shift = 0.0;
      if (s->lj_shift) {
        const double sr2 = c.sigma2 / (rc * rc);
        const double sr6 = sr2 * sr2 * sr2;
        c.shift = c.eps4 * (sr6 * sr6 - sr6);
      }
      t[static_cast<size_t>(a) * T + b] = c;
    }
  }
  return t;
}

struct Vec {
  double x, y, z;
};
inline double vdot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename Pos>
Vec mic_vec(const Pos& pos, uint32_t i, uint32_t j, const double L[3], const double hL[3]) {
  return {mic_d(static_cast<double>(pos.x(i)) - static_cast<double>(pos.x(j)), L[0], hL[0]),
          mic_d(static_cast<double>(pos.y(i)) - static_cast<double>(pos.y(j)), L[1], hL[1]),
          mic_d(static_cast<double>(pos.z(i)) - static_cast<double>(pos.z(j)), L[2], hL[2])};
}

// Extended precision (x87, 64-bit significand) for the listed terms: the
// reference's acos-based angle and dihedral are ill-conditioned near 0 and pi,
// where binary64 would leave the checker's own rounding above the error of
// the engine's stable formulation.
using X = long double;
constexpr X kPiX = 3.141592653589793238462643383279502884L;
struct XVec {
  X x, y, z;
};
inline XVec xsub(XVec a, XVec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline XVec xadd(XVec a, XVec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline XVec xscale(XVec a, X s) { return {a.x * s, a.y * s, a.z * s}; }
inline X xdot(XVec a, XVec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline XVec xcross(XVec a, XVec b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline X mic_x(X d, X L, X hL) { return d > hL ? d - L : (d < -hL ? d + L : d); }

template <typename Pos>
XVec mic_xvec(const Pos& pos, uint32_t i, uint32_t j, const X L[3], const X hL[3]) {
  return {mic_x(X(pos.x(i)) - X(pos.x(j)), L[0], hL[0]),
          mic_x(X(pos.y(i)) - X(pos.y(j)), L[1], hL[1]),
          mic_x(X(pos.z(i)) - X(pos.z(j)), L[2], hL[2])};
}

// Forces + observables. lists[i] holds the stored neighbours of slot i;
// the pair is evaluated if the decision arithmetic (Real) gives r2 < rc2
// (pair_lj.cu:54), and its value in binary64 (pair_lj.cu:56-66, :128-150).
// Listed terms follow bond_harmonic.cu:41-79, angle_harmonic.cu:61-158 and
// dihedral_harmonic.cu:83-206 in extended precision, rounded to binary64 when
// added; sin_floor is the angle's real_t{0.001} (angle_harmonic.cu:25) of the
// build under test.
template <typename Real, typename Pos, typename Box>
void compute_forces(const orc_system_t* s, const Pos& pos, const Box& box,
                    const std::vector<std::vector<uint32_t>>& lists, double* f,
                    orc_observables_t* obs, double sin_floor) {
  const uint32_t n = s->n;
  const double Ld[3] = {box.L[0], box.L[1], box.L[2]};
  const double hLd[3] = {box.hL[0], box.hL[1], box.hL[2]};
  const Real Lr[3] = {static_cast<Real>(box.L[0]), static_cast<Real>(box.L[1]), static_cast<Real>(box.L[2])};
  const Real hLr[3] = {static_cast<Real>(box.hL[0]), static_cast<Real>(box.hL[1]), static_cast<Real>(box.hL[2])};
  const Real rc = static_cast<Real>(s->lj_cutoff);
  const Real rc2 = rc * rc;
  const auto table = lj_table(s);
  const int T = s->n_types;
  double pair_pe = 0.0, pair_vir = 0.0;

#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pair_pe, pair_vir)
  for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
    const uint32_t i = static_cast<uint32_t>(ii);
    double fx = 0, fy = 0, fz = 0, pe = 0, vir = 0;
    const int ti = s->type[i] - 1;
    for (uint32_t j : lists[i]) {
      // The r^2 < r_c^2 decision (pair_lj.cu:54) on the minimum-image
      // displacement with exact intermediates: (xi - L) - xj or xi - (xj - L)
      // when the raw difference crosses half the box. Same value as the
      // reference's MIC(xi - xj) in exact arithmetic, rounded once instead of
      // at the scale of L; the device's force kernel evaluates it identically.
      const Real dxr = mic_diff_any<Real>(pos.x(i), pos.x(j), Lr[0], hLr[0]);
      const Real dyr = mic_diff_any<Real>(pos.y(i), pos.y(j), Lr[1], hLr[1]);
      const Real dzr = mic_diff_any<Real>(pos.z(i), pos.z(j), Lr[2], hLr[2]);
      const Real r2r = dxr * dxr + dyr * dyr + dzr * dzr;
      if (r2r >= rc2) {
        continue;
      }
      const LjCoeff& c = table[static_cast<size_t>(ti) * T + (s->type[j] - 1)];
      const Vec d = mic_vec(pos, i, j, Ld, hLd);
      const double r2 = vdot(d, d);
      const double sr2 = c.sigma2 / r2;
      const double sr6 = sr2 * sr2 * sr2;
      const double sr12 = sr6 * sr6;
      const double sc = c.eps24 / r2 * (2.0 * sr12 - sr6);
      fx += sc * d.x;
      fy += sc * d.y;
      fz += sc * d.z;
      pe += c.eps4 * (sr12 - sr6) - c.shift;
      vir += sc * r2;
    }
    f[3 * i] = fx;
    f[3 * i + 1] = fy;
    f[3 * i + 2] = fz;
    pair_pe += 0.5 * pe;
    pair_vir += 0.5 * vir;
  }

  std::vector<uint32_t> slot_of_tag(static_cast<size_t>(n) + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    slot_of_tag[s->tag[i]] = i;
  }
  auto add = [&](uint32_t slot, XVec v) {
    f[3 * slot] += static_cast<double>(v.x);
    f[3 * slot + 1] += static_cast<double>(v.y);
    f[3 * slot + 2] += static_cast<double>(v.z);
  };
  const X Lx[3] = {Ld[0], Ld[1], Ld[2]};
  const X hLx[3] = {hLd[0], hLd[1], hLd[2]};

  X bond_pe = 0, bond_vir = 0;
  for (uint32_t b = 0; b < s->n_bonds; ++b) {
    const uint32_t si = slot_of_tag[s->bonds[3 * b]];
    const uint32_t sj = slot_of_tag[s->bonds[3 * b + 1]];
    const int ty = static_cast<int>(s->bonds[3 * b + 2]) - 1;
    const X k = s->bond_k[ty], r0 = s->bond_r0[ty];
    const XVec d = mic_xvec(pos, si, sj, Lx, hLx);
    const X r = std::sqrt(xdot(d, d));
    const X stretch = r - r0;
    bond_pe += X(0.5) * k * stretch * stretch;
    if (r == 0) {
      continue;
    }
    const XVec fi = xscale(d, -k * stretch / r);
    add(si, fi);
    add(sj, xscale(fi, -1));
    bond_vir += xdot(d, fi);
  }

  const X sf = sin_floor;
  X angle_pe = 0, angle_vir = 0;
  for (uint32_t a = 0; a < s->n_angles; ++a) {
    const uint32_t si = slot_of_tag[s->angles[4 * a]];
    const uint32_t sj = slot_of_tag[s->angles[4 * a + 1]];
    const uint32_t sk = slot_of_tag[s->angles[4 * a + 2]];
    const int ty = static_cast<int>(s->angles[4 * a + 3]) - 1;
    const X k = s->angle_k[ty];
    const X theta0 = static_cast<X>(s->angle_theta0_deg[ty]) * kPiX / 180;
    const XVec va = mic_xvec(pos, si, sj, Lx, hLx);
    const XVec vb = mic_xvec(pos, sk, sj, Lx, hLx);
    const X ra2 = xdot(va, va), rb2 = xdot(vb, vb);
    if (ra2 == 0 || rb2 == 0) {
      angle_pe += X(0.5) * k * theta0 * theta0;
      continue;
    }
    const X ra = std::sqrt(ra2), rb = std::sqrt(rb2);
    X c = xdot(va, vb) / (ra * rb);
    c = std::max(X(-1), std::min(X(1), c));
    const X theta = c == 1 ? X(0) : (c == -1 ? kPiX : std::acos(c));
    const X dth = theta - theta0;
    angle_pe += X(0.5) * k * dth * dth;
    const X sin_sq = 1 - c * c;
    const X sn = sin_sq <= sf * sf ? sf : std::sqrt(sin_sq);
    const X A = -k * dth / sn;
    const X a11 = A * c / ra2, a12 = -A / (ra * rb), a22 = A * c / rb2;
    const XVec fi = xadd(xscale(va, a11), xscale(vb, a12));
    const XVec fk = xadd(xscale(vb, a22), xscale(va, a12));
    add(si, fi);
    add(sk, fk);
    add(sj, xscale(xadd(fi, fk), -1));
    angle_vir += xdot(va, fi) + xdot(vb, fk);
  }

  X dih_pe = 0, dih_vir = 0;
  for (uint32_t q = 0; q < s->n_dihedrals; ++q) {
    const uint32_t si = slot_of_tag[s->dihedrals[5 * q]];
    const uint32_t sj = slot_of_tag[s->dihedrals[5 * q + 1]];
    const uint32_t sk = slot_of_tag[s->dihedrals[5 * q + 2]];
    const uint32_t sl = slot_of_tag[s->dihedrals[5 * q + 3]];
    const int ty = static_cast<int>(s->dihedrals[5 * q + 4]) - 1;
    const X k = s->dihedral_k[ty];
    const int dd = s->dihedral_d[ty], nn = s->dihedral_n[ty];
    const XVec v0 = mic_xvec(pos, si, sj, Lx, hLx);
    const XVec v1 = mic_xvec(pos, sk, sj, Lx, hLx);
    const XVec v2 = mic_xvec(pos, sk, sl, Lx, hLx);
    const XVec c0 = xcross(v0, v1), c1 = xcross(v1, v2);
    const X n1 = xdot(v1, v1), nc0 = xdot(c0, c0), nc1 = xdot(c1, c1);
    if (n1 == 0 || nc0 == 0 || nc1 == 0) {
      dih_pe += nn == 0 ? k * (1 + dd) : k;
      continue;
    }
    X c = xdot(c0, c1) / std::sqrt(nc0 * nc1);
    c = std::max(X(-1), std::min(X(1), c));
    X phi = c == 1 ? X(0) : (c == -1 ? kPiX : std::acos(c));
    if (xdot(v0, c1) < 0) {
      phi = -phi;
    }
    const X arg = nn * phi;
    dih_pe += k * (1 + dd * std::cos(arg));
    const X dE = -k * dd * nn * std::sin(arg);
    const X nb = std::sqrt(n1);
    const XVec fi = xscale(c0, -dE * nb / nc0);
    const XVec fl = xscale(c1, dE * nb / nc1);
    const XVec sv = xsub(xscale(fi, xdot(v0, v1) / n1), xscale(fl, xdot(v2, v1) / n1));
    const XVec fj = xsub(sv, fi);
    const XVec fk = xscale(xadd(sv, fl), -1);
    add(si, fi);
    add(sj, fj);
    add(sk, fk);
    add(sl, fl);
    dih_vir += xdot(v0, fi) + xdot(v1, fk) + xdot(xsub(v1, v2), fl);
  }

  if (obs != nullptr) {
    obs->pair_pe = pair_pe;
    obs->pair_virial = pair_vir;
    obs->bond_pe = static_cast<double>(bond_pe);
    obs->bond_virial = static_cast<double>(bond_vir);
    obs->angle_pe = static_cast<double>(angle_pe);
    obs->angle_virial = static_cast<double>(angle_vir);
    obs->dihedral_pe = static_cast<double>(dih_pe);
    obs->dihedral_virial = static_cast<double>(dih_vir);
  }
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

int orc_plan(const float lo[3], const float hi[3], float force_cutoff,
             float cutoff_buffer, uint32_t n, orc_plan_t* out) {
  return make_plan(lo, hi, force_cutoff, cutoff_buffer, n, out) ? 0 : 1;
}

void orc_cell_ids(uint32_t n, const float* pos, const float lo[3],
                  const orc_plan_t* plan, uint32_t* cell) {
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t ix = axis_cell(pos[3 * i], lo[0], plan->inv_cell[0], plan->nx);
    const uint32_t iy = axis_cell(pos[3 * i + 1], lo[1], plan->inv_cell[1], plan->ny);
    const uint32_t iz = axis_cell(pos[3 * i + 2], lo[2], plan->inv_cell[2], plan->nz);
    cell[i] = ix + plan->nx * (iy + plan->ny * iz);
  }
}

void orc_stable_cell_order(uint32_t n, const uint32_t* cell, uint32_t n_cells, const float* pos,
                           uint32_t* order, uint32_t* counts, uint32_t* offsets) {
  std::vector<uint32_t> o, c, off;
  stable_order(n, cell, n_cells, o, c, off, pos, 3);
  std::copy(o.begin(), o.end(), order);
  std::copy(c.begin(), c.end(), counts);
  std::copy(off.begin(), off.end(), offsets);
}

uint64_t orc_exclusion_pairs(uint32_t n, uint32_t n_bonds, const uint32_t* bonds,
                             int32_t distance, uint32_t* pairs, uint64_t cap) {
  const auto p = exclusion_pairs(n, n_bonds, bonds, distance);
  for (uint64_t k = 0; k < p.size() && k < cap; ++k) {
    pairs[2 * k] = p[k].first;
    pairs[2 * k + 1] = p[k].second;
  }
  return p.size();
}

int orc_neighbor_build(const orc_system_t* sys, float cutoff_buffer,
                       uint32_t max_neighbors, int32_t path, uint32_t* counts,
                       uint32_t* list, int32_t* overflow) {
  orc_plan_t plan;
  if (!make_plan(sys->lo, sys->hi, sys->lj_cutoff, cutoff_buffer, sys->n, &plan)) {
    return 1;
  }
  if (path < 0) {
    path = plan.path;
  }
  if (path == ORC_PATH_REGULAR && (plan.nx < 3 || plan.ny < 3 || plan.nz < 3)) {
    g_error = "regular path requires at least 3 cells per axis.";
    return 1;
  }
  const BoxF box = make_box(sys->lo, sys->hi);
  const SlotExclusions excl = slot_exclusions(sys, sys->tag);
  std::vector<std::vector<uint32_t>> lists;
  build_lists<float>(sys->n, PosF{sys->pos}, box, plan, path, excl, plan.search_cutoff_sq, lists);
  *overflow = 0;
  for (uint32_t i = 0; i < sys->n; ++i) {
    const auto& v = lists[i];
    const uint32_t m = static_cast<uint32_t>(std::min<size_t>(v.size(), max_neighbors));
    if (v.size() > max_neighbors) {
      *overflow = 1;
    }
    counts[i] = m;
    std::copy(v.begin(), v.begin() + m, list + static_cast<size_t>(i) * max_neighbors);
  }
  return 0;
}

int orc_forces(const orc_system_t* sys, const uint32_t* counts, const uint32_t* list,
               uint32_t max_neighbors, double* forces, orc_observables_t* obs) {
  std::vector<std::vector<uint32_t>> lists;
  if (counts != nullptr && list != nullptr) {
    lists.assign(sys->n, {});
    for (uint32_t i = 0; i < sys->n; ++i) {
      const uint32_t* row = list + static_cast<size_t>(i) * max_neighbors;
      lists[i].assign(row, row + counts[i]);
    }
  } else {
    orc_plan_t plan;
    if (!make_plan(sys->lo, sys->hi, sys->lj_cutoff, 0.0f, sys->n, &plan)) {
      return 1;
    }
    const BoxF box = make_box(sys->lo, sys->hi);
    const SlotExclusions excl = slot_exclusions(sys, sys->tag);
    build_lists<float>(sys->n, PosF{sys->pos}, box, plan, plan.path, excl,
                       plan.search_cutoff_sq, lists);
  }
  compute_forces<float>(sys, PosF{sys->pos}, make_box(sys->lo, sys->hi), lists, forces, obs,
                        kSinFloor32);
  return 0;
}

int orc_forces_f64(const orc_system_t* sys, const double* pos64, double* forces,
                   orc_observables_t* obs) {
  orc_plan_t plan;
  if (!make_plan(sys->lo, sys->hi, sys->lj_cutoff, 0.0f, sys->n, &plan)) {
    return 1;
  }
  const BoxF box = make_box(sys->lo, sys->hi);
  const SlotExclusions excl = slot_exclusions(sys, sys->tag);
  std::vector<std::vector<uint32_t>> lists;
  const double rc = sys->lj_cutoff;
  build_lists<double>(sys->n, PosD{pos64}, box, plan, plan.path, excl, rc * rc, lists);
  compute_forces<double>(sys, PosD{pos64}, box, lists, forces, obs, kSinFloor32);
  return 0;
}

int orc_plan_f64(const float lo[3], const float hi[3], double force_cutoff,
                 double cutoff_buffer, uint32_t n, orc_plan_f64_t* out) {
  return make_plan_t<double>(lo, hi, force_cutoff, cutoff_buffer, n, out) ? 0 : 1;
}

void orc_cell_ids_f64(uint32_t n, const double* pos, const float lo[3],
                      const orc_plan_f64_t* plan, uint32_t* cell) {
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t ix = axis_cell<double>(pos[3 * i], lo[0], plan->inv_cell[0], plan->nx);
    const uint32_t iy = axis_cell<double>(pos[3 * i + 1], lo[1], plan->inv_cell[1], plan->ny);
    const uint32_t iz = axis_cell<double>(pos[3 * i + 2], lo[2], plan->inv_cell[2], plan->nz);
    cell[i] = ix + plan->nx * (iy + plan->ny * iz);
  }
}

int orc_neighbor_build_f64(const orc_system_t* sys, const double* pos64, double cutoff_buffer,
                           uint32_t max_neighbors, int32_t path, uint32_t* counts,
                           uint32_t* list, int32_t* overflow) {
  PlanD plan;
  if (!make_plan_t<double>(sys->lo, sys->hi, static_cast<double>(sys->lj_cutoff), cutoff_buffer,
                           sys->n, &plan)) {
    return 1;
  }
  if (path < 0) {
    path = plan.path;
  }
  if (path == ORC_PATH_REGULAR && (plan.nx < 3 || plan.ny < 3 || plan.nz < 3)) {
    g_error = "regular path requires at least 3 cells per axis.";
    return 1;
  }
  const BoxD box = make_box_t<double>(sys->lo, sys->hi);
  const SlotExclusions excl = slot_exclusions(sys, sys->tag);
  std::vector<std::vector<uint32_t>> lists;
  build_lists<double>(sys->n, PosD{pos64}, box, plan, path, excl, plan.search_cutoff_sq, lists);
  *overflow = 0;
  for (uint32_t i = 0; i < sys->n; ++i) {
    const auto& v = lists[i];
    const uint32_t m = static_cast<uint32_t>(std::min<size_t>(v.size(), max_neighbors));
    if (v.size() > max_neighbors) {
      *overflow = 1;
    }
    counts[i] = m;
    std::copy(v.begin(), v.begin() + m, list + static_cast<size_t>(i) * max_neighbors);
  }
  return 0;
}

int orc_forces_f64_list(const orc_system_t* sys, const double* pos64, const uint32_t* counts,
                        const uint32_t* list, uint32_t max_neighbors, double* forces,
                        orc_observables_t* obs) {
  std::vector<std::vector<uint32_t>> lists(sys->n);
  for (uint32_t i = 0; i < sys->n; ++i) {
    const uint32_t* row = list + static_cast<size_t>(i) * max_neighbors;
    lists[i].assign(row, row + counts[i]);
  }
  compute_forces<double>(sys, PosD{pos64}, make_box_t<double>(sys->lo, sys->hi), lists, forces,
                         obs, kSinFloor64);
  return 0;
}

// velocity_verlet.cu:59-95; wrap box_geometry_view.cuh:92-102.
void orc_vv_pre(uint32_t n, const float* pos, const float* vel, const float* force,
                const float* mass, const int32_t* image, const float lo[3],
                const float hi[3], double dt, double* pos_out, double* vel_out,
                int32_t* image_out) {
  const BoxF box = make_box(lo, hi);
  for (uint32_t i = 0; i < n; ++i) {
    const double h = 0.5 * dt / static_cast<double>(mass[i]);
    for (int a = 0; a < 3; ++a) {
      const double v = static_cast<double>(vel[3 * i + a]) + h * force[3 * i + a];
      double x = static_cast<double>(pos[3 * i + a]) + dt * v;
      const double c = std::floor((x - box.lo[a]) / static_cast<double>(box.L[a]));
      x -= c * static_cast<double>(box.L[a]);
      vel_out[3 * i + a] = v;
      pos_out[3 * i + a] = x;
      image_out[3 * i + a] = image[3 * i + a] + static_cast<int32_t>(c);
    }
  }
}

// velocity_verlet.cu:97-111.
void orc_vv_post(uint32_t n, const float* vel, const float* force, const float* mass,
                 double dt, double* vel_out) {
  for (uint32_t i = 0; i < n; ++i) {
    const double h = 0.5 * dt / static_cast<double>(mass[i]);
    for (int a = 0; a < 3; ++a) {
      vel_out[3 * i + a] = static_cast<double>(vel[3 * i + a]) + h * force[3 * i + a];
    }
  }
}

// berendsen_thermostat.cu:109-133 (kinetic energy).
double orc_kinetic_energy(uint32_t n, const float* vel, const float* mass) {
  double ke = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
    ke += 0.5 * static_cast<double>(mass[i]) * (vx * vx + vy * vy + vz * vz);
  }
  return ke;
}

// berendsen_thermostat.cu:104-107 (dof) and :135-163 (lambda).
double orc_berendsen_lambda(double ke, uint32_t n, double kb, double temperature,
                            double dt, double tau) {
  const double dof = std::max(0.0, 3.0 * n - 3.0);
  if (!(ke > 0.0) || !(dof > 0.0)) {
    return 1.0;
  }
  const double current = 2.0 * ke / dof;
  const double target = kb * temperature;
  const double arg = 1.0 + (dt / tau) * (target / current - 1.0);
  return std::sqrt(std::max(0.0, arg));
}

// rebuild_tracker.cu:80-92 with the threshold of neighbor_system.cpp:213-216,
// in binary32 like the reference's real_t: the unwrapped displacement
// (x - x_ref) + (float)(image - image_ref) * L per axis, L = hi - lo
// (box_geometry.cpp:9-17), and the strict test dx*dx + dy*dy + dz*dz >
// (0.5 * skin)^2, evaluated in source order without contraction.
int orc_displacement_exceeds(uint32_t n, const float* pos, const int32_t* image,
                             const float* ref_pos, const int32_t* ref_image,
                             const float lo[3], const float hi[3], float cutoff_buffer) {
  const BoxF box = make_box(lo, hi);
  const float half_buffer = 0.5f * cutoff_buffer;
  const float threshold_sq = half_buffer * half_buffer;
  for (uint32_t i = 0; i < n; ++i) {
    float d[3];
    for (int a = 0; a < 3; ++a) {
      const long long di = static_cast<long long>(image[3 * i + a]) -
                           static_cast<long long>(ref_image[3 * i + a]);
      d[a] = (pos[3 * i + a] - ref_pos[3 * i + a]) + static_cast<float>(di) * box.L[a];
    }
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > threshold_sq) {
      return 1;
    }
  }
  return 0;
}

}  // extern "C"

// A whole run in binary64 (simulation_run.cpp:189-235 step order; neighbour
// policy neighbor_system.cpp:79-146, 198-216; displacement test
// rebuild_tracker.cu:55-94; reorder cadence counted in rebuilds, or in steps
// when sort_every_steps is set, the engine's step-based option). Used as the
// CPU baseline and for trajectory-level invariants. The cell grid and box are
// the binary32 build's (D = false) or the BEADS_REAL64 build's (D = true);
// the state, the neighbour test and the displacement test are binary64 in
// both. x0, v0: initial positions and velocities [n][3].
namespace {

// Tangent of the forces, df = J dx with J = dF/dx (the Hessian of -U), for LJ
// pairs over the lists (same r^2 < r_c^2 decision as compute_forces<double>)
// and harmonic bonds: the linearisation of pair_lj.cu:56-66 and
// bond_harmonic.cu:41-79. For a pair, d = x_i - x_j, u = r^2,
// F_i = s(u) d, s = 24 eps / u (2 sr12 - sr6), ds/du = -24 eps / u^2 (14 sr12 -
// 4 sr6): dF_i = s e + 2 ds/du (d . e) d with e = dx_i - dx_j. For a bond,
// F_i = phi(r) d, phi = -k (r - r0) / r: dF_i = phi e - k r0 / r^3 (d . e) d.
template <typename Pos, typename Box>
void force_tangent(const orc_system_t* s, const Pos& pos, const Box& box,
                   const std::vector<std::vector<uint32_t>>& lists, const double* dx,
                   double* df) {
  const uint32_t n = s->n;
  const double Ld[3] = {static_cast<double>(box.L[0]), static_cast<double>(box.L[1]),
                        static_cast<double>(box.L[2])};
  const double hLd[3] = {static_cast<double>(box.hL[0]), static_cast<double>(box.hL[1]),
                         static_cast<double>(box.hL[2])};
  const double rc = static_cast<double>(s->lj_cutoff);
  const double rc2 = rc * rc;
  const auto table = lj_table(s);
  const int T = s->n_types;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
    const uint32_t i = static_cast<uint32_t>(ii);
    double gx = 0, gy = 0, gz = 0;
    const int ti = s->type[i] - 1;
    for (uint32_t j : lists[i]) {
      const double dxr = mic_diff_any<double>(pos.x(i), pos.x(j), Ld[0], hLd[0]);
      const double dyr = mic_diff_any<double>(pos.y(i), pos.y(j), Ld[1], hLd[1]);
      const double dzr = mic_diff_any<double>(pos.z(i), pos.z(j), Ld[2], hLd[2]);
      if (dxr * dxr + dyr * dyr + dzr * dzr >= rc2) {
        continue;
      }
      const LjCoeff& c = table[static_cast<size_t>(ti) * T + (s->type[j] - 1)];
      const Vec d = mic_vec(pos, i, j, Ld, hLd);
      const double u = vdot(d, d);
      const double sr2 = c.sigma2 / u;
      const double sr6 = sr2 * sr2 * sr2;
      const double sr12 = sr6 * sr6;
      const double sc = c.eps24 / u * (2.0 * sr12 - sr6);
      const double sp = -c.eps24 / (u * u) * (14.0 * sr12 - 4.0 * sr6);
      const double ex = dx[3 * i] - dx[3 * j], ey = dx[3 * i + 1] - dx[3 * j + 1],
                   ez = dx[3 * i + 2] - dx[3 * j + 2];
      const double de = d.x * ex + d.y * ey + d.z * ez;
      gx += sc * ex + 2.0 * sp * de * d.x;
      gy += sc * ey + 2.0 * sp * de * d.y;
      gz += sc * ez + 2.0 * sp * de * d.z;
    }
    df[3 * i] = gx;
    df[3 * i + 1] = gy;
    df[3 * i + 2] = gz;
  }
  std::vector<uint32_t> slot_of_tag(static_cast<size_t>(n) + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    slot_of_tag[s->tag[i]] = i;
  }
  for (uint32_t b = 0; b < s->n_bonds; ++b) {
    const uint32_t si = slot_of_tag[s->bonds[3 * b]];
    const uint32_t sj = slot_of_tag[s->bonds[3 * b + 1]];
    const int ty = static_cast<int>(s->bonds[3 * b + 2]) - 1;
    const double k = s->bond_k[ty], r0 = s->bond_r0[ty];
    const Vec d = mic_vec(pos, si, sj, Ld, hLd);
    const double r = std::sqrt(vdot(d, d));
    if (r == 0.0) {
      continue;
    }
    const double phi = -k * (r - r0) / r;
    const double dphi_r = -k * r0 / (r * r * r);
    const double ex = dx[3 * si] - dx[3 * sj], ey = dx[3 * si + 1] - dx[3 * sj + 1],
                 ez = dx[3 * si + 2] - dx[3 * sj + 2];
    const double de = d.x * ex + d.y * ey + d.z * ez;
    const double t[3] = {phi * ex + dphi_r * de * d.x, phi * ey + dphi_r * de * d.y,
                         phi * ez + dphi_r * de * d.z};
    for (int a = 0; a < 3; ++a) {
      df[3 * si + a] += t[a];
      df[3 * sj + a] -= t[a];
    }
  }
}

// tan (optional): the tangent run (forward mode) beside the state, NVE only:
// dv += h dF, dx += dt dv, dF = J(x) dx, permuted with the state by reorders.
struct Tangent {
  const double* dx0;
  const double* dv0;
  double* dx_out;
  double* dv_out;
};

template <bool D>
int run_impl(const orc_system_t* sys, const orc_run_params_t* p, const double* x0,
             const double* v0, double* pos_out, double* vel_out, int32_t* image_out,
             uint32_t* tags_out, orc_run_stats_t* st, const Tangent* tan = nullptr) {
  using CR = std::conditional_t<D, double, float>;
  using Plan = std::conditional_t<D, PlanD, orc_plan_t>;
  const double sin_floor = D ? kSinFloor64 : kSinFloor32;
#ifdef _OPENMP
  if (p->threads > 0) {
    omp_set_num_threads(p->threads);
  }
  st->threads_used = omp_get_max_threads();
#else
  st->threads_used = 1;
#endif
  const uint32_t n = sys->n;
  Plan plan;
  if (!make_plan_t<CR>(sys->lo, sys->hi, static_cast<CR>(sys->lj_cutoff),
                       static_cast<CR>(p->cutoff_buffer), n, &plan)) {
    return 1;
  }
  const BoxT<CR> box = make_box_t<CR>(sys->lo, sys->hi);
  std::vector<double> x(3 * static_cast<size_t>(n)), v(3 * static_cast<size_t>(n)), m(n);
  std::vector<int32_t> img(3 * static_cast<size_t>(n)), type(sys->type, sys->type + n);
  std::vector<uint32_t> tag(sys->tag, sys->tag + n);
  std::vector<float> mass_f(sys->mass, sys->mass + n);
  for (size_t k = 0; k < 3 * static_cast<size_t>(n); ++k) {
    x[k] = x0[k];
    v[k] = v0[k];
    img[k] = sys->image[k];
  }
  for (uint32_t i = 0; i < n; ++i) {
    m[i] = sys->mass[i];
  }
  std::vector<double> dx, dv, df;
  if (tan != nullptr) {
    dx.assign(tan->dx0, tan->dx0 + 3 * static_cast<size_t>(n));
    dv.assign(tan->dv0, tan->dv0 + 3 * static_cast<size_t>(n));
    df.assign(3 * static_cast<size_t>(n), 0.0);
  }
  orc_system_t cur = *sys;
  cur.type = type.data();
  cur.tag = tag.data();
  cur.mass = mass_f.data();

  std::vector<std::vector<uint32_t>> lists;
  std::vector<double> f(3 * static_cast<size_t>(n), 0.0), xref, imgref_dummy;
  std::vector<int32_t> imgref;
  SlotExclusions excl;
  bool excl_valid = false;
  uint64_t build_count = 0, dangerous = 0, reorders = 0, now = 0, last_sort = 0;
  bool checked_since_build = false;
  double neighbor_sum = 0.0;
  orc_observables_t obs{};

  auto rebuild = [&]() {
    ++build_count;
    const bool due = p->sort_every_steps != 0 ? now - last_sort >= p->sort_every_steps
                                              : build_count % p->sort_every_rebuild == 0;
    if (plan.path != ORC_PATH_DIRECT && due) {
      ++reorders;
      last_sort = now;
      std::vector<uint32_t> cell(n), order, counts, offsets;
      for (uint32_t i = 0; i < n; ++i) {
        const uint32_t ix = axis_cell<CR>(static_cast<CR>(x[3 * i]), box.lo[0], plan.inv_cell[0], plan.nx);
        const uint32_t iy = axis_cell<CR>(static_cast<CR>(x[3 * i + 1]), box.lo[1], plan.inv_cell[1], plan.ny);
        const uint32_t iz = axis_cell<CR>(static_cast<CR>(x[3 * i + 2]), box.lo[2], plan.inv_cell[2], plan.nz);
        cell[i] = ix + plan.nx * (iy + plan.ny * iz);
      }
      stable_order(n, cell.data(), plan.n_cells, order, counts, offsets);
      auto permute3 = [&](auto& a) {
        auto b = a;
        for (uint32_t s = 0; s < n; ++s) {
          for (int c = 0; c < 3; ++c) {
            a[3 * s + c] = b[3 * order[s] + c];
          }
        }
      };
      auto permute1 = [&](auto& a) {
        auto b = a;
        for (uint32_t s = 0; s < n; ++s) {
          a[s] = b[order[s]];
        }
      };
      permute3(x);
      permute3(v);
      permute3(img);
      if (tan != nullptr) {
        permute3(dx);
        permute3(dv);
      }
      permute1(m);
      permute1(type);
      permute1(tag);
      permute1(mass_f);
      excl_valid = false;
    }
    if (!excl_valid) {
      excl = slot_exclusions(&cur, tag.data());
      excl_valid = true;
    }
    build_lists<double>(n, PosD{x.data()}, box, plan, plan.path, excl,
                        static_cast<double>(plan.search_cutoff) * plan.search_cutoff, lists);
    size_t total = 0;
    for (const auto& l : lists) {
      total += std::min<size_t>(l.size(), p->max_neighbors);
      if (l.size() > p->max_neighbors) {
        g_error = "NeighborList overflow";
        return false;
      }
    }
    neighbor_sum += static_cast<double>(total) / n;
    xref = x;
    imgref = img;
    checked_since_build = false;
    return true;
  };

  const double dt = p->dt;
  const double half_skin = 0.5 * static_cast<double>(p->cutoff_buffer);
  const double thr2 = half_skin * half_skin;

  if (!rebuild()) {
    return 2;
  }
  compute_forces<double>(&cur, PosD{x.data()}, box, lists, f.data(), &obs, sin_floor);
  if (tan != nullptr) {
    force_tangent(&cur, PosD{x.data()}, box, lists, dx.data(), df.data());
  }

  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t step = 1; step <= p->steps; ++step) {
    now = step;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
      const size_t i = static_cast<size_t>(ii);
      const double h = 0.5 * dt / m[i];
      for (int a = 0; a < 3; ++a) {
        v[3 * i + a] += h * f[3 * i + a];
        double xa = x[3 * i + a] + dt * v[3 * i + a];
        const double c = std::floor((xa - box.lo[a]) / static_cast<double>(box.L[a]));
        xa -= c * static_cast<double>(box.L[a]);
        x[3 * i + a] = xa;
        img[3 * i + a] += static_cast<int32_t>(c);
        if (tan != nullptr) {
          dv[3 * i + a] += h * df[3 * i + a];
          dx[3 * i + a] += dt * dv[3 * i + a];
        }
      }
    }
    bool need = false;
    if (p->cutoff_buffer == 0.0f) {
      need = true;
    } else if (step % p->rebuild_check_every == 0) {
      const bool first = !checked_since_build;
      checked_since_build = true;
      int exceeded = 0;
#pragma omp parallel for schedule(static) reduction(| : exceeded)
      for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
        const size_t i = static_cast<size_t>(ii);
        double d2 = 0.0;
        for (int a = 0; a < 3; ++a) {
          const double d = (x[3 * i + a] - xref[3 * i + a]) +
                           static_cast<double>(img[3 * i + a] - imgref[3 * i + a]) * box.L[a];
          d2 += d * d;
        }
        exceeded |= d2 > thr2 ? 1 : 0;
      }
      if (exceeded) {
        need = true;
        if (p->rebuild_check_every > 1 && first) {
          ++dangerous;
        }
      }
    }
    if (need && !rebuild()) {
      return 2;
    }
    compute_forces<double>(&cur, PosD{x.data()}, box, lists, f.data(), &obs, sin_floor);
    if (tan != nullptr) {
      force_tangent(&cur, PosD{x.data()}, box, lists, dx.data(), df.data());
#pragma omp parallel for schedule(static)
      for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
        const size_t i = static_cast<size_t>(ii);
        const double h = 0.5 * dt / m[i];
        for (int a = 0; a < 3; ++a) dv[3 * i + a] += h * df[3 * i + a];
      }
    }
    double ke = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ke)
    for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii) {
      const size_t i = static_cast<size_t>(ii);
      const double h = 0.5 * dt / m[i];
      double v2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        v[3 * i + a] += h * f[3 * i + a];
        v2 += v[3 * i + a] * v[3 * i + a];
      }
      ke += 0.5 * m[i] * v2;
    }
    if (p->thermostat == 1) {
      const double lam = orc_berendsen_lambda(ke, n, p->kb, p->temperature, dt, p->tau);
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < static_cast<int64_t>(3 * static_cast<size_t>(n)); ++k) {
        v[static_cast<size_t>(k)] *= lam;
      }
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  st->wall_seconds = std::chrono::duration<double>(t1 - t0).count();
  st->build_count = build_count;
  st->dangerous_rebuild_count = dangerous;
  st->reorder_count = reorders;
  st->mean_neighbors = build_count ? neighbor_sum / static_cast<double>(build_count) : 0.0;
  double ke = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    const double v2 = v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2];
    ke += 0.5 * m[i] * v2;
  }
  st->final_ke = ke;
  st->final_obs = obs;
  std::copy(x.begin(), x.end(), pos_out);
  std::copy(v.begin(), v.end(), vel_out);
  std::copy(img.begin(), img.end(), image_out);
  std::copy(tag.begin(), tag.end(), tags_out);
  if (tan != nullptr) {
    std::copy(dx.begin(), dx.end(), tan->dx_out);
    std::copy(dv.begin(), dv.end(), tan->dv_out);
  }
  return 0;
}
}  // namespace

extern "C" int orc_run(const orc_system_t* sys, const orc_run_params_t* p,
                       double* pos_out, double* vel_out, int32_t* image_out,
                       uint32_t* tags_out, orc_run_stats_t* st) {
  const size_t m = 3 * static_cast<size_t>(sys->n);
  const std::vector<double> x0(sys->pos, sys->pos + m), v0(sys->vel, sys->vel + m);
  return run_impl<false>(sys, p, x0.data(), v0.data(), pos_out, vel_out, image_out, tags_out,
                         st);
}

extern "C" int orc_run_f64(const orc_system_t* sys, const orc_run_params_t* p,
                           const double* pos_in, const double* vel_in, double* pos_out,
                           double* vel_out, int32_t* image_out, uint32_t* tags_out,
                           orc_run_stats_t* st) {
  return run_impl<true>(sys, p, pos_in, vel_in, pos_out, vel_out, image_out, tags_out, st);
}

extern "C" int orc_run_tangent_f64(const orc_system_t* sys, const orc_run_params_t* p,
                                   const double* pos_in, const double* vel_in,
                                   const double* dpos_in, const double* dvel_in,
                                   double* pos_out, double* vel_out, int32_t* image_out,
                                   uint32_t* tags_out, double* dpos_out, double* dvel_out,
                                   orc_run_stats_t* st) {
  if (sys->n_angles != 0 || sys->n_dihedrals != 0 || p->thermostat != 0) {
    g_error = "the tangent run covers LJ pairs and harmonic bonds under NVE only.";
    return 1;
  }
  const Tangent tan{dpos_in, dvel_in, dpos_out, dvel_out};
  return run_impl<true>(sys, p, pos_in, vel_in, pos_out, vel_out, image_out, tags_out, st, &tan);
}
