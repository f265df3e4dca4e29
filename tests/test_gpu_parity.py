"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Bars (stated once, used below):
  * cell assignment, stable in-cell order and neighbour sets: bit-exact;
  * forces: |f_gpu - f_fp64| <= FORCE_RTOL * max(|f_fp64|, f_rms) per component,
    f_rms the root-mean-square force component of the configuration;
  * energies / virials: relative FP64 difference <= ENERGY_RTOL of the
    summed magnitude of the per-term contributions.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd import workloads as W
from beads_amd.capi import Engine
from gpu_cases import cases

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-5

CASES = cases()


def _engine(w: W.Workload) -> Engine:
    e = Engine(w)
    e.initialize(observables=True)
    return e


def _force_error(f_gpu: np.ndarray, f_ref: np.ndarray) -> float:
    scale = np.sqrt(np.mean(f_ref ** 2))
    denom = np.maximum(np.abs(f_ref), scale)
    return float(np.max(np.abs(f_gpu - f_ref) / denom))


@pytest.mark.parametrize("name", list(CASES))
def test_cells_and_neighbors_bit_exact(name):
    w = CASES[name]
    plan = O.plan(w)
    with _engine(w) as e:
        st = e.state()
        stats = e.stats()
        assert stats.path == plan.path
        assert (stats.nx, stats.ny, stats.nz) == (plan.nx, plan.ny, plan.nz)
        counts, lst = e.neighbors()
        if plan.path != O.PATH_DIRECT:
            cells = e.cells(plan.n_cells)
            ref_cells = O.cell_ids(w, positions=st["positions"])
            np.testing.assert_array_equal(cells["cell_of"], ref_cells)
            order, ref_counts, offsets = O.stable_cell_order(ref_cells, plan.n_cells,
                                                             st["positions"])
            np.testing.assert_array_equal(cells["cell_count"], ref_counts)
            np.testing.assert_array_equal(cells["cell_start"], offsets)
            np.testing.assert_array_equal(cells["cell_atoms"], order)
    rc, rl, ovf = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert not ovf
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        got = np.sort(lst[: counts[i], i])
        want = np.sort(rl[i, : rc[i]])
        assert np.array_equal(got, want), f"atom {i}: neighbour sets differ"


@pytest.mark.parametrize("name", list(CASES))
def test_forces_and_observables(name):
    w = CASES[name]
    with _engine(w) as e:
        st = e.state()
        stats = e.stats()
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    err = _force_error(st["forces"].astype(np.float64), f_ref)
    assert err <= FORCE_RTOL, f"force error {err:.3e}"
    names = ("pair", "bond", "angle", "dihedral")
    for q, term in enumerate(names):
        ref_pe = obs[f"{term}_pe"]
        ref_w = obs[f"{term}_virial"]
        assert abs(stats.pe[q] - ref_pe) <= ENERGY_RTOL * max(1.0, abs(ref_pe)), (term, stats.pe[q], ref_pe)
        assert abs(stats.virial[q] - ref_w) <= ENERGY_RTOL * max(1.0, abs(ref_w)) + 1e-6 * w.n, (
            term, stats.virial[q], ref_w)
    ke = O.kinetic_energy(st["velocities"], st["masses"])
    assert abs(stats.ke - ke) <= 1e-9 * ke


@pytest.mark.parametrize("name", ["lj_regular", "melt"])
def test_one_step_matches_oracle(name):
    w = CASES[name].with_(berendsen=False)
    with _engine(w) as e:
        s0 = e.state()
        e.advance(1)
        s1 = e.state()
    # Slot order is unchanged by one step unless the step rebuilt with a sort.
    assert np.array_equal(s0["tags"], s1["tags"])
    po, vo, io = O.vv_pre(w, s0["positions"], s0["velocities"], s0["forces"], s0["images"])
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    dx = s1["positions"] - po
    dx -= np.round(dx / L) * L
    assert np.max(np.abs(dx)) <= 2e-6 * float(L.max())
    f1, _ = O.forces(w, positions=s1["positions"], tags=s1["tags"], types=s1["types"])
    assert _force_error(s1["forces"].astype(np.float64), f1) <= FORCE_RTOL
    v_ref = vo + 0.5 * w.dt / w.masses.astype(np.float64)[:, None] * s1["forces"].astype(np.float64)
    vscale = np.sqrt(np.mean(v_ref ** 2))
    assert np.max(np.abs(s1["velocities"] - v_ref)) <= 1e-5 * vscale
    # Image counts agree except where a coordinate sits within rounding of a face.
    lo = w.box_bound[0].astype(np.float64)
    edge = np.minimum(np.abs(po - lo), np.abs(po - (lo + L))).min(axis=1) < 1e-4
    np.testing.assert_array_equal(s1["images"][~edge], io[~edge])


def test_berendsen_lambda_and_kinetic_energy():
    w = CASES["lj_regular"].with_(berendsen=True, temperature=1.4, tau=0.5)
    with _engine(w) as e:
        s0 = e.state()
        e.advance(1)
        stats = e.stats()
        s1 = e.state()
    # Stored velocities carry the second half-kick; the factor is still pending.
    ke = O.kinetic_energy(s1["velocities"], s1["masses"])
    lam = O.berendsen_lambda(ke, w.n, w.kb, w.temperature, w.dt, w.tau)
    assert abs(stats.lambda_ - lam) <= 1e-9
    assert lam > 1.0  # T0 above the initial temperature heats the system
    assert np.array_equal(s0["tags"], s1["tags"])


def test_rebuilds_reorders_and_long_run_stay_consistent():
    w = CASES["melt"].with_(sort_every_rebuild=1)
    with _engine(w) as e:
        e.advance(200, observables_on_last=True)
        stats = e.stats()
        st = e.state()
    assert stats.build_count >= 2 and stats.reorder_count >= 1
    assert stats.overflow == 0
    assert sorted(st["tags"].tolist()) == list(range(1, w.n + 1))
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert _force_error(st["forces"].astype(np.float64), f_ref) <= FORCE_RTOL
    # Masses and types travel with their tags through every reorder.
    order = np.argsort(st["tags"])
    np.testing.assert_array_equal(st["types"][order], w.types)
    np.testing.assert_array_equal(st["masses"][order], w.masses)


@pytest.mark.parametrize("n", [1200, 4000])  # 4 and 6 cells per axis
def test_neighbor_overflow_caps_lists_in_scan_order(n):
    """max_neighbors below the real counts: the overflow flag is set and every
    stored list holds the first max_neighbors accepted candidates in the
    builder's scan order, the same entries as the oracle's capped lists."""
    w = W.lj_box(n, 0.8, 1.0, seed=4).with_(max_neighbors=8)
    with Engine(w) as e:
        e.initialize(observables=False)
        stats = e.stats()
        st = e.state()
        counts, lst = e.neighbors()
    assert stats.overflow == 1
    rc, rl, ovf = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"],
                                   types=st["types"], max_neighbors=8)
    assert ovf
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        assert np.array_equal(np.sort(lst[: counts[i], i]), np.sort(rl[i, : rc[i]])), i


def test_nve_energy_conservation():
    w = W.lj_fluid_fcc(cells=8, steps=1000)  # 2048 beads, 4^3 cells
    with _engine(w) as e:
        s = e.stats()
        e0 = s.ke + sum(s.pe)
        drifts = []
        for _ in range(10):
            e.advance(100, observables_on_last=True)
            s = e.stats()
            drifts.append(s.ke + sum(s.pe) - e0)
    per_step = abs(drifts[-1]) / abs(e0) / 1000
    assert per_step < 1e-6, (drifts, e0)


@pytest.mark.parametrize("name", ["lj_regular", "melt"])
def test_step_graph_replays_match_eager_launches(name, monkeypatch):
    """Stretches of 16+ steps replay one captured hipGraph; with eager reads and
    synchronisations between the replays every replay must give exactly what
    launching the same kernels one by one gives."""
    w = CASES[name]
    out = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("BEADS_AMD_GRAPHS", graphs)
        states = []
        with _engine(w) as e:
            for chunk in (40, 37, 48):
                e.advance(chunk, observables_on_last=True)
                states.append((e.state(), e.stats()))
        out[graphs] = states
    for (a, sa), (b, sb) in zip(out["1"], out["0"]):
        assert sa.step == sb.step
        np.testing.assert_array_equal(a["positions"], b["positions"])
        np.testing.assert_array_equal(a["velocities"], b["velocities"])
        np.testing.assert_array_equal(a["forces"], b["forces"])
        assert list(sa.pe) == list(sb.pe) and sa.build_count == sb.build_count


def test_lj_fluid_fcc_config_full_run():
    """The 4,000-bead FCC LJ fluid exactly as configured for the small
    benchmark (check every 10 steps, 1000 steps of dt 0.005, NVE): total
    energy drift below 1e-6 per step, and at the end the neighbour sets are
    bit-exact and the forces within the force bar of the FP64 oracle at the
    engine's final positions. The lattice starts with zero force on every
    bead, so the force bar's scale there is the end state's rms force."""
    w = W.lj_fluid_fcc()
    assert w.n == 4000 and w.rebuild_check_every == 10 and w.steps == 1000
    with _engine(w) as e:
        s = e.stats()
        e0 = s.ke + sum(s.pe)
        for _ in range(10):
            e.advance(100, observables_on_last=True)
        s = e.stats()
        st = e.state()
        counts, lst = e.neighbors()
    assert s.step == 1000 and s.overflow == 0
    assert s.build_count >= 2 and s.dangerous_rebuild_count < s.build_count
    assert abs(s.ke + sum(s.pe) - e0) / abs(e0) / 1000 < 1e-6
    rc, rl, _ = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        assert np.array_equal(np.sort(lst[: counts[i], i]), np.sort(rl[i, : rc[i]])), i
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert _force_error(st["forces"].astype(np.float64), f_ref) <= FORCE_RTOL
    assert abs(s.pe[0] - obs["pair_pe"]) <= ENERGY_RTOL * abs(obs["pair_pe"])
