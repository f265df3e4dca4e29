"""Batched replicas on the GPU, in both execution modes: one persistent
launch (one workgroup per replica) and phased per-step batched kernels over
64-atom chunks (a replica spread over many workgroups).

Each replica must match the single-system engine (same decisions and FP64
pair arithmetic, so identical initial forces) and the FP64 oracle to the same
force bar as the single path, in every mode and force team shape.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd import workloads as W
from beads_amd.capi import Batch, Engine

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5

# Configuration indices (beads_batch_config_info): persistent 512-lane
# workgroups, phased with 4x4, 2x8 and 8x2 force teams.
MODES = [1, 3, 4, 5]


def replicas():
    sizes = [300, 1800, 4000, 700, 2600, 520]
    dens = [0.9, 0.7, 0.8442, 0.65, 0.85, 0.6]
    temps = [1.0, 1.3, 0.9, 1.5, 0.8, 1.1]
    return [W.lj_box(n, d, t, seed=40 + k, berendsen=True, tau=0.5)
            for k, (n, d, t) in enumerate(zip(sizes, dens, temps))]


def _force_error(f_gpu, f_ref):
    scale = np.sqrt(np.mean(f_ref ** 2))
    return float(np.max(np.abs(f_gpu - f_ref) / np.maximum(np.abs(f_ref), scale)))


@pytest.mark.parametrize("config", MODES)
def test_initial_forces_match_single_engine_and_oracle(config):
    ws = replicas()
    with Batch(ws) as b:
        b.set_config(config)
        b.run(0)
        for r, w in enumerate(ws):
            st = b.state(r)
            with Engine(w) as e:
                e.initialize(observables=False)
                single = e.state()
            np.testing.assert_array_equal(st["tags"], single["tags"])
            np.testing.assert_array_equal(st["forces"], single["forces"])
            f_ref, _ = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
            assert _force_error(st["forces"].astype(np.float64), f_ref) <= FORCE_RTOL


@pytest.mark.parametrize("config", MODES)
def test_replicas_follow_single_engine_trajectories(config):
    ws = replicas()
    steps = 60
    with Batch(ws) as b:
        b.set_config(config)
        b.run(steps, observables_on_last=True)
        for r, w in enumerate(ws):
            st = b.state(r)
            bs = b.stats(r)
            assert bs.step == steps and bs.overflow == 0
            with Engine(w) as e:
                e.initialize(observables=False)
                e.advance(steps, observables_on_last=True)
                single = e.state()
                ss = e.stats()
            assert bs.build_count == ss.build_count and bs.reorder_count == ss.reorder_count
            order_b = np.argsort(st["tags"])
            order_s = np.argsort(single["tags"])
            L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
            dx = st["positions"][order_b].astype(np.float64) - single["positions"][order_s]
            dx -= np.round(dx / L) * L
            assert np.max(np.abs(dx)) < 1e-4, r
            assert bs.pe[0] == pytest.approx(ss.pe[0], rel=1e-5)
            assert bs.ke == pytest.approx(ss.ke, rel=1e-5)
            f_ref, _ = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
            assert _force_error(st["forces"].astype(np.float64), f_ref) <= FORCE_RTOL


def test_autotuner_picks_a_candidate_and_restores_state():
    ws = replicas()
    with Batch(ws) as b:
        chosen, secs = b.autotune(trial_steps=10)
        assert 0 <= chosen < len(secs) and b.config() == chosen
        assert secs[chosen] == min(secs)
        assert all(s > 0 for s in secs)
        assert {c["mode"] for c in b.configs()} == {"persistent", "phased"}
        b.run(0)
        st = b.state(2)
        np.testing.assert_array_equal(np.sort(st["tags"]), np.arange(1, ws[2].n + 1))
        # after the autotuner the batch restarts from the initial state
        assert b.stats(2).step == 0


def test_modes_agree_over_a_run_in_pieces():
    """Persistent and phased runs of the same replicas, the phased one in
    pieces (initial step, then 7 + 23 steps), end with the same counters and
    positions within 1e-5, and each matches its own restart-free run."""
    ws = replicas()
    out = {}
    for config in (1, 3):
        with Batch(ws) as b:
            b.set_config(config)
            b.run(0)
            b.run(7)
            b.run(23, observables_on_last=True)
            out[config] = [(b.stats(r), b.state(r)) for r in range(len(ws))]
    for r, w in enumerate(ws):
        (sa, xa), (sb, xb) = out[1][r], out[3][r]
        assert (sa.step, sa.build_count, sa.reorder_count, sa.dangerous_rebuild_count) == \
            (sb.step, sb.build_count, sb.reorder_count, sb.dangerous_rebuild_count), r
        L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
        oa, ob = np.argsort(xa["tags"]), np.argsort(xb["tags"])
        dx = xa["positions"][oa].astype(np.float64) - xb["positions"][ob]
        dx -= np.round(dx / L) * L
        assert np.max(np.abs(dx)) < 1e-5, r
        assert sa.pe[0] == pytest.approx(sb.pe[0], rel=1e-6)


@pytest.mark.parametrize("config", [1, 3])
def test_full_replica_set_against_single_engines_and_oracle(config):
    """The 512-replica set of the batched benchmark (N log-uniform in
    [500, 20000], 2.6M atoms in all) for 20 steps in one launch. On a sample
    that includes the largest and the smallest replica: build and reorder
    counts equal to a single engine's over the same 20 steps, positions within
    1e-4 of its trajectory, and forces within the force bar of the FP64
    oracle at the batch's own final positions. Every replica ends without
    overflow at step 20."""
    ws = W.replica_set(512, steps=20)
    sizes = np.array([w.n for w in ws])
    sample = sorted({int(np.argmax(sizes)), int(np.argmin(sizes)), 1, 100, 257, 511})
    steps = 20
    with Batch(ws) as b:
        b.set_config(config)
        b.run(steps, observables_on_last=True)
        for r in range(len(ws)):
            s = b.stats(r)
            assert s.step == steps and s.overflow == 0, r
        for r in sample:
            w = ws[r]
            st = b.state(r)
            bs = b.stats(r)
            with Engine(w) as e:
                e.initialize(observables=False)
                e.advance(steps, observables_on_last=True)
                single = e.state()
                ss = e.stats()
            assert (bs.build_count, bs.reorder_count) == (ss.build_count, ss.reorder_count), r
            order_b, order_s = np.argsort(st["tags"]), np.argsort(single["tags"])
            L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
            dx = st["positions"][order_b].astype(np.float64) - single["positions"][order_s]
            dx -= np.round(dx / L) * L
            assert np.max(np.abs(dx)) < 1e-4, r
            f_ref, _ = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
            assert _force_error(st["forces"].astype(np.float64), f_ref) <= FORCE_RTOL, r


def bonded_replicas():
    """Polymer melts (harmonic bonds, angles, dihedrals, exclusions) on each
    neighbour path beside a pair-only replica: 1,200 beads (cell path, a
    reorder every 2nd build), 3,000 beads (regular cell path), 216 beads
    (two cells per axis), and 480 beads with two angle and dihedral types (the second
    d = -1, n = 2) and exclusions to bond distance 3."""
    a = W.polymer_melt(n_chains=40, beads_per_chain=30, seed=11, jitter=0.05)
    a = a.with_(sort_every_rebuild=2)
    b = W.polymer_melt(n_chains=100, beads_per_chain=30, seed=12, jitter=0.05)
    c = W.polymer_melt(n_chains=18, beads_per_chain=12, density=7.5, seed=13, jitter=0.05)
    d = W.polymer_melt(n_chains=24, beads_per_chain=20, seed=14, jitter=0.05)
    angles, dihedrals = d.angles.copy(), d.dihedrals.copy()
    angles[1::2, 3] = 2
    dihedrals[1::2, 4] = 2
    d = d.with_(angles=angles, dihedrals=dihedrals, exclusion_distance=3,
                angle_k=np.array([25.0, 40.0]), angle_theta0=np.array([120.0, 105.0]),
                dihedral_k=np.array([1.0, 2.5]), dihedral_d=np.array([1, -1]),
                dihedral_n=np.array([1, 2]))
    lj = W.lj_box(700, 0.65, 1.5, seed=43, berendsen=True, tau=0.5)
    return [a, b, c, d, lj]


def _tag_sorted_dx(w, xa, xb):
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    oa, ob = np.argsort(xa["tags"]), np.argsort(xb["tags"])
    dx = xa["positions"][oa].astype(np.float64) - xb["positions"][ob]
    dx -= np.round(dx / L) * L
    return float(np.max(np.abs(dx)))


@pytest.mark.parametrize("config", MODES)
def test_bonded_replicas_match_single_engine_and_oracle(config):
    """Step-0 forces of melts with listed terms and exclusions equal the
    single engine's bit for bit and the FP64 oracle's to the force bar; the
    four energies and virials match the engine's; after 60 steps the build
    and reorder counts equal the engine's and positions agree within 1e-4."""
    ws = bonded_replicas()
    steps = 60
    with Batch(ws) as b:
        b.set_config(config)
        b.run(0)
        initial = [(b.state(r), b.stats(r)) for r in range(len(ws))]
        b.run(steps, observables_on_last=True)
        final = [(b.state(r), b.stats(r)) for r in range(len(ws))]
    for r, w in enumerate(ws):
        with Engine(w) as e:
            e.initialize(observables=True)
            s0, ss0 = e.state(), e.stats()
            e.advance(steps, observables_on_last=True)
            s1, ss1 = e.state(), e.stats()
        (b0, bs0), (b1, bs1) = initial[r], final[r]
        np.testing.assert_array_equal(b0["tags"], s0["tags"])
        np.testing.assert_array_equal(b0["forces"], s0["forces"])
        for q in range(4):
            assert bs0.pe[q] == pytest.approx(ss0.pe[q], rel=1e-6, abs=1e-9), (r, q)
            assert bs0.virial[q] == pytest.approx(ss0.virial[q], rel=1e-6, abs=1e-9), (r, q)
        f_ref, obs = O.forces(w, positions=b0["positions"], tags=b0["tags"], types=b0["types"])
        assert _force_error(b0["forces"].astype(np.float64), f_ref) <= FORCE_RTOL, r
        for q, key in enumerate(("pair_pe", "bond_pe", "angle_pe", "dihedral_pe")):
            assert bs0.pe[q] == pytest.approx(obs[key], rel=1e-5, abs=1e-6), (r, key)
        assert bs1.step == steps and bs1.overflow == 0
        assert (bs1.build_count, bs1.reorder_count) == (ss1.build_count, ss1.reorder_count), r
        assert _tag_sorted_dx(w, b1, s1) < 1e-4, r
        for q in range(4):
            assert bs1.pe[q] == pytest.approx(ss1.pe[q], rel=1e-4, abs=1e-6), (r, q)
        f_ref, _ = O.forces(w, positions=b1["positions"], tags=b1["tags"], types=b1["types"])
        assert _force_error(b1["forces"].astype(np.float64), f_ref) <= FORCE_RTOL, r


def test_bonded_replicas_reorder_and_modes_agree():
    """The 1,200-bead melt reorders every 2nd build; persistent and phased
    runs of the bonded set, the phased one in pieces, end with the same
    counters and positions within 1e-5; the reordered replica's reorder count
    is positive, so its records went through the remap."""
    ws = bonded_replicas()
    out = {}
    for config in (1, 3):
        with Batch(ws) as b:
            b.set_config(config)
            b.run(0)
            b.run(11)
            b.run(29, observables_on_last=True)
            out[config] = [(b.stats(r), b.state(r)) for r in range(len(ws))]
    assert out[1][0][0].reorder_count > 0
    for r, w in enumerate(ws):
        (sa, xa), (sb, xb) = out[1][r], out[3][r]
        assert (sa.step, sa.build_count, sa.reorder_count, sa.dangerous_rebuild_count) == \
            (sb.step, sb.build_count, sb.reorder_count, sb.dangerous_rebuild_count), r
        assert _tag_sorted_dx(w, xa, xb) < 1e-5, r
        for q in range(4):
            assert sa.pe[q] == pytest.approx(sb.pe[q], rel=1e-6, abs=1e-9), (r, q)


def _exclusion_case(name):
    if name == "hub":
        from gpu_cases import hub_lattice

        return [hub_lattice(12, 2, 3, True),
                W.polymer_melt(n_chains=18, beads_per_chain=12, density=7.5, seed=22,
                               jitter=0.05)]
    distance = {"chains2": 2, "chains3": 3}[name]
    ws = [W.polymer_melt(n_chains=40, beads_per_chain=30, seed=21, jitter=0.05),
          W.polymer_melt(n_chains=18, beads_per_chain=12, density=7.5, seed=22, jitter=0.05)]
    return [w.with_(exclusion_distance=distance) for w in ws]


@pytest.mark.parametrize("case", ["chains2", "chains3", "hub"])
def test_bonded_replica_exclusion_variants(case):
    """Each exclusion variant of the phased build: up to 4 partners in
    registers (chains, bond distance 2), up to 8 (chains, distance 3: 6
    partners) and read from memory (a hub of 12 two-bead arms at distance 3,
    beside a melt: the hub and its first beads have 24 excluded partners).
    Step-0 forces equal the single engine's bit for bit in both modes, and
    after 30 steps the build and reorder counts are the engine's and
    positions agree within 1e-4."""
    ws = _exclusion_case(case)
    steps = 30
    for config in (1, 3):
        with Batch(ws) as b:
            b.set_config(config)
            b.run(0)
            initial = [b.state(r) for r in range(len(ws))]
            b.run(steps)
            final = [(b.state(r), b.stats(r)) for r in range(len(ws))]
        for r, w in enumerate(ws):
            with Engine(w) as e:
                e.initialize(observables=False)
                s0 = e.state()
                e.advance(steps)
                s1, ss1 = e.state(), e.stats()
            np.testing.assert_array_equal(initial[r]["forces"], s0["forces"])
            (b1, bs1) = final[r]
            assert (bs1.build_count, bs1.reorder_count) == (ss1.build_count, ss1.reorder_count)
            assert _tag_sorted_dx(w, b1, s1) < 1e-4, (config, r)
