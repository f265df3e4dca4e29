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
