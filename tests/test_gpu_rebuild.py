"""Rebuild policy on the GPU: the displacement test and the check cadence.

The engine decides on the device whether a step rebuilds (no host wait). Its
decisions are held to the reference's rule (rebuild_tracker.cu:55-94,
neighbor_system.cpp:79-105, 198-216) three ways:
  * closed form: one bead drifting at constant velocity through a lattice with
    no forces crosses half the skin at a known step, so build_count and
    dangerous_rebuild_count follow from the check interval (and are the FP64
    oracle run's counts too);
  * zero skin: a rebuild after every step, none dangerous;
  * step by step on interacting systems: after every step the engine rebuilt
    exactly when the oracle's binary32 restatement of the displacement test,
    applied to the engine's own positions and images against those at the last
    build, says the threshold was exceeded.
Counts are integers and decisions booleans: all bars are exact.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd import workloads as W
from beads_amd.capi import Batch, Engine
from gpu_cases import drifting_lattice, expected_build_steps, expected_builds, expected_reorders

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("v_dt,check_every,steps", [
    (0.15 / 10.5, 1, 60),
    (0.15 / 10.5, 3, 60),
    (0.15 / 2.5, 3, 24),
    (0.15 / 10.5, 10, 60),
    (0.15 / 9.5, 10, 60),
    (0.15 / 4.5, 7, 60),
])
def test_rebuild_cadence_closed_form(v_dt, check_every, steps):
    w = drifting_lattice(v_dt, check_every=check_every, steps=steps)
    want = expected_builds(0.15 / v_dt, check_every, steps, 0.3)
    _, ost = O.run(w)
    assert (ost.build_count, ost.dangerous_rebuild_count) == want
    with Engine(w) as e:
        e.initialize(observables=False)
        e.advance(steps)
        st = e.stats()
    assert st.step == steps
    assert (st.build_count, st.dangerous_rebuild_count) == want


@pytest.mark.parametrize("sort_every_rebuild,sort_every_steps", [(3, 0), (10, 1), (10, 12),
                                                                  (10, 25)])
def test_reorder_cadence_closed_form(sort_every_rebuild, sort_every_steps):
    """Sorting rebuilds: every sort_every_rebuild-th build (neighbor_system.cpp:
    198-202), or with the step-based option the first build at least
    sort_every_steps steps after the last sort; the plain engine, the batched
    replica kernel and the oracle run all follow the closed form."""
    v_dt = 0.15 / 4.5
    steps = 60
    w = drifting_lattice(v_dt, check_every=1, steps=steps).with_(
        sort_every_rebuild=sort_every_rebuild, sort_every_steps=sort_every_steps)
    build_steps, _ = expected_build_steps(0.15 / v_dt, 1, steps, 0.3)
    want = expected_reorders(build_steps, sort_every_rebuild, sort_every_steps)
    _, ost = O.run(w)
    assert (ost.build_count, ost.reorder_count) == (len(build_steps), want)
    with Engine(w) as e:
        e.initialize(observables=False)
        e.advance(steps)
        st = e.stats()
    assert (st.build_count, st.reorder_count) == (len(build_steps), want)
    with Batch([w]) as b:
        b.run(steps)
        bs = b.stats(0)
    assert (bs.build_count, bs.reorder_count) == (len(build_steps), want)


def test_step_based_sort_cadence_through_the_public_api(tmp_path):
    """neighbor sort_every_steps travels from set_neighbor through the spec dict
    and the C ABI: the run-start log names it, and on an interacting system
    (the mixture, skin 0.1, rebuilding every few steps) the engine sorts
    exactly at the builds the rule selects from its own build steps."""
    w = W.lj_mixture(n=20000, seed=2).with_(steps=40, sort_every_steps=15)
    sim = w.to_simulation()
    sim.set_logging(echo="log", prefix=str(tmp_path / "run"))
    sim.execute()
    log = (tmp_path / "run.log").read_text()
    line = next(ln for ln in log.splitlines() if "sort_every_steps" in ln)
    assert line.rstrip().endswith(": 15"), line
    with Engine(w) as e:
        e.initialize(observables=False)
        build_steps, builds = [0], 1
        for step in range(1, 41):
            e.advance(1)
            st = e.stats()
            if st.build_count > builds:
                build_steps.append(step)
                builds = st.build_count
    assert len(build_steps) > 4
    assert st.reorder_count == expected_reorders(build_steps, w.sort_every_rebuild, 15) >= 2


def test_zero_skin_rebuilds_after_every_step():
    steps = 40
    w = drifting_lattice(0.01, skin=0.0, steps=steps)
    with Engine(w) as e:
        e.initialize(observables=False)
        e.advance(steps)
        st = e.stats()
    assert (st.build_count, st.dangerous_rebuild_count) == (steps + 1, 0)
    # the same at zero skin on an interacting system, where the list must also
    # stay exact: the sets equal the oracle's at the final positions
    w = W.lj_box(1200, 0.8442, 1.5, seed=12, skin=0.0)
    with Engine(w) as e:
        e.initialize(observables=False)
        e.advance(25)
        st = e.stats()
        s = e.state()
        counts, lst = e.neighbors()
    assert (st.build_count, st.dangerous_rebuild_count) == (26, 0)
    rc, rl, _ = O.neighbor_lists(w, positions=s["positions"], tags=s["tags"], types=s["types"])
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        assert np.array_equal(np.sort(lst[: counts[i], i]), np.sort(rl[i, : rc[i]])), i


def _by_tag(state):
    order = np.argsort(state["tags"])
    return state["positions"][order], state["images"][order]


@pytest.mark.parametrize("name", ["lj_hot", "melt"])
def test_each_step_rebuilds_exactly_when_binary32_test_exceeds(name):
    if name == "lj_hot":
        w = W.lj_box(2000, 0.8442, 2.0, seed=21, check_every=1)
        steps = 120
    else:
        w = W.polymer_melt(n_chains=60, beads_per_chain=50, seed=9)
        steps = 80
    w = w.with_(sort_every_rebuild=2)  # reorders in between: compare by tag
    decisions = []
    with Engine(w) as e:
        e.initialize(observables=False)
        ref_pos, ref_img = _by_tag(e.state())
        builds = e.stats().build_count
        for _ in range(steps):
            e.advance(1)
            pos, img = _by_tag(e.state())
            st = e.stats()
            rebuilt = st.build_count == builds + 1
            assert st.build_count in (builds, builds + 1)
            want = O.displacement_exceeds(w, pos, img, ref_pos, ref_img)
            assert rebuilt == want, f"step {st.step}: engine rebuilt={rebuilt}, oracle {want}"
            decisions.append(rebuilt)
            if rebuilt:
                ref_pos, ref_img, builds = pos, img, st.build_count
    assert any(decisions) and not all(decisions)  # both outcomes were exercised
