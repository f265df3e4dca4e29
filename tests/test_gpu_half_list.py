"""Half neighbour list on the GPU: each pair stored once (j > i by slot), the
reaction added with FP64 atomics before the fused force pass.

Same bars as the full-list path: the stored half sets are bit-exact against
the oracle's full sets restricted to j > i; forces |f - f_fp64| <= 1e-5 *
max(|f_fp64|, f_rms); energies and virials relative 1e-5. Atomic summation
order varies between runs, so the half path is compared with the full path
to the same force bar rather than bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd.capi import Engine
from gpu_cases import cases

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-5

CASES = cases()


def _force_error(f_gpu, f_ref):
    scale = np.sqrt(np.mean(f_ref ** 2))
    return float(np.max(np.abs(f_gpu - f_ref) / np.maximum(np.abs(f_ref), scale)))


@pytest.mark.parametrize("name", list(CASES))
def test_half_list_sets_forces_and_observables(name):
    w = CASES[name].with_(half_list=True)
    with Engine(w) as e:
        e.initialize(observables=True)
        st = e.state()
        stats = e.stats()
        counts, lst = e.neighbors()
    rc, rl, ovf = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert not ovf
    for i in range(w.n):
        want = np.sort(rl[i, : rc[i]])
        want = want[want > i]
        got = np.sort(lst[: counts[i], i])
        assert np.array_equal(got, want), f"atom {i}: half sets differ"
    assert int(counts.sum()) * 2 == int(rc.sum())
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert _force_error(st["forces"].astype(np.float64), f_ref) <= FORCE_RTOL
    for q, term in enumerate(("pair", "bond", "angle", "dihedral")):
        ref_pe, ref_w = obs[f"{term}_pe"], obs[f"{term}_virial"]
        assert abs(stats.pe[q] - ref_pe) <= ENERGY_RTOL * max(1.0, abs(ref_pe)), (term, stats.pe[q], ref_pe)
        assert abs(stats.virial[q] - ref_w) <= ENERGY_RTOL * max(1.0, abs(ref_w)) + 1e-6 * w.n


@pytest.mark.parametrize("name", ["lj_regular", "melt"])
def test_half_list_trajectory_follows_full_list(name):
    w = CASES[name]
    steps = 40
    out = {}
    for half in (False, True):
        with Engine(w.with_(half_list=half)) as e:
            e.initialize(observables=False)
            e.advance(steps, observables_on_last=True)
            st = e.state()
            out[half] = (st, e.stats())
    (sf, tf), (sh, th) = out[False], out[True]
    assert th.build_count == tf.build_count and th.reorder_count == tf.reorder_count
    of, oh = np.argsort(sf["tags"]), np.argsort(sh["tags"])
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    dx = sh["positions"][oh].astype(np.float64) - sf["positions"][of]
    dx -= np.round(dx / L) * L
    assert np.max(np.abs(dx)) < 1e-4
    assert th.pe[0] == pytest.approx(tf.pe[0], rel=1e-5)
    assert th.ke == pytest.approx(tf.ke, rel=1e-5)
    f_ref, _ = O.forces(w, positions=sh["positions"], tags=sh["tags"], types=sh["types"])
    assert _force_error(sh["forces"].astype(np.float64), f_ref) <= FORCE_RTOL
