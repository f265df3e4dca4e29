"""Parity at the full benchmark sizes: the 1,000,000-bead polymer melt and the
4,000,000-bead four-type LJ mixture, initial neighbour build and forces.

The FP64 OpenMP oracle evaluates these sizes in seconds, so the bars are the
same as for the small cases: neighbour counts equal for every atom and sorted
neighbour sets equal for a seeded sample of 20,000 atoms (bit-exact); forces
|f - f_fp64| <= 1e-5 max(|f_fp64|, f_rms) per component; energies and virials
relative 1e-5. Newton's third law holds for the pair sums (each pair value is
computed from the same exactly-rounded displacement from both sides), so the
summed force is zero to binary32 rounding of the per-atom totals.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd import workloads as W
from beads_amd.capi import Engine

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-5


def _force_error(f_gpu, f_ref):
    scale = np.sqrt(np.mean(f_ref ** 2))
    return float(np.max(np.abs(f_gpu - f_ref) / np.maximum(np.abs(f_ref), scale)))


@pytest.mark.parametrize("name", ["melt_1m", "mixture_4m"])
def test_full_size_lists_forces_and_observables(name):
    w = W.polymer_melt() if name == "melt_1m" else W.lj_mixture()
    print(f"{name}: {w.n} beads", flush=True)
    with Engine(w) as e:
        e.initialize(observables=True)
        st = e.state()
        stats = e.stats()
        counts, lst = e.neighbors()
    assert stats.overflow == 0
    print("device done", flush=True)
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    f = st["forces"].astype(np.float64)
    assert _force_error(f, f_ref) <= FORCE_RTOL
    for q, term in enumerate(("pair", "bond", "angle", "dihedral")):
        ref_pe, ref_w = obs[f"{term}_pe"], obs[f"{term}_virial"]
        assert abs(stats.pe[q] - ref_pe) <= ENERGY_RTOL * max(1.0, abs(ref_pe)), (term, stats.pe[q], ref_pe)
        assert abs(stats.virial[q] - ref_w) <= ENERGY_RTOL * max(1.0, abs(ref_w)) + 1e-6 * w.n
    total = np.abs(f.sum(axis=0))
    assert np.all(total <= 1e-6 * np.abs(f).sum(axis=0)), total
    print("forces done", flush=True)
    rc, rl, ovf = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert not ovf
    np.testing.assert_array_equal(counts, rc)
    rng = np.random.default_rng(17)
    for i in rng.choice(w.n, size=20000, replace=False):
        got = np.sort(lst[: counts[i], i])
        want = np.sort(rl[i, : rc[i]])
        assert np.array_equal(got, want), f"atom {i}: neighbour sets differ"
