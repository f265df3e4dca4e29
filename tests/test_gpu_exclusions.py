"""Bonded exclusions on the GPU: every variant of the neighbour build.

The build keeps a bead's excluded partners in registers when the largest
per-bead count is at most 4 or at most 8, and reads them from memory above
that. Each variant runs here on a topology that selects it, for exclusion
distances 1, 2 and 3 (exclusions.cu:514-598, native_spec_parser.cpp:730-753):
  * the reference's 8-bead exclusion example on the direct path (distance 1:
    two partners, 2: four, 3: six);
  * serpentine chains in an LJ lattice on the regular cell path (distance 1:
    two partners; 3: six, the 8-entry variant);
  * six-armed stars in the same lattice (12 partners at the hub for distances
    2 and 3: the memory variant).
Bars: neighbour counts and sets bit-exact against the FP64 oracle's build,
forces within 1e-5 of FP64 per component, bond energy relative 1e-5, at the
initial build and again after 100 steps at zero skin, which rebuild and
reorder after every step (the exclusion records follow the beads through each
permutation).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd.capi import Engine
from gpu_cases import bonded_lattice, example_chain, exclusion_counts

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5

CASES = {
    "example_d1": (lambda: example_chain(1), (2, 2)),
    "example_d2": (lambda: example_chain(2), (4, 4)),
    "example_d3": (lambda: example_chain(3), (6, 6)),
    "chains_d1": (lambda: bonded_lattice("chains", 1), (2, 2)),
    "chains_d3": (lambda: bonded_lattice("chains", 3), (5, 8)),
    "stars_d2": (lambda: bonded_lattice("stars", 2), (9, 10 ** 9)),
    "stars_d3": (lambda: bonded_lattice("stars", 3), (9, 10 ** 9)),
}


def _check_against_oracle(w, e):
    st = e.state()
    counts, lst = e.neighbors()
    assert e.stats().overflow == 0
    rc, rl, ovf = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    assert not ovf
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        assert np.array_equal(np.sort(lst[: counts[i], i]), np.sort(rl[i, : rc[i]])), i
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    f = st["forces"].astype(np.float64)
    scale = max(float(np.sqrt(np.mean(f_ref ** 2))), 1e-30)
    err = float(np.max(np.abs(f - f_ref) / np.maximum(np.abs(f_ref), scale)))
    assert err <= FORCE_RTOL, err
    return obs


@pytest.mark.parametrize("name", list(CASES))
def test_exclusion_variant_sets_and_forces(name):
    make, (lo, hi) = CASES[name]
    w = make()
    ecap = int(exclusion_counts(w).max())
    assert lo <= ecap <= hi, ecap  # the topology selects the intended build variant
    with Engine(w) as e:
        e.initialize(observables=True)
        obs = _check_against_oracle(w, e)
        pe = e.stats().pe
        assert abs(pe[1] - obs["bond_pe"]) <= 1e-5 * max(1.0, abs(obs["bond_pe"]))
    if w.n > 100:
        # zero skin: a rebuild (with a reorder) after every step, so the final
        # list is the one the oracle builds at the final positions
        w2 = w.with_(sort_every_rebuild=1, cutoff_buffer=0.0)
        with Engine(w2) as e:
            e.initialize(observables=False)
            e.advance(100, observables_on_last=True)
            st = e.stats()
            assert st.build_count == 101 and st.reorder_count >= 100
            _check_against_oracle(w2, e)
