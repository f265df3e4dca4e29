"""How many r < r_c decisions the exact-intermediate minimum image changes.

The force pass (and the oracle) form a pair's minimum-image displacement as
(xi - L) - xj or xi - (xj - L) when the raw difference crosses half the box,
rounded once; the reference forms MIC(xi - xj) in binary32 (pair_lj.cu:44-54),
rounding at the scale of L first. Both then take r^2 = dx dx + dy dy + dz dz
in binary32 and test r^2 < r_c^2 (pair_lj.cu:54). The two displacements can
differ only for pairs across a periodic boundary, by about an ulp of L; this
counts, over every stored pair of the benchmark systems (seeded positions),
the pairs whose decision differs, and bounds it. CPU only.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from beads_amd import workloads as W

f32 = np.float32


def _mic_reference(xi, xj, L, hL):
    d = xi - xj
    return np.where(d > hL, d - L, np.where(d < -hL, d + L, d))


def _mic_exact(xi, xj, L, hL):
    d = xi - xj
    return np.where(d > hL, (xi - L) - xj, np.where(d < -hL, xi - (xj - L), d))


def _decisions(w):
    counts, lst, ovf = O.neighbor_lists(w)
    assert not ovf
    pos = np.asarray(w.positions, dtype=f32)
    lo = np.asarray(w.box_bound[0], dtype=f32)
    L = (np.asarray(w.box_bound[1], dtype=f32) - lo).astype(f32)
    hL = (f32(0.5) * L).astype(f32)
    i = np.repeat(np.arange(w.n), counts.astype(np.int64))
    mask = np.arange(lst.shape[1])[None, :] < counts[:, None]
    j = lst[mask].astype(np.int64)
    rc2 = f32(w.cutoff) * f32(w.cutoff)
    out = {}
    for name, mic in (("reference", _mic_reference), ("exact", _mic_exact)):
        d = [mic(pos[i, a], pos[j, a], L[a], hL[a]).astype(f32) for a in range(3)]
        r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        out[name] = r2 < rc2
    raw = pos[i] - pos[j]
    crossing = np.any(np.abs(raw) > hL, axis=1)
    return out["reference"], out["exact"], crossing


@pytest.mark.parametrize("make, bound", [
    (lambda: W.lj_fluid_fcc(), 0),
    (lambda: W.polymer_melt(n_chains=2000), 0),
    (lambda: W.lj_mixture(n=200_000), 0),
])
def test_exact_minimum_image_changes_almost_no_cutoff_decision(make, bound):
    w = make()
    ref, exact, crossing = _decisions(w)
    differ = int(np.count_nonzero(ref != exact))
    print(f"{w.name}: {ref.size} stored pairs, {int(crossing.sum())} across a boundary, "
          f"{int(ref.sum())} inside r_c, {differ} decisions differ")
    assert differ == int(np.count_nonzero((ref != exact) & crossing))  # only boundary pairs
    assert differ <= bound
