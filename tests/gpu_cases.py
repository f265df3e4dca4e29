"""Small seeded workloads that cover every neighbour path and force term.

Shared by the GPU parity tests; sizes keep the FP64 oracle to seconds.
"""

from __future__ import annotations

import numpy as np

from beads_amd import workloads as W


def elongated_lj(n: int = 1500, lengths=(6.0, 6.0, 45.0), seed: int = 9) -> W.Workload:
    """LJ beads in a box with two short axes: the small-dimension path."""
    base = W.lj_box(n, 0.8, 1.0, seed=seed)
    rng = np.random.default_rng(seed)
    L = np.asarray(lengths, np.float64)
    # jittered anisotropic lattice with enough sites
    counts = np.maximum(1, np.floor(L / (L.prod() / n) ** (1 / 3))).astype(int)
    while counts.prod() < n:
        counts[np.argmax(L / counts)] += 1
    sites = np.sort(rng.choice(int(counts.prod()), size=n, replace=False))
    g = np.stack([sites % counts[0], (sites // counts[0]) % counts[1],
                  sites // (counts[0] * counts[1])], axis=1).astype(np.float64)
    spacing = L / counts
    pos = (g + 0.5) * spacing + rng.uniform(-0.05, 0.05, (n, 3)) * spacing
    box = np.stack([np.zeros(3), L]).astype(np.float32)
    return base.with_(name="elongated_lj", positions=W._wrap_into(pos, np.zeros(3), L),
                      box_bound=box)


def jittered_fcc(cells: int = 5, amplitude: float = 0.05, seed: int = 1) -> W.Workload:
    """FCC LJ fluid with seeded displacements (a perfect lattice has zero net
    forces, which leaves no scale to measure relative force errors against)."""
    w = W.lj_fluid_fcc(cells=cells, steps=0)
    rng = np.random.default_rng(seed)
    lo = w.box_bound[0].astype(np.float64)
    hi = w.box_bound[1].astype(np.float64)
    pos = w.positions.astype(np.float64) + rng.uniform(-amplitude, amplitude, w.positions.shape)
    return w.with_(name="jittered_fcc", positions=W._wrap_into(pos, lo, hi))


def cases() -> dict[str, W.Workload]:
    return {
        "fcc_direct": jittered_fcc(),
        "lj_regular": W.lj_box(4000, 0.8442, 1.0, seed=3),
        # regular path with 4 cells per axis: the build takes the minimum
        # image per candidate (image shifts per cell need at least 5)
        "lj_regular_4cells": W.lj_box(1200, 0.8442, 1.0, seed=6),
        "lj_small_dim": elongated_lj(),
        "melt": W.polymer_melt(n_chains=80, beads_per_chain=50, seed=3),
        "mixture": W.lj_mixture(n=30000, seed=4),
    }
