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


def example_chain(distance: int, beads: int = 8) -> W.Workload:
    """The reference's bonded-exclusion example (examples/05_bonded_exclusion_switch.py
    with examples/_systems.py:123-183): one zig-zag chain of 8 beads, bond
    length 0.35 nm, LJ eps 0.5 sigma 0.34 r_c 1.2, harmonic bonds k 0 r0 0.35,
    nm_kjmol, exclusions to the given bond distance. Every axis has fewer than
    three cells, so the engine takes the direct path."""
    pos = np.array([[1.0 + b * 0.35, 1.0 + (0.12 if b % 2 == 0 else -0.12),
                     1.0 + 0.08 * float((b % 3) - 1)] for b in range(beads)])
    box = np.array([[0.0, 0.0, 0.0], [2.0 + 0.35 * (beads - 1), 4.0, 4.0]])
    tags = np.arange(1, beads + 1, dtype=np.uint32)
    bonds = np.stack([tags[:-1], tags[1:], np.ones(beads - 1, np.uint32)], axis=1)
    rng = np.random.default_rng(5)
    return W.Workload(
        name=f"example_chain_d{distance}", units="nm_kjmol", positions=pos,
        velocities=rng.uniform(-0.01, 0.01, (beads, 3)), masses=np.ones(beads),
        types=np.ones(beads, np.int32), box_bound=box, lj_epsilon=[[0.5]], lj_sigma=[[0.34]],
        cutoff=1.2, bonds=bonds, bond_k=[0.0], bond_r0=[0.35], exclusion_distance=distance,
        dt=0.001, steps=0)


def bonded_lattice(kind: str, distance: int, side: int = 15, density: float = 0.8442,
                   seed: int = 31) -> W.Workload:
    """LJ beads (reduced units) on a jittered full cubic lattice of side^3 sites,
    with harmonic bonds (k 100, r0 = lattice spacing) between lattice
    neighbours, exclusions to the given bond distance:
      chains  serpentine runs of `side` beads: 2 d partners per interior bead
              (distance 3: six, the engine's 8-entry register variant);
      stars   a hub every 5 sites per axis with six arms of two beads along
              the lattice axes: the hub has 12 partners at distance 2 (the
              engine's variant that reads exclusions from memory)."""
    n = side ** 3
    L = (n / density) ** (1.0 / 3.0)
    spacing = L / side
    rng = np.random.default_rng(seed)
    if kind == "chains":
        g = W._serpentine(side).astype(np.float64)
        tags = np.arange(1, n + 1, dtype=np.uint32).reshape(-1, side)
        bonds = np.stack([tags[:, :-1], tags[:, 1:], np.ones_like(tags[:, 1:])],
                         axis=-1).reshape(-1, 3)
    elif kind == "stars":
        assert side % 5 == 0
        g = np.indices((side, side, side)).reshape(3, -1).T[:, ::-1].astype(np.float64)
        # site index s = x + side (y + side z); tag = s + 1
        def tag(x, y, z):
            return int((x % side) + side * ((y % side) + side * (z % side))) + 1
        bl = []
        for hx in range(2, side, 5):
            for hy in range(2, side, 5):
                for hz in range(2, side, 5):
                    hub = tag(hx, hy, hz)
                    for ax in range(3):
                        for sgn in (-1, 1):
                            d1 = [0, 0, 0]
                            d1[ax] = sgn
                            a1 = tag(hx + d1[0], hy + d1[1], hz + d1[2])
                            a2 = tag(hx + 2 * d1[0], hy + 2 * d1[1], hz + 2 * d1[2])
                            bl.append((min(hub, a1), max(hub, a1), 1))
                            bl.append((min(a1, a2), max(a1, a2), 1))
        bonds = np.array(sorted(bl), np.uint32)
    else:
        raise ValueError(kind)
    pos = (g + 0.5) * spacing + rng.uniform(-0.05, 0.05, (n, 3)) * spacing
    lo, hi = np.zeros(3), np.full(3, L)
    masses = np.ones(n)
    return W.Workload(
        name=f"bonded_{kind}_d{distance}", units="reduced", positions=W._wrap_into(pos, lo, hi),
        velocities=W.maxwell_boltzmann(rng, masses.astype(np.float32), 1.0, 1.0), masses=masses,
        types=np.ones(n, np.int32), box_bound=np.stack([lo, hi]), lj_epsilon=np.ones((1, 1)),
        lj_sigma=np.ones((1, 1)), cutoff=2.5, shift=True, bonds=bonds, bond_k=[100.0],
        bond_r0=[spacing], exclusion_distance=distance, dt=0.005, steps=100)


def hub_lattice(arms: int, arm_length: int, distance: int, listed: bool, side: int = 15,
                density: float = 0.8442, seed: int = 37) -> W.Workload:
    """LJ beads (reduced units) on a jittered full cubic lattice with one hub
    at the centre site carrying `arms` arms of `arm_length` beads each (the
    sites nearest the hub, bonded hub-a1-a2...), harmonic bonds (k 100, r0 =
    spacing) and, when `listed`, every angle and dihedral the arms define
    (a1-hub-b1 for every pair of arms, hub-a1-a2, a2-a1-hub-b1 ...), with
    exclusions to the given distance. The hub takes part in far more bonds,
    angles, dihedrals and exclusions than the other beads, so its records go
    past the engine's inline capacities into spill segments."""
    n = side ** 3
    L = (n / density) ** (1.0 / 3.0)
    spacing = L / side
    rng = np.random.default_rng(seed)
    g = np.indices((side, side, side)).reshape(3, -1).T[:, ::-1].astype(np.float64)
    pos = (g + 0.5) * spacing + rng.uniform(-0.05, 0.05, (n, 3)) * spacing
    c = side // 2
    hub = c + side * (c + side * c)
    d = np.linalg.norm((g - g[hub]), axis=1)
    order = [int(k) for k in np.argsort(d, kind="stable") if k != hub]
    tags = np.arange(1, n + 1)
    chains = []
    for a in range(arms):
        chains.append([hub] + [order[a + arms * m] for m in range(arm_length)])
    bonds, angles, dihedrals = set(), set(), set()
    for ch in chains:
        for u, v in zip(ch[:-1], ch[1:]):
            bonds.add((min(tags[u], tags[v]), max(tags[u], tags[v])))
    if listed:
        for a, ca in enumerate(chains):
            for b, cb in enumerate(chains):
                if a < b:
                    angles.add((tags[ca[1]], tags[hub], tags[cb[1]]))
                if a != b and arm_length >= 2:
                    dihedrals.add((tags[ca[2]], tags[ca[1]], tags[hub], tags[cb[1]]))
            for k in range(len(ca) - 2):
                angles.add((tags[ca[k]], tags[ca[k + 1]], tags[ca[k + 2]]))
    b = np.array([(i, j, 1) for i, j in sorted(bonds)], np.uint32)
    an = np.array([(i, j, k, 1) for i, j, k in sorted(angles)], np.uint32).reshape(-1, 4)
    dh = np.array([(i, j, k, m, 1) for i, j, k, m in sorted(dihedrals)], np.uint32).reshape(-1, 5)
    lo, hi = np.zeros(3), np.full(3, L)
    masses = np.ones(n)
    extra = {}
    if listed:
        extra = dict(angles=an, dihedrals=dh, angle_k=[5.0], angle_theta0=[109.5],
                     dihedral_k=[0.5], dihedral_d=[1], dihedral_n=[2])
    return W.Workload(
        name=f"hub_{arms}x{arm_length}_d{distance}", units="reduced",
        positions=W._wrap_into(pos, lo, hi),
        velocities=W.maxwell_boltzmann(rng, masses.astype(np.float32), 1.0, 1.0), masses=masses,
        types=np.ones(n, np.int32), box_bound=np.stack([lo, hi]), lj_epsilon=np.ones((1, 1)),
        lj_sigma=np.ones((1, 1)), cutoff=2.5, shift=True, bonds=b, bond_k=[100.0],
        bond_r0=[spacing], exclusion_distance=distance, dt=0.002, steps=50, **extra)


def exclusion_counts(w: W.Workload) -> np.ndarray:
    """Per-tag number of beads within the exclusion bond distance (brute-force
    breadth-first search over the bond graph)."""
    adj = [[] for _ in range(w.n + 1)]
    for a, b, _ in np.asarray(w.bonds, np.int64):
        adj[a].append(b)
        adj[b].append(a)
    out = np.zeros(w.n + 1, np.int64)
    for s in range(1, w.n + 1):
        seen, frontier = {s}, [s]
        for _ in range(w.exclusion_distance):
            frontier = [q for t in frontier for q in adj[t] if q not in seen]
            seen.update(frontier)
        out[s] = len(seen) - 1
    return out[1:]


def drifting_lattice(v_dt: float, check_every: int = 1, steps: int = 60, skin: float = 0.3,
                     spacing: float = 5.0, side: int = 4) -> W.Workload:
    """side^3 LJ beads (reduced units, r_c 2.5) on a simple-cubic lattice whose
    spacing exceeds r_c plus the largest drift, so no pair is ever inside r_c
    and every force is zero. All beads rest except bead 1, which drifts along
    +x by v_dt per step: its displacement from the last build is k * v_dt after
    k steps, so the rebuild steps follow in closed form (expected_builds)."""
    n = side ** 3
    g = np.indices((side, side, side)).reshape(3, -1).T.astype(np.float64)
    L = side * spacing
    pos = (g + 0.5) * spacing
    vel = np.zeros((n, 3))
    dt = 0.005
    vel[0, 0] = v_dt / dt
    lo, hi = np.zeros(3), np.full(3, L)
    return W.Workload(
        name=f"drifting_lattice_{n}", units="reduced", positions=W._wrap_into(pos, lo, hi),
        velocities=vel, masses=np.ones(n), types=np.ones(n, np.int32),
        box_bound=np.stack([lo, hi]), lj_epsilon=np.ones((1, 1)), lj_sigma=np.ones((1, 1)),
        cutoff=2.5, shift=True, dt=dt, cutoff_buffer=skin, rebuild_check_every=check_every,
        sort_every_rebuild=10, max_neighbors=64, steps=steps)


def expected_build_steps(steps_to_cross: float, check_every: int, steps: int,
                         skin: float) -> tuple[list[int], int]:
    """(steps at which a build happens, 0 being the initial build; dangerous
    rebuild count) when the largest displacement since the last build is
    k / steps_to_cross times half the skin after k steps. The rule of
    neighbor_system.cpp:79-105, 198-216: skin 0 rebuilds after every step;
    otherwise the displacement is checked when step % check_every == 0, a
    rebuild follows when it exceeds half the skin (strictly), and a rebuild at
    the first check after a build counts as dangerous when check_every > 1."""
    if skin == 0.0:
        return list(range(steps + 1)), 0
    builds, dangerous, last, seen = [0], 0, 0, False
    for step in range(1, steps + 1):
        if step % check_every:
            continue
        first, seen = not seen, True
        if step - last > steps_to_cross:
            dangerous += int(check_every > 1 and first)
            builds.append(step)
            last, seen = step, False
    return builds, dangerous


def expected_builds(steps_to_cross: float, check_every: int, steps: int,
                    skin: float) -> tuple[int, int]:
    """(build_count, dangerous_rebuild_count) of expected_build_steps; the
    initial build counts once."""
    builds, dangerous = expected_build_steps(steps_to_cross, check_every, steps, skin)
    return len(builds), dangerous


def expected_reorders(build_steps: list[int], sort_every_rebuild: int,
                      sort_every_steps: int = 0) -> int:
    """Reorders of a run whose builds happen at build_steps: every
    sort_every_rebuild-th build (neighbor_system.cpp:198-202), or, with
    sort_every_steps, the first build at least that many steps after the last
    sort (the start of the run counting as a sort at step 0)."""
    count, last = 0, 0
    for k, step in enumerate(build_steps, start=1):
        due = (step - last >= sort_every_steps) if sort_every_steps else k % sort_every_rebuild == 0
        if due:
            count, last = count + 1, step
    return count


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
