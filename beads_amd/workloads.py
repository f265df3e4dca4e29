"""Seeded synthetic workloads for the benchmark configurations and tests.

A ``Workload`` holds every array and parameter of one run in the engine's
dtypes: reals in ``real`` ("float32", or "float64" for the BEADS_REAL64
build's runs; generators make binary32 values, which widen exactly). It can become a ``beads_amd.Simulation`` (the public API), an engine
handle through the C ABI (``beads_amd.capi.Engine``), or an oracle system.

Generators (sizes are arguments so tests can shrink them):
  lj_fluid_fcc   single-type LJ fluid on an FCC lattice, reduced units
  polymer_melt   bead-spring chains, nm_kjmol, overlap-free lattice build
  lj_mixture     multi-type LJ mixture on a jittered lattice, nm_kjmol
  replica_set    many small LJ systems of log-uniform sizes, reduced units
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

KB = {"reduced": 1.0, "nm_kjmol": 0.00831446261815324}


@dataclass
class Workload:
    name: str
    units: str
    positions: np.ndarray
    velocities: np.ndarray
    masses: np.ndarray
    types: np.ndarray
    box_bound: np.ndarray
    lj_epsilon: np.ndarray
    lj_sigma: np.ndarray
    cutoff: float
    shift: bool = False
    tags: np.ndarray | None = None
    molecule_ids: np.ndarray | None = None
    images: np.ndarray | None = None
    bonds: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.uint32))
    angles: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.uint32))
    dihedrals: np.ndarray = field(default_factory=lambda: np.zeros((0, 5), np.uint32))
    bond_k: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    bond_r0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    angle_k: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    angle_theta0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    dihedral_k: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    dihedral_d: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    dihedral_n: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    exclusion_distance: int = 0
    dt: float = 0.005
    berendsen: bool = False
    temperature: float = 1.0
    tau: float = 1.0
    cutoff_buffer: float = 0.3
    rebuild_check_every: int = 1
    sort_every_rebuild: int = 10
    sort_every_steps: int = 0  # nonzero: step-based sort cadence (replaces the rebuild count)
    max_neighbors: int = 512
    half_list: bool = False
    steps: int = 100
    seed: int = 0
    real: str = "float32"

    def __post_init__(self) -> None:
        n = self.positions.shape[0]
        real = np.dtype(self.real)
        if real not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise ValueError("Workload.real must be float32 or float64")
        self.positions = np.ascontiguousarray(self.positions, real)
        self.velocities = np.ascontiguousarray(self.velocities, real)
        self.masses = np.ascontiguousarray(self.masses, real)
        self.types = np.ascontiguousarray(self.types, np.int32)
        self.box_bound = np.ascontiguousarray(self.box_bound, real)
        self.lj_epsilon = np.ascontiguousarray(self.lj_epsilon, real)
        self.lj_sigma = np.ascontiguousarray(self.lj_sigma, real)
        if self.tags is None:
            self.tags = np.arange(1, n + 1, dtype=np.uint32)
        if self.molecule_ids is None:
            self.molecule_ids = np.ones(n, np.uint32)
        if self.images is None:
            self.images = np.zeros((n, 3), np.int32)
        self.tags = np.ascontiguousarray(self.tags, np.uint32)
        self.molecule_ids = np.ascontiguousarray(self.molecule_ids, np.uint32)
        self.images = np.ascontiguousarray(self.images, np.int32)
        for name, dt in (("bonds", np.uint32), ("angles", np.uint32), ("dihedrals", np.uint32),
                         ("bond_k", real), ("bond_r0", real), ("angle_k", real),
                         ("angle_theta0", real), ("dihedral_k", real),
                         ("dihedral_d", np.int32), ("dihedral_n", np.int32)):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dt))

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n_types(self) -> int:
        return int(self.lj_epsilon.shape[0])

    @property
    def kb(self) -> float:
        return KB[self.units]

    def with_(self, **changes) -> "Workload":
        return replace(self, **changes)

    def as_float64(self, jitter: float = 0.0, seed: int = 0) -> "Workload":
        """The same run with binary64 reals; jitter > 0 adds uniform
        [-jitter, jitter) offsets to the positions in binary64 (kept inside the
        box), so that they are no longer binary32 values."""
        pos = self.positions.astype(np.float64)
        if jitter > 0.0:
            rng = np.random.default_rng(seed)
            lo = self.box_bound[0].astype(np.float64)
            hi = self.box_bound[1].astype(np.float64)
            pos = pos + rng.uniform(-jitter, jitter, pos.shape)
            pos = lo + np.mod(pos - lo, hi - lo)
            pos = np.where(pos >= hi, lo, pos)
        return replace(self, real="float64", positions=pos)

    def to_simulation(self):
        """The same run expressed through the public builders."""
        import beads_amd as beads

        system = beads.System(positions=self.positions, velocities=self.velocities,
                              masses=self.masses, types=self.types, tags=self.tags,
                              molecule_ids=self.molecule_ids, images=self.images,
                              box_bound=self.box_bound, units=self.units)
        if len(self.bonds) or len(self.angles) or len(self.dihedrals):
            system.set_topology(bonds=np.asarray(self.bonds, np.int64).reshape(-1, 3),
                                angles=np.asarray(self.angles, np.int64).reshape(-1, 4),
                                dihedrals=np.asarray(self.dihedrals, np.int64).reshape(-1, 5))
        ff = beads.ForceField()
        ff.pair_style("lj", cutoff=float(self.cutoff), shift=bool(self.shift))
        for a in range(self.n_types):
            for b in range(a, self.n_types):
                ff.pair_coeff("lj", type_i=a + 1, type_j=b + 1,
                              epsilon=float(self.lj_epsilon[a, b]), sigma=float(self.lj_sigma[a, b]))
        if len(self.bonds):
            ff.bond_style("harmonic")
            for t in range(len(self.bond_k)):
                ff.bond_coeff("harmonic", type=t + 1, k=float(self.bond_k[t]), r0=float(self.bond_r0[t]))
        if len(self.angles):
            ff.angle_style("harmonic")
            for t in range(len(self.angle_k)):
                ff.angle_coeff("harmonic", type=t + 1, k=float(self.angle_k[t]),
                               theta0=float(self.angle_theta0[t]))
        if len(self.dihedrals):
            ff.dihedral_style("harmonic")
            for t in range(len(self.dihedral_k)):
                ff.dihedral_coeff("harmonic", type=t + 1, k=float(self.dihedral_k[t]),
                                  d=int(self.dihedral_d[t]), n=int(self.dihedral_n[t]))
        if len(self.bonds) and self.exclusion_distance:
            ff.exclude_bonded(distance=int(self.exclusion_distance))
        dyn = beads.Dynamics("velocity_verlet", dt=float(self.dt))
        if self.berendsen:
            dyn.set_thermostat("berendsen", temperature=float(self.temperature), tau=float(self.tau))
        sim = beads.Simulation(system=system, forcefield=ff, dynamics=dyn)
        sim.set_neighbor(cutoff_buffer=float(self.cutoff_buffer),
                         rebuild_check_every=int(self.rebuild_check_every),
                         sort_every_rebuild=int(self.sort_every_rebuild),
                         sort_every_steps=int(self.sort_every_steps) or None,
                         max_neighbors=int(self.max_neighbors), half_list=bool(self.half_list))
        sim.set_runsteps(int(self.steps))
        return sim


def maxwell_boltzmann(rng: np.random.Generator, masses: np.ndarray, temperature: float,
                      kb: float) -> np.ndarray:
    """Gaussian velocities at the given temperature with zero total momentum."""
    sd = np.sqrt(kb * temperature / masses.astype(np.float64))[:, None]
    v = rng.standard_normal((masses.shape[0], 3)) * sd
    p = (masses[:, None] * v).sum(axis=0) / masses.sum()
    return (v - p).astype(np.float32)


def _wrap_into(pos: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Wrap float64 positions into [lo, hi) and make sure the float32 values are too."""
    L = hi - lo
    p = lo + np.mod(pos - lo, L)
    p32 = p.astype(np.float32)
    lo32, hi32 = lo.astype(np.float32), hi.astype(np.float32)
    p32 = np.where(p32 >= hi32, lo32, p32)
    p32 = np.where(p32 < lo32, lo32, p32)
    return p32


def lj_fluid_fcc(cells: int = 10, density: float = 0.8442, temperature: float = 1.0,
                 seed: int = 2024, steps: int = 1000) -> Workload:
    """Single-type LJ fluid: 4 * cells^3 beads on an FCC lattice, reduced units,
    eps = sigma = 1, r_c = 2.5 (shifted), skin 0.3, check every 10, dt 0.005."""
    a = (4.0 / density) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], np.float64)
    g = np.indices((cells, cells, cells)).reshape(3, -1).T.astype(np.float64)
    pos = ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3) + 0.25) * a
    L = cells * a
    lo, hi = np.zeros(3), np.full(3, L)
    n = pos.shape[0]
    rng = np.random.default_rng(seed)
    masses = np.ones(n, np.float32)
    return Workload(
        name=f"lj_fluid_fcc_{n}", units="reduced", positions=_wrap_into(pos, lo, hi),
        velocities=maxwell_boltzmann(rng, masses, temperature, 1.0), masses=masses,
        types=np.ones(n, np.int32), box_bound=np.stack([lo, hi]).astype(np.float32),
        lj_epsilon=np.ones((1, 1)), lj_sigma=np.ones((1, 1)), cutoff=2.5, shift=True,
        dt=0.005, cutoff_buffer=0.3, rebuild_check_every=10, sort_every_rebuild=10,
        max_neighbors=512, steps=steps, seed=seed)


def _serpentine(n_side: int) -> np.ndarray:
    """Lattice coordinates of a Hamiltonian path through an n^3 cubic lattice
    (boustrophedon in x inside y rows inside z layers): consecutive entries are
    nearest neighbours."""
    idx = np.arange(n_side ** 3)
    z = idx // (n_side * n_side)
    rem = idx % (n_side * n_side)
    y_raw = rem // n_side
    x_raw = rem % n_side
    y = np.where(z % 2 == 0, y_raw, n_side - 1 - y_raw)
    row = z * n_side + y_raw
    x = np.where(row % 2 == 0, x_raw, n_side - 1 - x_raw)
    return np.stack([x, y, z], axis=1)


def polymer_melt(n_chains: int = 20000, beads_per_chain: int = 50, density: float = 8.5,
                 bond_length: float = 0.47, mass: float = 72.0, temperature: float = 300.0,
                 seed: int = 7, steps: int = 500, jitter: float = 0.02) -> Workload:
    """Bead-spring melt in nm_kjmol: harmonic bonds (k 1250, r0 0.47), angles
    (k 25, theta0 120 deg), dihedrals (k 1, d 1, n 1), LJ eps 3.5 sigma 0.47 r_c 1.2,
    exclusions to distance 2, Berendsen 300 K tau 1 ps, dt 0.01 ps.

    Overlap-free build: chains are consecutive runs of a serpentine path through a
    cubic lattice whose spacing is the box length over floor(L / bond_length) sites;
    the lattice vacancies are spread as seeded random gaps between chains, then every
    bead is jittered by up to +-jitter nm."""
    n = n_chains * beads_per_chain
    L = (n / density) ** (1.0 / 3.0)
    side = int(math.floor(L / bond_length))
    while side ** 3 < n:
        side += 1
    spacing = L / side
    path = _serpentine(side)
    rng = np.random.default_rng(seed)
    vacancies = side ** 3 - n
    gaps = rng.multinomial(vacancies, np.full(n_chains + 1, 1.0 / (n_chains + 1)))
    starts = np.concatenate([[gaps[0]], gaps[0] + np.cumsum(gaps[1:n_chains] + beads_per_chain)])
    site = (starts[:, None] + np.arange(beads_per_chain)[None, :]).reshape(-1)
    pos = (path[site].astype(np.float64) + 0.5) * spacing
    pos += rng.uniform(-jitter, jitter, pos.shape)
    lo, hi = np.zeros(3), np.full(3, L)
    masses = np.full(n, mass, np.float32)
    tags = np.arange(1, n + 1, dtype=np.uint32).reshape(n_chains, beads_per_chain)
    b = np.stack([tags[:, :-1], tags[:, 1:], np.ones_like(tags[:, 1:])], axis=-1).reshape(-1, 3)
    a = np.stack([tags[:, :-2], tags[:, 1:-1], tags[:, 2:], np.ones_like(tags[:, 2:])],
                 axis=-1).reshape(-1, 4)
    d = np.stack([tags[:, :-3], tags[:, 1:-2], tags[:, 2:-1], tags[:, 3:], np.ones_like(tags[:, 3:])],
                 axis=-1).reshape(-1, 5)
    mol = np.repeat(np.arange(1, n_chains + 1, dtype=np.uint32), beads_per_chain)
    return Workload(
        name=f"polymer_melt_{n}", units="nm_kjmol", positions=_wrap_into(pos, lo, hi),
        velocities=maxwell_boltzmann(rng, masses, temperature, KB["nm_kjmol"]), masses=masses,
        types=np.ones(n, np.int32), box_bound=np.stack([lo, hi]).astype(np.float32),
        lj_epsilon=np.full((1, 1), 3.5), lj_sigma=np.full((1, 1), 0.47), cutoff=1.2, shift=False,
        molecule_ids=mol, bonds=b, angles=a, dihedrals=d,
        bond_k=np.array([1250.0]), bond_r0=np.array([0.47]), angle_k=np.array([25.0]),
        angle_theta0=np.array([120.0]), dihedral_k=np.array([1.0]), dihedral_d=np.array([1]),
        dihedral_n=np.array([1]), exclusion_distance=2, dt=0.01, berendsen=True,
        temperature=temperature, tau=1.0, cutoff_buffer=0.3, rebuild_check_every=1,
        sort_every_rebuild=10, max_neighbors=512, steps=steps, seed=seed)


def polymer_solvent(n_chains: int = 500, beads_per_chain: int = 100, n_solvent: int = 50_000,
                    density: float = 8.5, bond_length: float = 0.47, mass: float = 72.0,
                    temperature: float = 300.0, seed: int = 21, jitter: float = 0.02,
                    shift: bool = True) -> Workload:
    """Four-type polymer/solvent system in nm_kjmol for force matching: chain
    beads alternate types 1/2, solvent beads are types 3/4 at random; harmonic
    bonds only (k 1250, r0 0.47), exclusions to distance 2, LJ r_c 1.2 shifted,
    eps_ab ~ U[2,4] kJ/mol and sigma_ab ~ U[0.43,0.51] nm symmetric, dt 0.01 ps,
    no thermostat. Chains are consecutive runs of the serpentine lattice path
    separated by seeded gaps; solvent takes seeded random free sites; every bead
    is jittered by up to +-jitter nm."""
    n_chain_beads = n_chains * beads_per_chain
    n = n_chain_beads + n_solvent
    L = (n / density) ** (1.0 / 3.0)
    side = int(math.floor(L / bond_length))
    while side ** 3 < n:
        side += 1
    spacing = L / side
    path = _serpentine(side)
    rng = np.random.default_rng(seed)
    free = side ** 3 - n_chain_beads
    gaps = rng.multinomial(free, np.full(n_chains + 1, 1.0 / (n_chains + 1)))
    starts = np.concatenate([[gaps[0]], gaps[0] + np.cumsum(gaps[1:n_chains] + beads_per_chain)])
    chain_sites = (starts[:, None] + np.arange(beads_per_chain)[None, :]).reshape(-1)
    used = np.zeros(side ** 3, bool)
    used[chain_sites] = True
    solvent_sites = np.sort(rng.choice(np.flatnonzero(~used), size=n_solvent, replace=False))
    site = np.concatenate([chain_sites, solvent_sites])
    pos = (path[site].astype(np.float64) + 0.5) * spacing
    pos += rng.uniform(-jitter, jitter, pos.shape)
    lo, hi = np.zeros(3), np.full(3, L)
    masses = np.full(n, mass, np.float32)
    types = np.concatenate([np.tile(np.array([1, 2], np.int32), n_chain_beads // 2 + 1)[:n_chain_beads],
                            rng.integers(3, 5, n_solvent).astype(np.int32)])
    eps = rng.uniform(2.0, 4.0, (4, 4))
    sig = rng.uniform(0.43, 0.51, (4, 4))
    eps = np.triu(eps) + np.triu(eps, 1).T
    sig = np.triu(sig) + np.triu(sig, 1).T
    tags = np.arange(1, n_chain_beads + 1, dtype=np.uint32).reshape(n_chains, beads_per_chain)
    b = np.stack([tags[:, :-1], tags[:, 1:], np.ones_like(tags[:, 1:])], axis=-1).reshape(-1, 3)
    mol = np.concatenate([np.repeat(np.arange(1, n_chains + 1, dtype=np.uint32), beads_per_chain),
                          np.arange(n_chains + 1, n_chains + 1 + n_solvent, dtype=np.uint32)])
    return Workload(
        name=f"polymer_solvent_{n}", units="nm_kjmol", positions=_wrap_into(pos, lo, hi),
        velocities=maxwell_boltzmann(rng, masses, temperature, KB["nm_kjmol"]), masses=masses,
        types=types, box_bound=np.stack([lo, hi]).astype(np.float32), lj_epsilon=eps,
        lj_sigma=sig, cutoff=1.2, shift=shift, molecule_ids=mol, bonds=b,
        bond_k=np.array([1250.0]), bond_r0=np.array([0.47]), exclusion_distance=2, dt=0.01,
        cutoff_buffer=0.3, rebuild_check_every=1, sort_every_rebuild=10, max_neighbors=512,
        steps=200, seed=seed, temperature=temperature)


def lj_mixture(n: int = 4_000_000, n_types: int = 4, density: float = 8.5, seed: int = 11,
               steps: int = 200, temperature: float = 300.0, mass: float = 72.0) -> Workload:
    """Multi-type LJ mixture in nm_kjmol: types uniform at random, eps_ab ~ U[1,4]
    kJ/mol, sigma_ab ~ U[0.43,0.53] nm, r_c 1.1, skin 0.1, max_neighbors 128,
    spatial sort every 100 steps (the step-based cadence), dt 0.01 ps.
    Positions: n distinct sites drawn from the smallest cubic lattice with at
    least n sites, jittered by a tenth of the spacing."""
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    side = int(math.ceil(n ** (1.0 / 3.0)))
    spacing = L / side
    sites = rng.choice(side ** 3, size=n, replace=False)
    sites.sort()
    g = np.stack([sites % side, (sites // side) % side, sites // (side * side)], axis=1)
    pos = (g.astype(np.float64) + 0.5) * spacing + rng.uniform(-0.1, 0.1, (n, 3)) * spacing
    lo, hi = np.zeros(3), np.full(3, L)
    eps = rng.uniform(1.0, 4.0, (n_types, n_types))
    sig = rng.uniform(0.43, 0.53, (n_types, n_types))
    eps = np.triu(eps) + np.triu(eps, 1).T
    sig = np.triu(sig) + np.triu(sig, 1).T
    masses = np.full(n, mass, np.float32)
    return Workload(
        name=f"lj_mixture_{n}", units="nm_kjmol", positions=_wrap_into(pos, lo, hi),
        velocities=maxwell_boltzmann(rng, masses, temperature, KB["nm_kjmol"]), masses=masses,
        types=rng.integers(1, n_types + 1, n).astype(np.int32),
        box_bound=np.stack([lo, hi]).astype(np.float32), lj_epsilon=eps, lj_sigma=sig,
        cutoff=1.1, shift=False, dt=0.01, cutoff_buffer=0.1, rebuild_check_every=1,
        sort_every_rebuild=10, sort_every_steps=100, max_neighbors=128, steps=steps, seed=seed)


def lj_box(n: int, density: float, temperature: float, seed: int, cutoff: float = 2.5,
           skin: float = 0.3, steps: int = 100, dt: float = 0.005, berendsen: bool = False,
           tau: float = 0.5, check_every: int = 1) -> Workload:
    """n LJ beads (reduced units) on a jittered simple-cubic lattice at the given
    density, Maxwell-Boltzmann velocities."""
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    side = int(math.ceil(n ** (1.0 / 3.0)))
    spacing = L / side
    sites = np.sort(rng.choice(side ** 3, size=n, replace=False))
    g = np.stack([sites % side, (sites // side) % side, sites // (side * side)], axis=1)
    pos = (g.astype(np.float64) + 0.5) * spacing + rng.uniform(-0.05, 0.05, (n, 3)) * spacing
    lo, hi = np.zeros(3), np.full(3, L)
    masses = np.ones(n, np.float32)
    return Workload(
        name=f"lj_box_{n}", units="reduced", positions=_wrap_into(pos, lo, hi),
        velocities=maxwell_boltzmann(rng, masses, temperature, 1.0), masses=masses,
        types=np.ones(n, np.int32), box_bound=np.stack([lo, hi]).astype(np.float32),
        lj_epsilon=np.ones((1, 1)), lj_sigma=np.ones((1, 1)), cutoff=cutoff, shift=True, dt=dt,
        berendsen=berendsen, temperature=temperature, tau=tau, cutoff_buffer=skin,
        rebuild_check_every=check_every, sort_every_rebuild=10, max_neighbors=512, steps=steps,
        seed=seed)


def melt_replica_set(count: int = 128, side_min: int = 10, side_max: int = 27, seed: int = 9,
                     steps: int = 1000) -> list[Workload]:
    """Independent bead-spring melts of polymer_melt's force field and uneven
    size: lattice side s log-uniform in [side_min, side_max] (s^3 sites),
    chains of 20, 30, 40 or 50 beads filling 95 % of the sites, box 0.48 s nm
    (spacing 0.48 nm, density about 8.6 nm^-3), 300 K."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        side = int(round(math.exp(rng.uniform(math.log(side_min), math.log(side_max)))))
        per_chain = int(rng.choice([20, 30, 40, 50]))
        chains = max(1, int(0.95 * side ** 3) // per_chain)
        n = chains * per_chain
        out.append(polymer_melt(n_chains=chains, beads_per_chain=per_chain,
                                density=n / (0.48 * side) ** 3, seed=seed * 100003 + k,
                                steps=steps))
    return out


def replica_set(count: int = 512, n_min: int = 500, n_max: int = 20000, seed: int = 5,
                steps: int = 1000) -> list[Workload]:
    """Independent reduced-unit LJ systems: N log-uniform in [n_min, n_max],
    density U[0.6, 0.9], T U[0.8, 1.5] with Berendsen tau 0.5, r_c 2.5, dt 0.005."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), count))).astype(int)
    dens = rng.uniform(0.6, 0.9, count)
    temps = rng.uniform(0.8, 1.5, count)
    return [lj_box(int(sizes[k]), float(dens[k]), float(temps[k]), seed=seed * 100003 + k,
                   steps=steps, berendsen=True, tau=0.5) for k in range(count)]
