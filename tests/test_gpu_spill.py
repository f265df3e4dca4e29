"""Particles with more records than the inline capacities, on the GPU.

Listed-term roles and exclusions are stored per slot up to an inline
capacity per kind; a particle with more (a hub) keeps the rest in a spill
segment, which moves with it through every reorder. The reference has no
per-particle limit either: exclusions past 6 inline partners go to an
overflow region (exclusions.cu:131-210, 337-445) and bonded terms are term
lists. Two hubs at the centre of a 3,375-bead LJ lattice:
  * 300 one-bead arms, exclusions to distance 2: the hub has 300 bonds, and
    the hub and every arm bead 300 exclusions (all past the capacities);
  * 30 two-bead arms with every angle and dihedral the arms define,
    exclusions to distance 3: the hub takes part in 30 bonds, 465 angles,
    870 dihedrals and 60 exclusions.
Bars: neighbour counts and sets bit-exact against the FP64 oracle, forces
within 1e-5 of FP64 per component, energies relative 1e-5; at the initial
build and after 30 steps at zero skin that rebuild and reorder after every
step. The adjoint reads spilled bonds too: dE/dx = -F and dE/dk, dE/dr0 in
closed form.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from beads_amd.adjoint import FrameModel
from beads_amd.capi import Engine
from gpu_cases import hub_lattice

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5
ENERGY_RTOL = 1e-5

CASES = {"hub300": (300, 1, 2, False), "hub30x2": (30, 2, 3, True)}


def _check(w, e):
    st = e.state()
    stats = e.stats()
    assert stats.overflow == 0
    counts, lst = e.neighbors()
    rc, rl, _ = O.neighbor_lists(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        assert np.array_equal(np.sort(lst[: counts[i], i]), np.sort(rl[i, : rc[i]])), i
    f_ref, obs = O.forces(w, positions=st["positions"], tags=st["tags"], types=st["types"])
    scale = np.sqrt(np.mean(f_ref ** 2))
    err = np.max(np.abs(st["forces"] - f_ref) / np.maximum(np.abs(f_ref), scale))
    assert err <= FORCE_RTOL, err
    for q, term in enumerate(("pair", "bond", "angle", "dihedral")):
        ref = obs[f"{term}_pe"]
        assert abs(stats.pe[q] - ref) <= ENERGY_RTOL * max(1.0, abs(ref)), (term, stats.pe[q], ref)


@pytest.mark.parametrize("name", list(CASES))
def test_hub_records_past_inline_capacity(name):
    w = hub_lattice(*CASES[name])
    with Engine(w) as e:
        e.initialize(observables=True)
        _check(w, e)
    w0 = w.with_(cutoff_buffer=0.0, sort_every_rebuild=1)
    with Engine(w0) as e:
        e.initialize(observables=False)
        e.advance(30, observables_on_last=True)
        st = e.stats()
        assert st.build_count == 31 and st.reorder_count >= 30
        _check(w0, e)


def test_adjoint_reads_spilled_bonds():
    w = hub_lattice(300, 1, 2, False)
    x = torch.tensor(w.positions, device="cuda", requires_grad=True)
    with FrameModel(w) as m:
        p = m.initial_params()
        f, energy = m.forces(x, p)
        energy.backward()
        m.check()
        gx = x.grad.cpu().numpy().astype(np.float64)
        g_k = float(p.bond_k.grad[0])
        g_r0 = float(p.bond_r0.grad[0])
        f = f.detach().cpu().numpy().astype(np.float64)
    with Engine(w) as e:
        e.initialize(observables=False)
        np.testing.assert_array_equal(e.state()["forces"], f.astype(np.float32))
    scale = np.sqrt(np.mean(f ** 2))
    assert np.max(np.abs(-gx - f) / np.maximum(np.abs(f), scale)) <= FORCE_RTOL
    # E_bond = sum k/2 (r - r0)^2: dE/dk = sum (r - r0)^2 / 2, dE/dr0 = -k sum (r - r0)
    pos = w.positions.astype(np.float64)
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    d = pos[w.bonds[:, 0].astype(int) - 1] - pos[w.bonds[:, 1].astype(int) - 1]
    d -= np.round(d / L) * L
    r = np.linalg.norm(d, axis=1)
    r0 = float(w.bond_r0[0])
    k = float(w.bond_k[0])
    assert g_k == pytest.approx(0.5 * np.sum((r - r0) ** 2), rel=1e-5)
    assert g_r0 == pytest.approx(-k * np.sum(r - r0), rel=1e-5)
