"""beads_engine_frame_backward_noalloc against beads_engine_frame_backward.

The allocation-free variant stages dL/dF in the library's static per-GPU
scratch array and clears the parameter gradients with a kernel; it runs the
same adjoint kernel on the same inputs. Bars: dL/dx bit-identical (each atom's
value is a fixed-order FP64 team sum rounded once); FP64 parameter gradients
within 1e-12 relative (they are summed with per-workgroup FP64 atomics, whose
order varies). Two calls in a row, the second on a smaller system, so the
second reuses the scratch the first left behind; a system above the scratch's
capacity is rejected before anything is launched.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from beads_amd import workloads as W
from beads_amd.capi import BeadsError, Engine

pytestmark = pytest.mark.gpu

MAX_ATOMS = 1 << 20


def _system(n_chains, n_solvent, seed):
    return W.polymer_solvent(n_chains=n_chains, beads_per_chain=20, n_solvent=n_solvent,
                             seed=seed, jitter=0.03)


def _backward_pair(w, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.tensor(w.positions, device="cuda")
    g = torch.randn((w.n, 3), generator=gen, device="cuda")
    n_p = 2 * w.n_types * w.n_types + 2 * int(w.bond_k.shape[0])
    with Engine(w) as e:
        f = torch.empty_like(x)
        e.frame_forces(x.data_ptr(), f.data_ptr(), reuse_list=False)
        gx_ref = torch.empty_like(x)
        gp_ref = torch.empty(n_p, dtype=torch.float64, device="cuda")
        e.frame_backward(g.data_ptr(), 0.25, gx_ref.data_ptr(), gp_ref.data_ptr())
        # poison the outputs: the variant must overwrite every entry
        gx = torch.full_like(x, float("nan"))
        gp = torch.full((n_p,), float("nan"), dtype=torch.float64, device="cuda")
        e.frame_backward_noalloc(g.data_ptr(), 0.25, gx.data_ptr(), gp.data_ptr())
        e.synchronize()
    return (gx.cpu().numpy(), gp.cpu().numpy()), (gx_ref.cpu().numpy(), gp_ref.cpu().numpy())


def test_noalloc_backward_matches_allocating_backward_twice():
    for w, seed in ((_system(40, 3000, 3), 1), (_system(10, 600, 5), 2)):
        (gx, gp), (gx_ref, gp_ref) = _backward_pair(w, seed)
        np.testing.assert_array_equal(gx, gx_ref)
        assert np.all(np.isfinite(gp))
        np.testing.assert_allclose(gp, gp_ref, rtol=1e-12, atol=1e-12 * np.abs(gp_ref).max())


def test_noalloc_backward_rejects_systems_above_scratch_capacity():
    n = MAX_ATOMS + 1
    rng = np.random.default_rng(0)
    L = (n / 0.05) ** (1.0 / 3.0)  # dilute: the list stays short
    w = W.Workload(name="dilute", units="reduced", positions=rng.uniform(0, L, (n, 3)),
                   velocities=np.zeros((n, 3)), masses=np.ones(n), types=np.ones(n, np.int32),
                   box_bound=np.array([[0, 0, 0], [L, L, L]]), lj_epsilon=[[1.0]],
                   lj_sigma=[[1.0]], cutoff=2.5, max_neighbors=64)
    x = torch.tensor(w.positions, device="cuda")
    with Engine(w) as e:
        f = torch.empty_like(x)
        e.frame_forces(x.data_ptr(), f.data_ptr(), reuse_list=False)
        g = torch.zeros_like(x)
        gx = torch.empty_like(x)
        gp = torch.empty(2, dtype=torch.float64, device="cuda")
        with pytest.raises(BeadsError, match="at most 1048576 particles"):
            e.frame_backward_noalloc(g.data_ptr(), 0.0, gx.data_ptr(), gp.data_ptr())
