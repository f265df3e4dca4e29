"""GPU parity of the double-precision engine build (BEADS_REAL64,
libbeads_amd_real64.so) against the oracle's binary64 entry points.

Inputs are binary64: every workload is widened and its positions jittered by
up to 1e-4 in binary64, so that no position is a binary32 value (a build that
stored binary32 would miss every bar below by seven orders of magnitude). Box
bounds and force-field parameters, the cutoff and the skin included, are
binary32 values (exact in binary64), which is what the oracle takes: with
r_c = 1.2 given in binary64 the oracle's binary32 r_c is 5e-8 larger, and a
pair that passes through that band during a run is counted by one side only.

Bars (stated once, used below):
  * cell assignment, in-cell order and neighbour sets: bit-exact against the
    binary64 decisions of the oracle (cell index, r^2 <= r_s^2);
  * forces: |f_gpu - f_ref| <= FORCE_RTOL * max(|f_ref|, f_rms) per component;
  * energies, virials, kinetic energy: relative ENERGY_RTOL;
  * one velocity-Verlet step: positions to STEP_ATOL (absolute, in the box's
    length unit), velocities to FORCE_RTOL of their rms;
  * a whole run against the oracle's binary64 run (same step order, own cell
    order): positions and velocities to RUN_RTOL of their scale after the
    stated steps (both sides round differently only in summation order);
  * adjoint gradients against central differences of the oracle in binary64:
    GRAD_RTOL; the trajectory adjoint against the oracle's tangent run by the
    dot-product identity: DOT_RTOL.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from beads_amd import workloads as W
from beads_amd.capi import Batch, Engine
from gpu_cases import cases

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-12
ENERGY_RTOL = 1e-12
STEP_ATOL = 1e-13
RUN_RTOL = 1e-10
GRAD_RTOL = 1e-7
DOT_RTOL = 1e-9

def _binary64_case(w: W.Workload, seed: int) -> W.Workload:
    return w.as_float64(jitter=1e-4, seed=seed).with_(
        cutoff=float(np.float32(w.cutoff)), cutoff_buffer=float(np.float32(w.cutoff_buffer)))


CASES = {name: _binary64_case(w, k) for k, (name, w) in enumerate(cases().items())}


def _engine(w: W.Workload) -> Engine:
    e = Engine(w, precision="double")
    e.initialize(observables=True)
    return e


def _force_error(f_gpu: np.ndarray, f_ref: np.ndarray) -> float:
    scale = np.sqrt(np.mean(f_ref ** 2))
    return float(np.max(np.abs(f_gpu - f_ref) / np.maximum(np.abs(f_ref), scale)))


def test_double_engine_reports_binary64_storage():
    w = CASES["lj_regular"]
    assert w.positions.dtype == np.float64
    assert (w.positions != w.positions.astype(np.float32)).mean() > 0.99
    with _engine(w) as e:
        st = e.state()
    assert st["positions"].dtype == np.float64 and st["forces"].dtype == np.float64
    # the device keeps the binary64 positions as given (no step has run)
    order = np.argsort(st["tags"])
    np.testing.assert_array_equal(st["positions"][order], w.positions)


@pytest.mark.parametrize("name", list(CASES))
def test_cells_and_neighbors_bit_exact(name):
    w = CASES[name]
    plan = O.plan_f64(w)
    with _engine(w) as e:
        st = e.state()
        stats = e.stats()
        assert stats.path == plan.path
        assert (stats.nx, stats.ny, stats.nz) == (plan.nx, plan.ny, plan.nz)
        counts, lst = e.neighbors()
        if plan.path != O.PATH_DIRECT:
            cells = e.cells(plan.n_cells)
            ref_cells = O.cell_ids_f64(w, st["positions"])
            np.testing.assert_array_equal(cells["cell_of"], ref_cells)
            ref_counts = np.bincount(ref_cells, minlength=plan.n_cells)
            np.testing.assert_array_equal(cells["cell_count"], ref_counts)
            # members of a cell in (x, slot) order, x in binary64
            x = st["positions"][:, 0]
            want = np.lexsort((np.arange(w.n), x, ref_cells))
            np.testing.assert_array_equal(cells["cell_atoms"], want)
    rc, rl, ovf = O.neighbor_lists_f64(w, st["positions"], tags=st["tags"], types=st["types"])
    assert not ovf
    np.testing.assert_array_equal(counts, rc)
    for i in range(w.n):
        got = np.sort(lst[: counts[i], i])
        want = np.sort(rl[i, : rc[i]])
        assert np.array_equal(got, want), f"atom {i}: neighbour sets differ"


@pytest.mark.parametrize("name", list(CASES))
def test_forces_and_observables(name):
    w = CASES[name]
    with _engine(w) as e:
        st = e.state()
        stats = e.stats()
    f_ref, obs = O.forces_real64(w, st["positions"], tags=st["tags"], types=st["types"])
    err = _force_error(st["forces"], f_ref)
    print(f"{name}: force error {err:.2e}")
    assert err <= FORCE_RTOL, f"force error {err:.3e}"
    for q, term in enumerate(("pair", "bond", "angle", "dihedral")):
        ref_pe, ref_w = obs[f"{term}_pe"], obs[f"{term}_virial"]
        assert abs(stats.pe[q] - ref_pe) <= ENERGY_RTOL * max(1.0, abs(ref_pe)), (term, stats.pe[q], ref_pe)
        assert abs(stats.virial[q] - ref_w) <= ENERGY_RTOL * max(1.0, abs(ref_w)), (
            term, stats.virial[q], ref_w)
    v, m = st["velocities"], st["masses"]
    ke = 0.5 * float(np.sum(m[:, None] * v * v))
    assert abs(stats.ke - ke) <= ENERGY_RTOL * ke


@pytest.mark.parametrize("name", ["lj_regular", "melt"])
def test_one_step_matches_oracle(name):
    w = CASES[name].with_(berendsen=False)
    with _engine(w) as e:
        s0 = e.state()
        e.advance(1)
        s1 = e.state()
    assert np.array_equal(s0["tags"], s1["tags"])
    dt = float(w.dt)
    h = 0.5 * dt / s0["masses"][:, None]
    v_half = s0["velocities"] + h * s0["forces"]
    lo = w.box_bound[0]
    L = w.box_bound[1] - w.box_bound[0]
    x = s0["positions"] + dt * v_half
    c = np.floor((x - lo) / L)
    x = x - c * L
    dx = s1["positions"] - x
    dx -= np.round(dx / L) * L
    print(f"{name}: one-step position error {np.max(np.abs(dx)):.2e}")
    assert np.max(np.abs(dx)) <= STEP_ATOL * float(L.max())
    f1, _ = O.forces_real64(w, s1["positions"], tags=s1["tags"], types=s1["types"])
    assert _force_error(s1["forces"], f1) <= FORCE_RTOL
    v_ref = v_half + h * f1
    vscale = np.sqrt(np.mean(v_ref ** 2))
    assert np.max(np.abs(s1["velocities"] - v_ref)) <= FORCE_RTOL * vscale
    np.testing.assert_array_equal(s1["images"] - s0["images"], c.astype(np.int32))


def _by_tag(state):
    order = np.argsort(state["tags"])
    return {k: state[k][order] for k in ("positions", "velocities", "images")}


@pytest.mark.parametrize("name,steps,thermostat", [("lj_regular", 200, False),
                                                   ("melt", 200, False),
                                                   ("melt", 100, True)])
def test_run_matches_binary64_oracle_run(name, steps, thermostat):
    """The dynamics loop (integrate, rebuild checks, reorders, forces,
    Berendsen) against the oracle's binary64 run from the same state."""
    w = CASES[name].with_(berendsen=thermostat, temperature=1.3 * CASES[name].temperature,
                          tau=0.5, sort_every_rebuild=2)
    with _engine(w) as e:
        e.advance(steps, observables_on_last=True)
        st = e.state()
        stats = e.stats()
    ref, rstats = O.run_f64(w, w.positions, w.velocities, steps=steps)
    a, b = _by_tag(st), _by_tag(ref)
    L = (w.box_bound[1] - w.box_bound[0])
    ua = a["positions"] + a["images"] * L
    ub = b["positions"] + b["images"] * L
    x_err = float(np.max(np.abs(ua - ub))) / float(L.max())
    # the engine holds the last Berendsen factor pending; the oracle applied it
    va = a["velocities"] * stats.lambda_
    vscale = float(np.sqrt(np.mean(b["velocities"] ** 2)))
    v_err = float(np.max(np.abs(va - b["velocities"]))) / vscale
    print(f"{name} {steps} steps thermostat={thermostat}: position error {x_err:.2e} (of L), "
          f"velocity error {v_err:.2e} (of rms), builds {stats.build_count}/{rstats.build_count}")
    assert stats.reorder_count >= 1 and stats.overflow == 0
    assert x_err <= RUN_RTOL and v_err <= RUN_RTOL
    pe = sum(stats.pe)
    pe_ref = sum(rstats.final_obs.as_dict()[f"{t}_pe"] for t in ("pair", "bond", "angle",
                                                                   "dihedral"))
    assert abs(pe - pe_ref) <= RUN_RTOL * max(1.0, abs(pe_ref))


def test_batched_replicas_match_single_double_engines():
    ws = [W.lj_box(n, 0.8, 1.0, seed=10 + k).as_float64(jitter=1e-4, seed=k)
          for k, n in enumerate((500, 1500, 4000))]
    steps = 20
    with Batch(ws, precision="double") as b:
        b.run(steps, observables_on_last=True)
        got = [b.state(r) for r in range(len(ws))]
    for w, g in zip(ws, got):
        with Engine(w, precision="double") as e:
            e.initialize(observables=True)
            e.advance(steps)
            st = e.state()
        a, c = _by_tag(g), _by_tag(st)
        L = w.box_bound[1] - w.box_bound[0]
        dx = (a["positions"] + a["images"] * L) - (c["positions"] + c["images"] * L)
        assert np.max(np.abs(dx)) <= RUN_RTOL * float(L.max())


def test_batched_bonded_replicas_match_single_double_engines():
    """Melts with bonds, angles, dihedrals and exclusions (1,200 beads on the
    cell path, reordering every 2nd build; 216 beads on two cells per axis)
    beside a pair-only box, batched in binary64 in both modes: after 20 steps
    positions and the four energies match single binary64 engines."""
    ws = [W.polymer_melt(n_chains=40, beads_per_chain=30, seed=11, jitter=0.05)
          .with_(sort_every_rebuild=2).as_float64(jitter=1e-4, seed=1),
          W.polymer_melt(n_chains=18, beads_per_chain=12, density=7.5, seed=13, jitter=0.05)
          .as_float64(jitter=1e-4, seed=2),
          W.lj_box(500, 0.8, 1.0, seed=14).as_float64(jitter=1e-4, seed=3)]
    steps = 20
    for config in (1, 3):
        with Batch(ws, precision="double") as b:
            b.set_config(config)
            b.run(steps, observables_on_last=True)
            got = [(b.state(r), b.stats(r)) for r in range(len(ws))]
        for w, (g, gs) in zip(ws, got):
            with Engine(w, precision="double") as e:
                e.initialize(observables=True)
                e.advance(steps, observables_on_last=True)
                st, ss = e.state(), e.stats()
            assert (gs.build_count, gs.reorder_count) == (ss.build_count, ss.reorder_count)
            a, c = _by_tag(g), _by_tag(st)
            L = w.box_bound[1] - w.box_bound[0]
            dx = (a["positions"] + a["images"] * L) - (c["positions"] + c["images"] * L)
            assert np.max(np.abs(dx)) <= RUN_RTOL * float(L.max()), (config, w.name)
            for q in range(4):
                assert abs(gs.pe[q] - ss.pe[q]) <= RUN_RTOL * max(1.0, abs(ss.pe[q])), (q, w.name)


def test_trajectory_and_snapshot_are_binary64_through_the_public_api(tmp_path):
    """BEADS_AMD_PRECISION=double selects _native_real64: the public builders
    validate binary64 arrays and the run writes real size 8 files whose
    contents equal the engine handle's binary64 state."""
    w = CASES["melt"].with_(steps=20)
    np.save(tmp_path / "pos.npy", w.positions)
    repo = Path(__file__).resolve().parents[1]
    run = str(tmp_path / "run")
    script = f"""
import sys
sys.path[:0] = [{str(repo)!r}, {str(repo / 'tests')!r}]
import numpy as np
import beads_amd as beads
assert beads.build_info()["real_dtype"] == "float64"
from gpu_cases import cases
w = cases()["melt"]
w = w.as_float64().with_(steps=20, positions=np.load({str(tmp_path / 'pos.npy')!r}),
                         cutoff=float(np.float32(w.cutoff)),
                         cutoff_buffer=float(np.float32(w.cutoff_buffer)))
sim = w.to_simulation()
assert sim._system.positions.dtype == np.float64
sim.set_trajectory(every=20, prefix={run!r},
                   fields=["tag", "type", "position", "image", "velocity", "force"])
sim.save_final_state(prefix={run!r})
sim.set_logging(echo="none")
sim.execute()
snap = beads.System.from_state({run + '.state.beadsbin'!r})
assert snap.positions.dtype == np.float64
order = np.argsort(np.asarray(snap.tags))
np.save({str(tmp_path / 'snap.npy')!r}, np.asarray(snap.positions)[order])
print("executed", flush=True)
"""
    env = dict(os.environ, BEADS_AMD_PRECISION="double")
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    from beads_amd.io import read_beads_trajectory

    traj = read_beads_trajectory(tmp_path / "run.trajectory.beadsbin")
    assert traj["real_size"] == 8 and traj["steps"].tolist() == [0, 20]
    rec = traj["records"][-1]
    assert rec["position"].dtype == np.float64
    order = np.argsort(rec["tag"])
    np.testing.assert_array_equal(np.load(tmp_path / "snap.npy"), rec["position"][order])
    rec0 = traj["records"][0]
    o0 = np.argsort(rec0["tag"])
    np.testing.assert_array_equal(rec0["position"][o0], w.positions)
    with _engine(w) as e:
        e.advance(20)
        st = e.state()
    a = _by_tag(st)
    L = w.box_bound[1] - w.box_bound[0]
    np.testing.assert_allclose(rec["position"][order], a["positions"], rtol=0,
                               atol=RUN_RTOL * float(L.max()))


# ------------------------------------------------------------- adjoint -----

def _rel_err(g, ref):
    g = np.asarray(g, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-2 * scale)))


def _adjoint_system():
    return _binary64_case(W.polymer_solvent(n_chains=12, beads_per_chain=20, n_solvent=900,
                                            seed=3, jitter=0.03, shift=True), 2)


def _pow2_step(theta: float) -> float:
    """2^(e - 20) for theta in [2^e, 2^(e+1)): theta +- h are binary32 values,
    so the oracle (binary32 parameters) sees the perturbation exactly."""
    return 2.0 ** (np.floor(np.log2(abs(theta))) - 20)


def test_adjoint_force_matching_gradients_match_binary64_differences():
    import torch

    from beads_amd.adjoint import FrameModel, force_matching_loss, force_matching_value_and_grad

    w = _adjoint_system()
    rng = np.random.default_rng(7)
    lo, L = w.box_bound[0], w.box_bound[1] - w.box_bound[0]
    frames = [lo + np.mod(w.positions + rng.uniform(-0.02, 0.02, w.positions.shape) - lo, L)
              for _ in range(3)]
    w_t = w.with_(lj_epsilon=(w.lj_epsilon * 1.1).astype(np.float32),
                  bond_r0=(w.bond_r0 * 1.02).astype(np.float32))
    targets = [O.forces_real64(w_t, x)[0] for x in frames]
    n3 = 3 * w.n * len(frames)

    def loss_at(wp, xs=frames):
        return sum(float(((O.forces_real64(wp, x)[0] - t) ** 2).sum())
                   for x, t in zip(xs, targets)) / n3

    with FrameModel(w, precision="double") as m:
        p = m.initial_params(requires_grad=False)
        xs = [torch.tensor(x, device="cuda") for x in frames]
        ts = [torch.tensor(t, device="cuda") for t in targets]
        loss, grads = force_matching_value_and_grad(m, xs, ts, p)
        m.check()
        g = [t.cpu().numpy() for t in grads.tensors()]
        # positions: autograd through the frame forces
        x0 = xs[0].clone().requires_grad_(True)
        lx = force_matching_loss(m, [x0] + xs[1:], ts, p)
        lx.backward()
        gx = x0.grad.cpu().numpy()
    assert float(loss) == pytest.approx(loss_at(w), rel=1e-12)
    T = w.n_types
    fd, got = [], []
    for which, name in ((0, "lj_epsilon"), (1, "lj_sigma")):
        for a in range(T):
            for b in range(a, T):
                base = getattr(w, name)
                theta = float(base[a, b])
                h = _pow2_step(theta)
                vals = []
                for sgn in (1, -1):
                    arr = base.copy()
                    arr[a, b] = arr[b, a] = theta + sgn * h
                    vals.append(loss_at(w.with_(**{name: arr})))
                fd.append((vals[0] - vals[1]) / (2 * h))
                got.append(g[which][a, b] + (g[which][b, a] if a != b else 0.0))
    for which, name in ((2, "bond_k"), (3, "bond_r0")):
        base = getattr(w, name)
        theta = float(base[0])
        h = _pow2_step(theta)
        vals = []
        for sgn in (1, -1):
            arr = base.copy()
            arr[0] = theta + sgn * h
            vals.append(loss_at(w.with_(**{name: arr})))
        fd.append((vals[0] - vals[1]) / (2 * h))
        got.append(g[which][0])
    err = _rel_err(got, fd)
    print(f"force-matching parameter gradient error {err:.2e}")
    assert err <= GRAD_RTOL
    # dL/dx on a few coordinates of frame 0 (central differences, h = 1e-7)
    h = 1e-7
    fdx, gotx = [], []
    for i in rng.choice(w.n, 6, replace=False):
        for c in range(3):
            vals = []
            for sgn in (1, -1):
                xp = [x.copy() for x in frames]
                xp[0][i, c] += sgn * h
                vals.append(loss_at(w, xp))
            fdx.append((vals[0] - vals[1]) / (2 * h))
            gotx.append(gx[i, c])
    errx = _rel_err(gotx, fdx)
    print(f"force-matching position gradient error {errx:.2e}")
    assert errx <= 1e-5  # the FD step in x is limited by the r^-14 curvature


@pytest.mark.parametrize("steps", [30, 200])
def test_trajectory_adjoint_passes_the_dot_product_test(steps):
    """<g, T d> = <T^T g, d> for the derivative T of a velocity-Verlet
    trajectory (x0, v0) -> (x_T, v_T): T^T g from the engine's reverse sweep
    (adjoint kernel over a trajectory recomputed from checkpoints every 20
    steps), T d from the oracle's tangent run in binary64 (its forward-mode
    derivative, pinned against central differences on the CPU). Finite
    differences of whole runs cannot check a 200-step gradient (rounding of
    the two perturbed runs, pairs crossing r_c between them); the dot-product
    identity holds to rounding at any horizon."""
    import torch

    from beads_amd.adjoint import FrameModel

    w = _adjoint_system().with_(berendsen=False, cutoff_buffer=0.25, rebuild_check_every=1)
    rng = np.random.default_rng(11)
    gx_T, gv_T = rng.standard_normal((w.n, 3)), rng.standard_normal((w.n, 3))
    dx0, dv0 = rng.standard_normal((w.n, 3)), rng.standard_normal((w.n, 3))
    L = w.box_bound[1] - w.box_bound[0]
    with FrameModel(w, precision="double") as m:
        p = m.initial_params(requires_grad=False)
        x0 = torch.tensor(w.positions, device="cuda", requires_grad=True)
        v0 = torch.tensor(w.velocities, device="cuda", requires_grad=True)
        xT, vT = m.trajectory(x0, v0, p, steps, checkpoint_every=20)
        loss = (xT * torch.tensor(gx_T, device="cuda")).sum() + \
            (vT * torch.tensor(gv_T, device="cuda")).sum()
        loss.backward()
        gx0, gv0 = x0.grad.cpu().numpy(), v0.grad.cpu().numpy()
        xT_np = xT.detach().cpu().numpy()
    ref, tan, _ = O.run_tangent_f64(w, w.positions, w.velocities, dx0, dv0, steps=steps)
    o = np.argsort(ref["tags"])
    fwd = float(np.max(np.abs(xT_np - (ref["positions"][o] + ref["images"][o] * L))))
    lhs_terms = np.concatenate([(gx_T * tan["dpos"][o]).ravel(), (gv_T * tan["dvel"][o]).ravel()])
    rhs_terms = np.concatenate([(gx0 * dx0).ravel(), (gv0 * dv0).ravel()])
    lhs, rhs = float(lhs_terms.sum()), float(rhs_terms.sum())
    rel = abs(lhs - rhs) / max(abs(lhs), 1e-300)
    print(f"{steps} steps: forward difference {fwd:.2e}, <g, T d> = {lhs:.12e}, "
          f"<T^T g, d> = {rhs:.12e}, relative difference {rel:.2e}")
    assert fwd <= 1e-10
    assert rel <= DOT_RTOL
