"""Adjoint kernel and trajectory backprop on the GPU, against the FP64 oracle.

Gradients of the force-matching loss, of the potential energy and of a loss
on the end state of a velocity-Verlet trajectory are compared with central
finite differences of the same quantity evaluated by the FP64 oracle
(``oracle.forces_f64``). Parameter steps are h = 1e-4 |theta| rounded to the
binary32 grid (the engine's parameters are binary32); the difference quotient
divides by the exact FP64 distance between the two binary32 values.

Tolerance (stated here and applied per entry): |g - g_fd| <= 1e-4 *
max(|g_fd|, 1e-2 * max|g_fd|) for the force-matching, energy and trajectory
gradients. The trajectory check runs 30 steps: the binary32 forward
trajectory departs from the FP64 one by rounding that the dynamics amplifies
(measured on this system: largest position difference 1.3e-6 nm after 10
steps, 2.4e-6 after 30, 2.6e-5 after 100, 9.5e-4 after 200; gradient error
against FP64 differences at most 4.3e-5 after 30 steps and 2.4e-4 after 100),
and after about 100 steps no finite-difference step keeps every pair on the
same side of r_c in both perturbed runs. Longer horizons are checked by
property instead: the gradient does not depend on the checkpoint interval.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from beads_amd import workloads as W
from beads_amd.adjoint import FrameModel, Params, force_matching_loss
from beads_amd.capi import Engine

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-4
TRAJ_RTOL = 1e-4
FORCE_RTOL = 1e-5


def small_system(seed=3, shift=True):
    return W.polymer_solvent(n_chains=12, beads_per_chain=20, n_solvent=900, seed=seed,
                             jitter=0.03, shift=shift)


def _rel_err(g, ref):
    g = np.asarray(g, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-2 * scale)))


def _force_error(f_gpu, f_ref):
    scale = np.sqrt(np.mean(f_ref ** 2))
    return float(np.max(np.abs(f_gpu - f_ref) / np.maximum(np.abs(f_ref), scale)))


def _frames(w, count, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    lo = w.box_bound[0].astype(np.float64)
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    out = []
    for _ in range(count):
        p = w.positions.astype(np.float64) + rng.uniform(-amp, amp, (w.n, 3))
        p = lo + np.mod(p - lo, L)
        out.append(p.astype(np.float32))
    return out


def _with_params(w, eps, sig, k, r0):
    return w.with_(lj_epsilon=np.asarray(eps, np.float32), lj_sigma=np.asarray(sig, np.float32),
                   bond_k=np.asarray(k, np.float32), bond_r0=np.asarray(r0, np.float32))


def _param_vector(w):
    return [w.lj_epsilon.astype(np.float32).copy(), w.lj_sigma.astype(np.float32).copy(),
            w.bond_k.astype(np.float32).copy(), w.bond_r0.astype(np.float32).copy()]


def _fd_param(loss_fn, params, which, idx, h_rel=1e-4):
    """Central difference of loss_fn(params) in one parameter (symmetric pair for
    the [T, T] tables), h = h_rel |theta| on the binary32 grid."""
    base = params[which]
    theta = float(base[idx])
    h = h_rel * abs(theta)
    plus = [p.copy() for p in params]
    minus = [p.copy() for p in params]
    tp, tm = np.float32(theta + h), np.float32(theta - h)
    for arr, val in ((plus[which], tp), (minus[which], tm)):
        arr[idx] = val
        if arr.ndim == 2:
            arr[idx[::-1]] = val
    return (loss_fn(plus) - loss_fn(minus)) / (float(tp) - float(tm))


def _param_indices(T):
    return [(0, (a, b)) for a in range(T) for b in range(a, T)] + \
           [(1, (a, b)) for a in range(T) for b in range(a, T)] + [(2, 0), (3, 0)]


def _gpu_param_grad(grads, which, idx):
    g = grads[which]
    if g.ndim == 2:
        a, b = idx
        return float(g[a, b] + g[b, a]) if a != b else float(g[a, b])
    return float(g[idx])


def test_frame_forces_match_dynamics_engine_and_oracle():
    w = small_system()
    with FrameModel(w) as m:
        x = torch.tensor(w.positions, device="cuda")
        p = m.initial_params(requires_grad=False)
        f, e = m.forces(x, p)
        f = f.cpu().numpy()
    with Engine(w) as eng:
        eng.initialize(observables=True)
        st = eng.state()
        stats = eng.stats()
    np.testing.assert_array_equal(st["tags"], w.tags)
    np.testing.assert_array_equal(f, st["forces"])
    assert float(e) == pytest.approx(stats.pe[0] + stats.pe[1], rel=1e-12)
    f_ref, obs = O.forces_f64(w, w.positions.astype(np.float64))
    assert _force_error(f.astype(np.float64), f_ref) <= FORCE_RTOL
    assert float(e) == pytest.approx(obs["pair_pe"] + obs["bond_pe"], rel=1e-5)


def test_stream_ordered_frames_match_synchronous_calls():
    """FrameModel on a side torch stream (stream-ordered entry points, no host
    wait per call) against the synchronous entry points of a second engine
    driven through the same call sequence: forward frames 0..2 keeping the
    list when the skin allows, then backward 2, reload 1, backward 1, reload
    0, backward 0. Forces, energies and dL/dx are bit-identical. Parameter
    gradients are FP64 per frame (summed with per-workgroup atomics) and
    autograd accumulates them over frames in the parameters' binary32 dtype,
    so they agree with the FP64 sum of the engine's outputs to 1e-6."""
    w = small_system(seed=7)
    frames = _frames(w, 3, seed=11)
    gen = torch.Generator(device="cuda").manual_seed(5)
    xs = [torch.tensor(x, device="cuda") for x in frames]
    gs = [torch.randn(x.shape, generator=gen, device="cuda") for x in xs]
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    with FrameModel(w) as m:
        with torch.cuda.stream(side):
            p = m.initial_params()
            xr = [x.clone().requires_grad_(True) for x in xs]
            fs, es, loss = [], [], None
            for x, g in zip(xr, gs):
                f, e = m.forces(x, p)
                fs.append(f)
                es.append(e)
                term = (f.double() * g.double()).sum() + 0.5 * e
                loss = term if loss is None else loss + term
            loss.backward()
            m.check()
            out_f = [f.detach().cpu().numpy() for f in fs]
            out_e = [float(e.detach()) for e in es]
            out_gx = [x.grad.cpu().numpy() for x in xr]
            out_gp = np.concatenate([t.grad.double().cpu().numpy().ravel()
                                     for t in p.tensors()])
        assert m.backward_calls == 3 and m.reload_calls == 2
    n_p = 2 * w.n_types * w.n_types + 2 * int(w.bond_k.shape[0])
    with Engine(w) as eng:
        ref_f, ref_e, ref_gx = [], [], [None] * 3
        gp_sum = np.zeros(n_p)
        for x in xs:
            f = torch.empty_like(x)
            ref_e.append(float(np.sum(eng.frame_forces(x.data_ptr(), f.data_ptr(),
                                                       reuse_list=True))))
            ref_f.append(f.cpu().numpy())
        for i in (2, 1, 0):
            if i != 2:
                eng.frame_load(xs[i].data_ptr(), reuse_list=True)
            gx = torch.empty_like(xs[i])
            gp = torch.empty(n_p, dtype=torch.float64, device="cuda")
            eng.frame_backward(gs[i].data_ptr(), 0.5, gx.data_ptr(), gp.data_ptr())
            ref_gx[i] = gx.cpu().numpy()
            gp_sum += gp.cpu().numpy()
    for i in range(3):
        np.testing.assert_array_equal(out_f[i], ref_f[i])
        assert out_e[i] == ref_e[i]
        np.testing.assert_array_equal(out_gx[i], ref_gx[i])
    # p.tensors() holds eps, sigma as [T, T] and k, r0 as [nb]: the same order
    # as the engine's parameter vector.
    np.testing.assert_allclose(out_gp, gp_sum, rtol=1e-6, atol=1e-6 * np.abs(gp_sum).max())


def test_stream_ordered_overflow_reported_by_check():
    """A list overflow in a stream-ordered frame call surfaces at check()."""
    w = W.lj_box(1200, 0.8, 1.0, seed=4).with_(max_neighbors=8)
    with FrameModel(w) as m:
        p = m.initial_params(requires_grad=False)
        m.forces(torch.tensor(w.positions, device="cuda"), p)
        with pytest.raises(Exception, match="overflow"):
            m.check()


def test_force_matching_parameter_gradients_vs_fd():
    w = small_system()
    frames = _frames(w, 3, seed=9)
    # Targets: FP64 forces at "true" parameters, 4 % away from the model's.
    rng = np.random.default_rng(4)
    true = _param_vector(w)
    true = [np.asarray(t * rng.uniform(0.96, 1.04, t.shape), np.float32) for t in true]
    true[0] = np.triu(true[0]) + np.triu(true[0], 1).T
    true[1] = np.triu(true[1]) + np.triu(true[1], 1).T
    wt = _with_params(w, *true)
    targets = [O.forces_f64(wt, x.astype(np.float64))[0] for x in frames]

    def loss64(params):
        wp = _with_params(w, *params)
        tot = 0.0
        for x, ft in zip(frames, targets):
            f = O.forces_f64(wp, x.astype(np.float64))[0]
            tot += np.sum((f - ft) ** 2)
        return tot / (len(frames) * w.n * 3)

    with FrameModel(w) as m:
        p = m.initial_params()
        xs = [torch.tensor(x, device="cuda") for x in frames]
        ts = [torch.tensor(t, device="cuda") for t in targets]
        loss = force_matching_loss(m, xs, ts, p)
        loss.backward()
        grads = [t.grad.cpu().numpy().astype(np.float64) for t in p.tensors()]
        fwd_s, bwd_s = m.engine.frame_timing()
    assert fwd_s > 0 and bwd_s > 0
    params = _param_vector(w)
    g_gpu, g_fd = [], []
    for which, idx in _param_indices(w.n_types):
        g_gpu.append(_gpu_param_grad(grads, which, idx))
        g_fd.append(_fd_param(loss64, params, which, idx))
    assert _rel_err(g_gpu, g_fd) <= GRAD_RTOL, (g_gpu, g_fd)


def test_force_matching_position_gradient_vs_fd():
    w = small_system(seed=5)
    x0 = _frames(w, 1, seed=2)[0]
    target = O.forces_f64(w, x0.astype(np.float64))[0] * 0.9
    rng = np.random.default_rng(1)
    # bonded beads, solvent beads and a few at random
    atoms = [0, 1, 19, 20, 250, 700, int(w.n - 1)] + list(rng.integers(0, w.n, 5))
    coords = [(a, int(c)) for a in atoms for c in (0, 1, 2)]

    def loss64(pos64):
        f = O.forces_f64(w, pos64)[0]
        return np.sum((f - target) ** 2) / (w.n * 3)

    with FrameModel(w) as m:
        p = m.initial_params(requires_grad=False)
        x = torch.tensor(x0, device="cuda", requires_grad=True)
        loss = force_matching_loss(m, [x], [torch.tensor(target, device="cuda")], p)
        loss.backward()
        gx = x.grad.cpu().numpy().astype(np.float64)
    base = x0.astype(np.float64)
    g_fd = []
    for a, c in coords:
        h = 1e-5
        pp, pm = base.copy(), base.copy()
        pp[a, c] += h
        pm[a, c] -= h
        g_fd.append((loss64(pp) - loss64(pm)) / (2 * h))
    g_gpu = [gx[a, c] for a, c in coords]
    # The device evaluates J^T g from binary32 coordinates: hold it to the
    # float32 displacement rounding (|dx| ulp / |dx| ~ 1e-6 relative per pair).
    assert _rel_err(g_gpu, g_fd) <= GRAD_RTOL, (g_gpu, g_fd)


@pytest.mark.parametrize("shift", [True, False])
def test_energy_gradients(shift):
    w = small_system(seed=7, shift=shift)
    x0 = _frames(w, 1, seed=3)[0]

    def energy64(params):
        o = O.forces_f64(_with_params(w, *params), x0.astype(np.float64))[1]
        return o["pair_pe"] + o["bond_pe"]

    with FrameModel(w) as m:
        p = m.initial_params()
        x = torch.tensor(x0, device="cuda", requires_grad=True)
        f, e = m.forces(x, p)
        e.backward()
        grads = [t.grad.cpu().numpy().astype(np.float64) for t in p.tensors()]
        gx = x.grad.cpu().numpy()
        f = f.detach().cpu().numpy()
    # dE/dx = -F: the adjoint's FP64 force against the forward kernel's binary32
    # force, held to the force tolerance.
    assert _force_error(-gx.astype(np.float64), f.astype(np.float64)) <= FORCE_RTOL
    params = _param_vector(w)
    g_gpu = [_gpu_param_grad(grads, which, idx) for which, idx in _param_indices(w.n_types)]
    g_fd = [_fd_param(energy64, params, which, idx) for which, idx in _param_indices(w.n_types)]
    assert _rel_err(g_gpu, g_fd) <= GRAD_RTOL, (g_gpu, g_fd)


def _oracle_trajectory(w, params, x0, v0, steps):
    """FP64 velocity Verlet with oracle forces; also returns, per force
    evaluation, the set of pairs closer than r_c (periodic KD-tree)."""
    from scipy.spatial import cKDTree

    wp = _with_params(w, *params)
    dt = float(w.dt)
    a = 0.5 * dt / w.masses.astype(np.float64)[:, None]
    lo = w.box_bound[0].astype(np.float64)
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    inside = []

    def force(x):
        xw = lo + np.mod(x - lo, L)
        tree = cKDTree(np.minimum(xw - lo, np.nextafter(L, 0)), boxsize=L)
        inside.append(tree.query_pairs(float(w.cutoff), output_type="ndarray"))
        return O.forces_f64(wp, xw)[0]

    x, v = x0.astype(np.float64).copy(), v0.astype(np.float64).copy()
    f = force(x)
    for _ in range(steps):
        v = v + a * f
        x = x + dt * v
        f = force(x)
        v = v + a * f
    return x, v, inside


def _same_pairs(a, b):
    if len(a) != len(b):
        return False
    for p, q in zip(a, b):
        if p.shape != q.shape:
            return False
        kp = np.sort(p[:, 0].astype(np.int64) * 1_000_003 + p[:, 1])
        kq = np.sort(q[:, 0].astype(np.int64) * 1_000_003 + q[:, 1])
        if not np.array_equal(kp, kq):
            return False
    return True


def _smooth_central_difference(run, steps_h):
    """run(h) -> (loss, inside-pair sets) at the perturbed point h (h may be
    negative). The cut-off makes the force jump at r_c, so a difference
    quotient whose two runs put some pair on opposite sides of r_c at some
    force evaluation measures a jump that the derivative does not have. Take
    the largest step whose two runs see identical pair sets everywhere."""
    for h in steps_h:
        lp, sp = run(h)
        lm, sm = run(-h)
        if _same_pairs(sp, sm):
            return (lp - lm) / (2 * h)
    raise AssertionError("no finite-difference step keeps the cut-off crossings unchanged")


def test_trajectory_backprop_vs_fd():
    w = small_system(seed=11)
    steps, every = 30, 7
    rng = np.random.default_rng(8)
    cx = rng.normal(size=(w.n, 3))
    cv = rng.normal(size=(w.n, 3)) * 0.1

    def loss_pairs(params, x0=w.positions, v0=w.velocities):
        x, v, inside = _oracle_trajectory(w, params, x0, v0, steps)
        return float(np.sum(cx * x) + np.sum(cv * v)), inside

    with FrameModel(w) as m:
        p = m.initial_params()
        x0 = torch.tensor(w.positions, device="cuda", requires_grad=True)
        v0 = torch.tensor(w.velocities, device="cuda", requires_grad=True)
        xT, vT = m.trajectory(x0, v0, p, steps=steps, checkpoint_every=every)
        loss = (torch.tensor(cx, device="cuda") * xT.double()).sum() + \
               (torch.tensor(cv, device="cuda") * vT.double()).sum()
        loss.backward()
        grads = [t.grad.cpu().numpy().astype(np.float64) for t in p.tensors()]
        gx0 = x0.grad.cpu().numpy().astype(np.float64)
        gv0 = v0.grad.cpu().numpy().astype(np.float64)
    params = _param_vector(w)
    idx = [(0, (0, 0)), (0, (0, 2)), (1, (1, 1)), (1, (2, 3)), (2, 0), (3, 0)]
    g_gpu = [_gpu_param_grad(grads, which, i) for which, i in idx]
    g_fd = []
    for which, i in idx:
        theta = float(params[which][i])

        def run(h, which=which, i=i, theta=theta):
            q = [t.copy() for t in params]
            val = np.float32(theta + h)
            q[which][i] = val
            if q[which].ndim == 2:
                q[which][i[::-1]] = val
            loss, inside = loss_pairs(q)
            return loss, inside

        # steps on the binary32 grid: the quotient divides by the exact distance
        for h_rel in (1e-4, 2e-5, 4e-6, 1e-6, 2.5e-7):
            h = h_rel * abs(theta)
            tp, tm = np.float32(theta + h), np.float32(theta - h)
            lp, sp = run(float(tp) - theta)
            lm, sm = run(float(tm) - theta)
            if _same_pairs(sp, sm):
                g_fd.append((lp - lm) / (float(tp) - float(tm)))
                break
        else:
            raise AssertionError(f"no smooth finite-difference step for parameter {which} {i}")
    assert _rel_err(g_gpu, g_fd) <= TRAJ_RTOL, (g_gpu, g_fd)
    # initial-state gradients for a few coordinates
    base_x = w.positions.astype(np.float64)
    base_v = w.velocities.astype(np.float64)
    fd_x, fd_v, gpu_x, gpu_v = [], [], [], []
    for a, c in [(0, 0), (5, 1), (300, 2), (900, 0)]:
        def run_x(h, a=a, c=c):
            q = base_x.copy()
            q[a, c] += h
            return loss_pairs(params, q, base_v)

        def run_v(h, a=a, c=c):
            q = base_v.copy()
            q[a, c] += h
            return loss_pairs(params, base_x, q)

        hs = (1e-5, 2e-6, 4e-7, 1e-7)
        fd_x.append(_smooth_central_difference(run_x, hs))
        fd_v.append(_smooth_central_difference(run_v, hs))
        gpu_x.append(gx0[a, c])
        gpu_v.append(gv0[a, c])
    assert _rel_err(gpu_x, fd_x) <= TRAJ_RTOL, (gpu_x, fd_x)
    assert _rel_err(gpu_v, fd_v) <= TRAJ_RTOL, (gpu_v, fd_v)


def test_frame_mode_rejects_dynamics_mixing():
    ws = small_system()
    with FrameModel(ws) as m:
        x = torch.tensor(ws.positions, device="cuda")
        m.forces(x, m.initial_params(requires_grad=False))
        with pytest.raises(Exception):
            m.engine.initialize()


def test_fused_value_and_grad_matches_autograd():
    """force_matching_value_and_grad (adjoint right after each forward, no
    autograd graph) gives the autograd loss to FP64 rounding and the autograd
    parameter gradients to 1e-6 (autograd accumulates the per-frame FP64
    gradients over frames in binary32)."""
    from beads_amd.adjoint import force_matching_value_and_grad

    w = small_system(seed=11)
    frames = _frames(w, 4, seed=13)
    gen = torch.Generator(device="cuda").manual_seed(3)
    with FrameModel(w) as m:
        p = m.initial_params()
        xs = [torch.tensor(x, device="cuda") for x in frames]
        with torch.no_grad():
            ts = [m.forces(x, p)[0] * 1.05 + 0.1 * torch.randn(x.shape, generator=gen, device="cuda")
                  for x in xs]
        loss = force_matching_loss(m, xs, ts, p)
        loss.backward()
        ref = [t.grad.double().cpu().numpy() for t in p.tensors()]
        value, grads = force_matching_value_and_grad(m, xs, ts, p)
        got = [t.double().cpu().numpy() for t in grads.tensors()]
        m.check()
    assert float(value) == pytest.approx(float(loss.detach()), rel=1e-12)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g, r, rtol=1e-6, atol=1e-6 * max(np.abs(r).max(), 1e-30))


def test_full_size_force_matching_and_energy_gradients():
    """The 100,000-bead, four-type polymer/solvent force-matching system at its
    benchmark size: the 22 parameter gradients of the force-matching loss over
    two frames against FP64 central differences of the oracle's loss (bar
    GRAD_RTOL), and dE/dx = -F on every bead (force bar)."""
    w = W.polymer_solvent()
    assert w.n == 100_000 and w.n_types == 4
    frames = _frames(w, 2, seed=23)
    rng = np.random.default_rng(29)
    true = _param_vector(w)
    true = [np.asarray(t * rng.uniform(0.97, 1.03, t.shape), np.float32) for t in true]
    true[0] = np.triu(true[0]) + np.triu(true[0], 1).T
    true[1] = np.triu(true[1]) + np.triu(true[1], 1).T
    wt = _with_params(w, *true)
    targets = [O.forces_f64(wt, x.astype(np.float64))[0] for x in frames]

    def loss64(params):
        wp = _with_params(w, *params)
        tot = 0.0
        for x, ft in zip(frames, targets):
            f = O.forces_f64(wp, x.astype(np.float64))[0]
            tot += np.sum((f - ft) ** 2)
        return tot / (len(frames) * w.n * 3)

    with FrameModel(w) as m:
        p = m.initial_params()
        xs = [torch.tensor(x, device="cuda") for x in frames]
        ts = [torch.tensor(t, device="cuda") for t in targets]
        loss = force_matching_loss(m, xs, ts, p)
        loss.backward()
        m.check()
        grads = [t.grad.cpu().numpy().astype(np.float64) for t in p.tensors()]
        print(f"loss {float(loss.detach()):.6e}", flush=True)
        q = m.initial_params(requires_grad=False)
        x = xs[0].clone().requires_grad_(True)
        f, e = m.forces(x, q)
        e.backward()
        m.check()
        gx = x.grad.cpu().numpy().astype(np.float64)
        f = f.detach().cpu().numpy().astype(np.float64)
    assert _force_error(-gx, f) <= FORCE_RTOL
    params = _param_vector(w)
    idx = _param_indices(w.n_types)
    assert len(idx) == 22
    g_gpu = [_gpu_param_grad(grads, which, i) for which, i in idx]
    g_fd = [_fd_param(loss64, params, which, i) for which, i in idx]
    assert _rel_err(g_gpu, g_fd) <= GRAD_RTOL, (g_gpu, g_fd)


def test_full_size_trajectory_gradient_independent_of_checkpoint_interval():
    """200 steps of velocity Verlet on the 100,000-bead polymer/solvent system,
    backpropagated with checkpoints every 20 steps, every step and only at the
    start. Recomputing a segment from its checkpoint replays the forward, so
    the three sweeps differentiate the same trajectory: positions' and
    velocities' gradients and the FP64 parameter gradients agree to 1e-6
    relative (the lists of the recomputed frames may be built at different
    positions than in the forward, so FP64 pair sums may differ in order; a
    binary32 force then differs in its last bit on rare occasions)."""
    w = W.polymer_solvent()
    steps = 200
    rng = np.random.default_rng(31)
    cx = torch.tensor(rng.normal(size=(w.n, 3)), device="cuda")
    cv = torch.tensor(rng.normal(size=(w.n, 3)) * 0.1, device="cuda")
    out = {}
    for every in (20, 1, 200):
        with FrameModel(w) as m:  # each sweep from a fresh engine: the same forward
            p = m.initial_params()
            x0 = torch.tensor(w.positions, device="cuda", requires_grad=True)
            v0 = torch.tensor(w.velocities, device="cuda", requires_grad=True)
            xT, vT = m.trajectory(x0, v0, p, steps=steps, checkpoint_every=every)
            loss = (cx * xT.double()).sum() + (cv * vT.double()).sum()
            loss.backward()
            m.check()
            out[every] = (xT.detach().cpu().numpy(), x0.grad.cpu().numpy().astype(np.float64),
                          v0.grad.cpu().numpy().astype(np.float64),
                          np.concatenate([t.grad.double().cpu().numpy().ravel()
                                          for t in p.tensors()]))
            print(f"every {every}: loss {float(loss.detach()):.9e}", flush=True)
    ref = out[20]
    for every in (1, 200):
        got = out[every]
        np.testing.assert_array_equal(got[0], ref[0])  # the same forward trajectory
        for g, r in zip(got[1:], ref[1:]):
            assert _rel_err(g, r) <= 1e-6, every


# ------------------------------------------------ angles and dihedrals ----

def melt_small(seed=3):
    """480 beads of the melt's force field (bonds, 120-degree angles, n = 1
    dihedrals, exclusions to distance 2), jittered off the lattice so that no
    angle sits near 0 or 180 degrees."""
    return W.polymer_melt(n_chains=24, beads_per_chain=20, density=8.5, seed=seed, jitter=0.06)


def _with_listed(w, ak, at, dk):
    return w.with_(angle_k=np.asarray(ak, np.float32), angle_theta0=np.asarray(at, np.float32),
                   dihedral_k=np.asarray(dk, np.float32))


def _listed_vector(w):
    return [w.angle_k.astype(np.float32).copy(), w.angle_theta0.astype(np.float32).copy(),
            w.dihedral_k.astype(np.float32).copy()]


def test_listed_terms_force_matching_gradients_vs_fd():
    """Positions and the angle k, theta0 (per degree) and dihedral k gradients
    of a force-matching loss through harmonic angles and dihedrals, against
    central differences of the FP64 oracle's loss, at the adjoint bar."""
    w = melt_small()
    x0 = _frames(w, 1, seed=4, amp=0.02)[0]
    # Targets from parameters well away from the model's, so that dL/dF = 2 (F - F_target)
    # / (3n) is large against the binary32 forward's own rounding (its angles come
    # from a binary32 atan2, about 1e-7 rad).
    wt = _with_listed(w, w.angle_k * 1.3, w.angle_theta0 * 0.9, w.dihedral_k * 1.5)
    target = O.forces_f64(wt, x0.astype(np.float64))[0]

    def loss64(pos64, listed):
        f = O.forces_f64(_with_listed(w, *listed), pos64)[0]
        return np.sum((f - target) ** 2) / (w.n * 3)

    with FrameModel(w) as m:
        p = m.initial_params()
        assert p.angle_k is not None and p.dihedral_k is not None
        x = torch.tensor(x0, device="cuda", requires_grad=True)
        loss = force_matching_loss(m, [x], [torch.tensor(target, device="cuda")], p)
        loss.backward()
        gx = x.grad.cpu().numpy().astype(np.float64)
        g_par = [float(p.angle_k.grad[0]), float(p.angle_theta0.grad[0]),
                 float(p.dihedral_k.grad[0])]
    listed = _listed_vector(w)
    base = x0.astype(np.float64)
    # chain ends, second and middle beads of a few chains
    atoms = [0, 1, 2, 9, 10, 19, 20, 21, 200, 239, 240, int(w.n - 1)]
    g_fd, g_gpu = [], []
    for a in atoms:
        for c in range(3):
            h = 1e-5
            pp, pm = base.copy(), base.copy()
            pp[a, c] += h
            pm[a, c] -= h
            g_fd.append((loss64(pp, listed) - loss64(pm, listed)) / (2 * h))
            g_gpu.append(gx[a, c])
    assert _rel_err(g_gpu, g_fd) <= GRAD_RTOL, (g_gpu, g_fd)
    p_fd = [_fd_param(lambda q: loss64(base, q), listed, which, 0) for which in range(3)]
    assert _rel_err(g_par, p_fd) <= GRAD_RTOL, (g_par, p_fd)


def test_listed_terms_energy_gradients():
    """dE/dx = -F with angles and dihedrals, and dE/d(angle k, theta0,
    dihedral k) against central differences of the FP64 oracle's energy."""
    w = melt_small(seed=5)
    x0 = _frames(w, 1, seed=6, amp=0.02)[0]

    def energy64(listed):
        o = O.forces_f64(_with_listed(w, *listed), x0.astype(np.float64))[1]
        return o["pair_pe"] + o["bond_pe"] + o["angle_pe"] + o["dihedral_pe"]

    with FrameModel(w) as m:
        p = m.initial_params()
        x = torch.tensor(x0, device="cuda", requires_grad=True)
        f, e = m.forces(x, p)
        e.backward()
        gx = x.grad.cpu().numpy().astype(np.float64)
        f = f.detach().cpu().numpy().astype(np.float64)
        g_par = [float(p.angle_k.grad[0]), float(p.angle_theta0.grad[0]),
                 float(p.dihedral_k.grad[0])]
    assert _force_error(-gx, f) <= FORCE_RTOL
    listed = _listed_vector(w)
    p_fd = [_fd_param(energy64, listed, which, 0) for which in range(3)]
    assert _rel_err(g_par, p_fd) <= GRAD_RTOL, (g_par, p_fd)


def melt_two_types(seed=7):
    """melt_small with angles and dihedrals alternating between two types of
    different parameters; the second dihedral type has d = -1, n = 2."""
    w = melt_small(seed=seed)
    angles = w.angles.copy()
    angles[1::2, 3] = 2
    dihedrals = w.dihedrals.copy()
    dihedrals[1::2, 4] = 2
    return w.with_(angles=angles, dihedrals=dihedrals,
                   angle_k=np.array([25.0, 40.0], np.float32),
                   angle_theta0=np.array([120.0, 105.0], np.float32),
                   dihedral_k=np.array([1.0, 2.5], np.float32),
                   dihedral_d=np.array([1, -1], np.int32), dihedral_n=np.array([1, 2], np.int32))


def test_listed_terms_gradients_per_type():
    """Every angle k, theta0 and dihedral k gradient of two-type angles and
    dihedrals (one with d = -1, n = 2), for the energy and for a force-matching
    loss, against central differences of the FP64 oracle; dE/dx = -F."""
    w = melt_two_types()
    x0 = _frames(w, 1, seed=8, amp=0.02)[0]
    pos64 = x0.astype(np.float64)
    wt = _with_listed(w, w.angle_k * np.array([1.3, 0.8]), w.angle_theta0 * np.array([0.9, 1.1]),
                      w.dihedral_k * np.array([1.5, 0.6]))
    target = O.forces_f64(wt, pos64)[0]

    def energy64(listed):
        o = O.forces_f64(_with_listed(w, *listed), pos64)[1]
        return o["pair_pe"] + o["bond_pe"] + o["angle_pe"] + o["dihedral_pe"]

    def loss64(listed):
        f = O.forces_f64(_with_listed(w, *listed), pos64)[0]
        return np.sum((f - target) ** 2) / (w.n * 3)

    def grads(p):
        return [float(g) for t in (p.angle_k, p.angle_theta0, p.dihedral_k) for g in t.grad]

    with FrameModel(w) as m:
        assert m.n_angle_types == 2 and m.n_dihedral_types == 2
        p = m.initial_params()
        x = torch.tensor(x0, device="cuda", requires_grad=True)
        f, e = m.forces(x, p)
        e.backward()
        f = f.detach().cpu().numpy().astype(np.float64)
        assert _force_error(-x.grad.cpu().numpy().astype(np.float64), f) <= FORCE_RTOL
        assert _force_error(f, O.forces_f64(w, pos64)[0]) <= FORCE_RTOL
        g_energy = grads(p)
        p = m.initial_params()
        force_matching_loss(m, [torch.tensor(x0, device="cuda")],
                            [torch.tensor(target, device="cuda")], p).backward()
        g_loss = grads(p)
    listed = _listed_vector(w)
    slots = [(which, t) for which in range(3) for t in range(2)]
    e_fd = [_fd_param(energy64, listed, which, t) for which, t in slots]
    l_fd = [_fd_param(loss64, listed, which, t) for which, t in slots]
    assert _rel_err(g_energy, e_fd) <= GRAD_RTOL, (g_energy, e_fd)
    assert _rel_err(g_loss, l_fd) <= GRAD_RTOL, (g_loss, l_fd)
