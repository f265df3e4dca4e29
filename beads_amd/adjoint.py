"""Differentiable forces on the gfx950 engine (frame mode + adjoint kernel).

No reference counterpart: the reference computes forces only inside its
dynamics loop (native/src/simulation/simulation_run.cpp:212-235) and has no
gradients. This module puts the engine's frame-mode entry points
(``beads_engine_frame_forces`` / ``beads_engine_frame_backward``) behind
``torch.autograd.Function`` so that

* ``FrameModel.forces(x, params)`` returns forces and the total potential
  energy; gradients of any loss of them flow to the positions (J^T g - gE F,
  the Hessian-vector product of the adjoint kernels) and to the per-type-pair
  LJ epsilon/sigma, the per-type bond k/r0, angle k/theta0 (degrees) and
  dihedral k (FP64 accumulation on the device);
* ``FrameModel.trajectory(x0, v0, params, steps, checkpoint_every)`` runs
  velocity Verlet (NVE) with engine forces and backpropagates through it by a
  hand-written reverse sweep that recomputes each segment of ``checkpoint_every``
  steps from its stored (x, v) checkpoint.

Forces and positions are in the engine's real dtype as in the dynamics path
(float32, or float64 with ``precision="double"``, the BEADS_REAL64 build); the
adjoint kernel evaluates in FP64 from those coordinates. Engine calls are enqueued on torch's
current stream (the engine's stream-ordered frame entry points), so tensors
flow between torch and the engine without host waits; energies come back as
device tensors. A neighbour-list overflow surfaces at ``FrameModel.check()``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .capi import Engine
from .workloads import Workload


@dataclass
class Params:
    """Differentiable force-field parameters: epsilon/sigma [T, T] (symmetric),
    bond k/r0 [n_bond_types], angle k and theta0 in degrees [n_angle_types],
    dihedral k [n_dihedral_types] (None where the system has no such term)."""

    epsilon: torch.Tensor
    sigma: torch.Tensor
    bond_k: torch.Tensor
    bond_r0: torch.Tensor
    angle_k: torch.Tensor | None = None
    angle_theta0: torch.Tensor | None = None
    dihedral_k: torch.Tensor | None = None

    def slots(self) -> tuple[torch.Tensor | None, ...]:
        """All seven parameter slots in the adjoint's order, None where absent."""
        return (self.epsilon, self.sigma, self.bond_k, self.bond_r0, self.angle_k,
                self.angle_theta0, self.dihedral_k)

    def tensors(self) -> tuple[torch.Tensor, ...]:
        """The parameters the system has, in the same order."""
        return tuple(t for t in self.slots() if t is not None)


class FrameModel:
    """One topology (types, bonds, exclusions, box, cutoff) evaluated at
    caller-supplied positions. Positions are [n, 3] in the workload's slot
    order; the engine never reorders them."""

    def __init__(self, workload: Workload, device: str | torch.device = "cuda",
                 precision: str | None = None):
        self.workload = workload
        self.device = torch.device(device)
        self.engine = Engine(workload, precision)
        self.real = torch.float64 if self.engine.precision == "double" else torch.float32
        self.n = workload.n
        self.n_types = workload.n_types
        self.n_bond_types = int(workload.bond_k.shape[0])
        self.n_angle_types = int(workload.angle_k.shape[0]) if workload.angles.shape[0] else 0
        self.n_dihedral_types = (int(workload.dihedral_k.shape[0]) if workload.dihedrals.shape[0]
                                 else 0)
        self.n_params = self.engine.adjoint_param_count()
        self.masses = torch.as_tensor(workload.masses, dtype=self.real,
                                      device=self.device).reshape(-1, 1)
        self._params_key: bytes | None = None
        self._params_ident: tuple | None = None
        self._loaded = 0   # token of the frame whose neighbour list the engine holds
        self._tokens = 0
        self._stream = None  # the torch stream the engine enqueues on
        self.forward_calls = 0
        self.backward_calls = 0
        self.reload_calls = 0

    def close(self) -> None:
        self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- params

    def initial_params(self, requires_grad: bool = True) -> Params:
        w = self.workload

        def t(a):
            return torch.tensor(np.asarray(a), dtype=self.real,
                                device=self.device, requires_grad=requires_grad)

        return Params(t(w.lj_epsilon), t(w.lj_sigma), t(w.bond_k), t(w.bond_r0),
                      t(w.angle_k) if self.n_angle_types else None,
                      t(w.angle_theta0) if self.n_angle_types else None,
                      t(w.dihedral_k) if self.n_dihedral_types else None)

    def _set_params(self, *params) -> None:
        """Upload (eps, sigma, bond k, r0, angle k, theta0 [deg], dihedral k)."""
        params = tuple(params) + (None,) * (7 - len(params))
        present = [p for p in params if p is not None]
        # Same tensors, unmodified since the last upload: nothing to do (no
        # device-to-host copy).
        ident = tuple((p.data_ptr(), p._version, str(p.device), tuple(p.shape)) for p in present)
        if ident == self._params_ident:
            return
        arrs = [None if p is None else np.ascontiguousarray(p.detach().to("cpu", self.real).numpy())
                for p in params]
        key = b"".join(a.tobytes() for a in arrs if a is not None)
        if key != self._params_key:
            self.engine.set_lj_params(arrs[0], arrs[1])
            if self.n_bond_types:
                self.engine.set_bond_params(arrs[2], arrs[3])
            if self.n_angle_types and arrs[4] is not None and arrs[5] is not None:
                self.engine.set_angle_params(arrs[4], arrs[5])
            if self.n_dihedral_types and arrs[6] is not None:
                self.engine.set_dihedral_params(arrs[6])
            self._params_key = key
        self._params_ident = ident

    # ------------------------------------------------------------ raw calls

    def _on_current_stream(self) -> None:
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            self.engine.set_stream(s)
            self._stream = s

    def check(self) -> None:
        """Wait for the enqueued frame work; raises on a neighbour-list overflow."""
        self.engine.synchronize()

    def _forward_raw(self, x: torch.Tensor, reuse_list: bool = True):
        self._on_current_stream()
        f = torch.empty_like(x)
        e = torch.empty(4, dtype=torch.float64, device=self.device)
        self.engine.frame_forces_async(x.data_ptr(), f.data_ptr(), e.data_ptr(),
                                       reuse_list=reuse_list)
        self._tokens += 1
        self._loaded = self._tokens
        self.forward_calls += 1
        return f, e, self._tokens

    def _load_raw(self, x: torch.Tensor, token: int) -> None:
        """Make frame x (evaluated earlier under token) current for a backward."""
        if self._loaded == token:
            return
        self._on_current_stream()
        self.engine.frame_load_async(x.data_ptr(), reuse_list=True)
        self._loaded = token
        self.reload_calls += 1

    def _backward_raw(self, g_forces: torch.Tensor, g_energy: float):
        self._on_current_stream()
        gx = torch.empty((self.n, 3), dtype=self.real, device=self.device)
        gp = torch.empty(max(self.n_params, 1), dtype=torch.float64, device=self.device)
        self.engine.frame_backward_async(g_forces.data_ptr(), g_energy, gx.data_ptr(),
                                         gp.data_ptr())
        self.backward_calls += 1
        return gx, gp

    def _split(self, gp: torch.Tensor, like: Params | tuple) -> tuple[torch.Tensor | None, ...]:
        """The device gradient vector as (eps, sigma, bond k, r0, angle k,
        theta0 per degree, dihedral k), in the dtypes and devices of `like`
        (None where `like` has None)."""
        T, nb, na, nd = self.n_types, self.n_bond_types, self.n_angle_types, self.n_dihedral_types
        like = tuple(like.slots() if isinstance(like, Params) else like)
        like = like + (None,) * (7 - len(like))
        sizes = (T * T, T * T, nb, nb, na, na, nd)
        offs = np.concatenate([[0], np.cumsum(sizes)])
        parts = [gp[offs[q]: offs[q + 1]] for q in range(7)]
        parts[0] = parts[0].reshape(T, T)
        parts[1] = parts[1].reshape(T, T)
        parts[5] = parts[5] * (np.pi / 180.0)  # the kernel's gradient is per radian
        return tuple(None if q is None else p.to(q.dtype).to(q.device)
                     for p, q in zip(parts, like))

    def _positions(self, x: torch.Tensor) -> torch.Tensor:
        if x.shape != (self.n, 3):
            raise ValueError(f"positions must have shape ({self.n}, 3)")
        return x.detach().to(self.device, self.real).contiguous()

    # ------------------------------------------------------------- public

    def forces(self, x: torch.Tensor, params: Params) -> tuple[torch.Tensor, torch.Tensor]:
        """Forces [n, 3] (the engine's real dtype) and total potential energy (FP64
        scalar)."""
        return _FrameForces.apply(self, x, *params.slots())

    def trajectory(self, x0: torch.Tensor, v0: torch.Tensor, params: Params, steps: int,
                   checkpoint_every: int = 20) -> tuple[torch.Tensor, torch.Tensor]:
        """Velocity Verlet (NVE, the workload's dt) from (x0, v0); returns the
        final positions and velocities, differentiable w.r.t. x0, v0, params."""
        if steps < 1 or checkpoint_every < 1:
            raise ValueError("steps and checkpoint_every must be positive")
        return _VerletTrajectory.apply(self, int(steps), int(checkpoint_every), x0, v0,
                                       *params.slots())


def force_matching_value_and_grad(model: FrameModel, frames: list[torch.Tensor],
                                  targets: list[torch.Tensor],
                                  params: Params) -> tuple[torch.Tensor, Params]:
    """force_matching_loss and its parameter gradients without an autograd
    graph: each frame's adjoint pass runs right after its forward, while its
    neighbour list is current, on dL/dF = 2 (F - F_target) / (frames * n * 3)
    (the gradient autograd hands the same kernels). The per-frame FP64
    parameter gradients are summed in FP64 and returned in the parameters'
    dtypes. For training loops that need only dL/dtheta."""
    _check_frames(frames, targets)
    ps = tuple(None if t is None else t.detach() for t in params.slots())
    model._set_params(*ps)
    scale = 1.0 / (len(frames) * model.n * 3)
    total = torch.zeros((), dtype=torch.float64, device=model.device)
    g_sum = torch.zeros(max(model.n_params, 1), dtype=torch.float64, device=model.device)
    for x, ft in zip(frames, targets):
        f, _, _ = model._forward_raw(model._positions(x))
        diff = f.double() - ft.to(model.device).double()
        total += (diff * diff).sum()
        _, gp = model._backward_raw((2.0 * scale * diff).to(model.real).contiguous(), 0.0)
        g_sum += gp
    return total * scale, Params(*model._split(g_sum, ps))


def _check_frames(frames, targets) -> None:
    if len(frames) == 0 or len(frames) != len(targets):
        raise ValueError(f"force matching needs one target per frame and at least one frame; "
                         f"got {len(frames)} frames and {len(targets)} targets")


def force_matching_loss(model: FrameModel, frames: list[torch.Tensor],
                        targets: list[torch.Tensor], params: Params) -> torch.Tensor:
    """mean over frames, atoms and components of (F(x; params) - F_target)^2."""
    _check_frames(frames, targets)
    total = None
    for x, ft in zip(frames, targets):
        f, _ = model.forces(x, params)
        term = ((f.double() - ft.double()) ** 2).sum()
        total = term if total is None else total + term
    return total / (len(frames) * model.n * 3)


class _FrameForces(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model: FrameModel, x, *params):
        model._set_params(*params)
        pos = model._positions(x)
        f, e, token = model._forward_raw(pos)
        # An output that the loss does not use arrives as None, not as a zero
        # tensor whose value the backward would read back to the host.
        ctx.set_materialize_grads(False)
        ctx.model = model
        ctx.token = token
        ctx.x_dtype = x.dtype
        ctx.save_for_backward(pos, *params)
        return f, e.sum()

    @staticmethod
    def backward(ctx, g_forces, g_energy):
        model: FrameModel = ctx.model
        pos, *params = ctx.saved_tensors
        model._set_params(*params)
        model._load_raw(pos, ctx.token)
        gF = (torch.zeros_like(pos) if g_forces is None
              else g_forces.detach().to(model.device, model.real).contiguous())
        gE = 0.0 if g_energy is None else float(g_energy)
        gx, gp = model._backward_raw(gF, gE)
        return (None, gx.to(ctx.x_dtype)) + model._split(gp, params)


class _VerletTrajectory(torch.autograd.Function):
    """x_{n+1} = x_n + dt v_n + dt^2/(2m) F_n, v_{n+1} = v_n + dt/(2m) (F_n + F_{n+1}).

    Reverse sweep for L(x_T, v_T), a = dt/(2m), carry = adjoint of F_n from
    the later half-kick:
        Fbar_n  = carry + a vbar_n
        xbar_n += J(x_n)^T Fbar_n,   thetabar += (dF_n/dtheta)^T Fbar_n
        vhbar   = vbar_n + dt xbar_n
        xbar_{n-1} = xbar_n, vbar_{n-1} = vhbar, carry = a vhbar
    and at n = 0: xbar_0 += J(x_0)^T carry. Positions of one segment are
    recomputed from its checkpoint before its reverse pass."""

    @staticmethod
    def forward(ctx, model: FrameModel, steps, every, x0, v0, *params):
        model._set_params(*params)
        dt = float(model.workload.dt)
        a = (0.5 * dt) / model.masses
        x = model._positions(x0)
        v = v0.detach().to(model.device, model.real).contiguous()
        checkpoints = [(x.clone(), v.clone())]
        f, _, _ = model._forward_raw(x)
        for s in range(steps):
            v = v + a * f
            x = (x + dt * v).contiguous()
            f, _, _ = model._forward_raw(x)
            v = v + a * f
            if (s + 1) % every == 0 and s + 1 < steps:
                checkpoints.append((x.clone(), v.clone()))
        ctx.model = model
        ctx.steps, ctx.every = steps, every
        ctx.checkpoints = checkpoints
        ctx.dtypes = (x0.dtype, v0.dtype)
        ctx.save_for_backward(*params)
        return x.to(x0.dtype), v.to(v0.dtype)

    @staticmethod
    def backward(ctx, gx_T, gv_T):
        model: FrameModel = ctx.model
        params = ctx.saved_tensors
        model._set_params(*params)
        dt = float(model.workload.dt)
        a = (0.5 * dt) / model.masses
        steps, every = ctx.steps, ctx.every
        dev = model.device
        xbar = (torch.zeros((model.n, 3), dtype=model.real, device=dev) if gx_T is None
                else gx_T.detach().to(dev, model.real).clone())
        vbar = (torch.zeros((model.n, 3), dtype=model.real, device=dev) if gv_T is None
                else gv_T.detach().to(dev, model.real).clone())
        carry = torch.zeros_like(xbar)
        pbar = torch.zeros(max(model.n_params, 1), dtype=torch.float64, device=dev)
        for seg in range(len(ctx.checkpoints) - 1, -1, -1):
            start = seg * every
            end = min(start + every, steps)
            # Recompute positions x_start..x_end of this segment.
            x, v = ctx.checkpoints[seg]
            f, _, tok = model._forward_raw(x)
            xs, toks = [x], [tok]
            for _ in range(start, end):
                v = v + a * f
                x = (x + dt * v).contiguous()
                f, _, tok = model._forward_raw(x)
                v = v + a * f
                xs.append(x)
                toks.append(tok)
            # Reverse over steps end-1 .. start (frames end .. start+1).
            for n in range(end, start, -1):
                model._load_raw(xs[n - start], toks[n - start])
                fbar = (carry + a * vbar).contiguous()
                gx, gp = model._backward_raw(fbar, 0.0)
                xbar = xbar + gx
                pbar = pbar + gp
                vbar = vbar + dt * xbar
                carry = a * vbar
            if seg == 0:
                model._load_raw(xs[0], toks[0])
                gx, gp = model._backward_raw(carry.contiguous(), 0.0)
                xbar = xbar + gx
                pbar = pbar + gp
        return ((None, None, None, xbar.to(ctx.dtypes[0]), vbar.to(ctx.dtypes[1]))
                + model._split(pbar, params))
