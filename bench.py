"""Benchmark: atom-steps/s of the MI355X engine on the 1M-bead polymer melt.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload melt1m|fcc4000|mixture4m]

One step = one velocity-Verlet step of the whole system (pre-force half-kick,
drift, wrap, displacement check, predicated neighbour rebuild, fused LJ +
bonded forces, post-force half-kick, Berendsen factor). Inputs are resident in
HBM before the timed region; the timed region is bracketed by device
synchronisation. Before the warm-up, --equilibrate untimed steps (default 300)
take the lattice-built melt past its relaxation, so the timed steps see the
steady rebuild rate (the equilibration's own rate is reported beside it).
Prints one JSON line.

roofline: the fused force kernel (the dominant kernel) against HBM peak.
  algorithmic bytes per atom = 60 + 4 K + A_bonded (read x+type 16, count 4,
  indices 4 K, velocity 12, mass 4; write velocity 12, force 12; listed terms
  12 per bond + 16 per angle + 20 per dihedral, per atom), K the mean stored
  neighbour count; every real is twice as wide with BEADS_AMD_PRECISION=double
  (the BEADS_REAL64 library): 116 + 4 K + 2 A_bonded. achieved = that x N /
  mean force-kernel duration, the duration measured with HIP events on the
  engine's stream.
cpu_baseline: the FP64 OpenMP host oracle on a bounded sample of the same
  workload (same system, fewer steps), on this host's cores.

batched: 512 reduced-unit LJ replicas (N log-uniform in [500, 20000]) advanced
  by one persistent hipGraph-captured launch (workgroup size from the
  autotuner) against the same replicas run one after another on the plain
  single-system path; both timed by the host clock around synchronised runs.
adjoint: the 100,000-bead polymer/solvent force-matching workload (64 seeded
  frames, MSE to seeded target forces) forward vs forward+backward, plus a
  200-step velocity-Verlet trajectory backprop with checkpoints every 20 steps.
other_workloads: the 4,000-bead FCC LJ fluid (1000 steps) and the 4,000,000-bead
  four-type mixture (200 steps), each with its own FP64 oracle sample.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Bytes per real of the engine library the run uses (BEADS_AMD_PRECISION).
REAL_SIZE = 8 if os.environ.get("BEADS_AMD_PRECISION", "single").strip().lower() == "double" else 4


def make_workload(name: str):
    from beads_amd import workloads as W

    if name == "melt1m":
        return W.polymer_melt(), "polymer melt, 20000 chains x 50 beads = 1,000,000 beads, nm_kjmol"
    if name == "fcc4000":
        return W.lj_fluid_fcc(), "LJ fluid, 4000 beads FCC, reduced units"
    if name == "mixture4m":
        return W.lj_mixture(), "4-type LJ mixture, 4,000,000 beads, nm_kjmol"
    raise SystemExit(f"unknown workload {name}")


def force_bytes_per_atom(w, mean_neighbors: float, real_size: int = 4) -> float:
    scale = real_size / 4.0
    bonded = (12.0 * len(w.bonds) + 16.0 * len(w.angles) + 20.0 * len(w.dihedrals)) / w.n
    return 56.0 * scale + 4.0 + 4.0 * mean_neighbors + bonded * scale


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, sample_steps: int) -> dict:
    # Thread placement for the OpenMP runtime, which reads it when the oracle
    # library is first loaded (nothing in this process loads it earlier).
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    _, st = O.run(w, steps=sample_steps, threads=threads)
    value = w.n * sample_steps / st.wall_seconds
    return {"value": value, "unit": "atom-steps/s", "cores": int(st.threads_used), "kind": "port",
            "sample": f"{w.name}: {sample_steps} steps of the same system, FP64 OpenMP oracle, "
                      f"{st.wall_seconds:.1f} s step loop",
            "cpu_model": cpu_model(), "omp_proc_bind": os.environ["OMP_PROC_BIND"],
            "omp_places": os.environ["OMP_PLACES"]}


def batched_leg(steps: int, serial_steps: int, count: int, melts: bool = False) -> dict:
    """Batched replicas vs serial plain runs of the same replicas: the LJ
    replica set, or (melts) bead-spring melts with bonds, angles, dihedrals
    and exclusions."""
    from beads_amd import workloads as W
    from beads_amd.capi import Batch, Engine

    ws = W.melt_replica_set(count) if melts else W.replica_set(count)
    atoms = sum(w.n for w in ws)
    kind = "melt" if melts else "LJ"
    print(f"bench: batched leg, {count} {kind} replicas, {atoms} atoms", file=sys.stderr,
          flush=True)
    with Batch(ws) as b:
        chosen, trial = b.autotune(trial_steps=20)
        config = b.config_info(chosen)
        configs = b.configs()
        b.run(0)
        b.synchronize()
        t0 = time.perf_counter()
        dev = b.profile(steps)
        b.synchronize()
        wall = time.perf_counter() - t0
    batched = atoms * steps / wall
    # Serial: the plain path, one replica at a time, each warmed up first.
    serial_wall = 0.0
    for w in ws:
        with Engine(w) as e:
            e.initialize(observables=False)
            # two 16-step graphs: the step graph is captured, instantiated and
            # replayed before the timed steps
            e.profile(32)
            e.synchronize()
            t0 = time.perf_counter()
            e.profile(serial_steps)
            e.synchronize()
            serial_wall += time.perf_counter() - t0
    serial = atoms * serial_steps / serial_wall
    if melts:
        sizes = [w.n for w in ws]
        label = (f"{count} bead-spring melts (bonds, angles, dihedrals, exclusions to distance 2),"
                 f" N {min(sizes)}-{max(sizes)}, {atoms} beads, nm_kjmol, Berendsen 300 K")
    else:
        label = (f"{count} LJ replicas, N log-uniform [500, 20000], {atoms} atoms, "
                 f"reduced units, Berendsen")
    return {"workload": label,
            "steps": steps, "config": config,
            "autotune_s": [dict(c, seconds=round(x, 5)) for c, x in zip(configs, trial)],
            "batched_atom_steps_per_s": batched, "batched_device_s": dev,
            "batched_replica_steps_per_s": count * steps / wall,
            "serial_atom_steps_per_s": serial, "serial_steps_per_replica": serial_steps,
            "serial_replica_steps_per_s": count * serial_steps / serial_wall,
            "speedup": batched / serial}


def adjoint_leg(frames: int, traj_steps: int, every: int) -> dict:
    """Force-matching gradient and trajectory backprop cost vs the forward pass."""
    import numpy as np
    import torch

    from beads_amd import workloads as W
    from beads_amd.adjoint import FrameModel, force_matching_loss, force_matching_value_and_grad

    w = W.polymer_solvent()
    print(f"bench: adjoint leg, {w.n} beads, {frames} frames", file=sys.stderr, flush=True)
    rng = np.random.default_rng(123)
    lo = w.box_bound[0].astype(np.float64)
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    xs = []
    for _ in range(frames):
        p = w.positions.astype(np.float64) + rng.uniform(-0.03, 0.03, (w.n, 3))
        xs.append(torch.tensor((lo + np.mod(p - lo, L)).astype(np.float32), device="cuda"))
    with FrameModel(w) as m:
        p = m.initial_params()
        # Seeded targets: forces at parameters 4 % away from the model's.
        with torch.no_grad():
            true = [t.detach() * torch.tensor(rng.uniform(0.96, 1.04, tuple(t.shape)),
                                               dtype=torch.float32, device="cuda")
                    for t in p.tensors()]
            true[0] = torch.triu(true[0]) + torch.triu(true[0], 1).T
            true[1] = torch.triu(true[1]) + torch.triu(true[1], 1).T
            from beads_amd.adjoint import Params
            tp = Params(*true)
            targets = [m.forces(x, tp)[0].clone() for x in xs]
        # warm-up
        force_matching_loss(m, xs[:2], targets[:2], p).backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            force_matching_loss(m, xs, targets, p)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        fwd_dev = bwd_dev = 0.0
        calls = 0
        for x in xs[:8]:
            f, _ = m.forces(x, p)
            fd, _ = m.engine.frame_timing()
            (f.double() ** 2).sum().backward()
            _, bd = m.engine.frame_timing()
            fwd_dev += fd
            bwd_dev += bd
            calls += 1
        for t in p.tensors():
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = force_matching_loss(m, xs, targets, p)
        loss.backward()
        torch.cuda.synchronize()
        t_fb = time.perf_counter() - t0
        grad_norm = float(sum((t.grad.double() ** 2).sum() for t in p.tensors()) ** 0.5)
        # the same loss and parameter gradients without an autograd graph
        force_matching_value_and_grad(m, xs[:2], targets[:2], p)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fused_loss, _ = force_matching_value_and_grad(m, xs, targets, p)
        torch.cuda.synchronize()
        t_fused = time.perf_counter() - t0
        # trajectory backprop, after one short warm-up trajectory
        x0 = torch.tensor(w.positions, device="cuda", requires_grad=True)
        v0 = torch.tensor(w.velocities, device="cuda", requires_grad=True)
        xw, vw = m.trajectory(x0, v0, p, steps=every, checkpoint_every=every)
        ((xw.double() ** 2).sum() * 1e-3 + (vw.double() ** 2).sum()).backward()
        for t in (x0, v0, *p.tensors()):
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xT, vT = m.trajectory(x0, v0, p, steps=traj_steps, checkpoint_every=every)
        torch.cuda.synchronize()
        t_traj_f = time.perf_counter() - t0
        t0 = time.perf_counter()
        ((xT.double() ** 2).sum() * 1e-3 + (vT.double() ** 2).sum()).backward()
        torch.cuda.synchronize()
        t_traj_b = time.perf_counter() - t0
    return {"workload": f"polymer/solvent {w.n} beads (500 x 100 chains + 50,000 solvent), "
                        f"4 types, 22 parameters, nm_kjmol",
            "frames": frames, "loss": float(loss.detach()), "grad_norm": grad_norm,
            "forward_s": t_fwd, "forward_backward_s": t_fb,
            "gradient_over_forward": t_fb / t_fwd,
            "backward_over_forward": (t_fb - t_fwd) / t_fwd,
            "fused_value_and_grad_s": t_fused,
            "fused_gradient_over_forward": t_fused / t_fwd,
            "fused_loss_rel_diff": abs(float(fused_loss) - float(loss.detach()))
            / abs(float(loss.detach())),
            "kernel_backward_over_forward": bwd_dev / fwd_dev,
            "frame_forward_device_us": fwd_dev / calls * 1e6,
            "frame_backward_device_us": bwd_dev / calls * 1e6,
            "trajectory": {"steps": traj_steps, "checkpoint_every": every,
                           "forward_s": t_traj_f, "backward_s": t_traj_b,
                           "backward_over_forward": t_traj_b / t_traj_f}}


def adjoint_melt_leg(frames: int) -> dict:
    """Force-matching gradient cost on the 1M-bead melt's force field (LJ,
    bonds, angles, dihedrals, exclusions; 7 parameters): per-frame kernel time
    of the backward against the forward, and autograd forward + backward
    against the forward."""
    import numpy as np
    import torch

    from beads_amd import workloads as W
    from beads_amd.adjoint import FrameModel, Params, force_matching_loss

    w = W.polymer_melt()
    print(f"bench: adjoint melt leg, {w.n} beads, {frames} frames", file=sys.stderr, flush=True)
    rng = np.random.default_rng(321)
    lo = w.box_bound[0].astype(np.float64)
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    xs = []
    for _ in range(frames):
        q = w.positions.astype(np.float64) + rng.uniform(-0.01, 0.01, (w.n, 3))
        xs.append(torch.tensor((lo + np.mod(q - lo, L)).astype(np.float32), device="cuda"))
    with FrameModel(w) as m:
        p = m.initial_params()
        with torch.no_grad():
            tp = Params(*[t.detach() * float(rng.uniform(0.96, 1.04)) for t in p.tensors()])
            targets = [m.forces(x, tp)[0].clone() for x in xs]
        force_matching_loss(m, xs[:1], targets[:1], p).backward()
        torch.cuda.synchronize()
        fwd_dev = bwd_dev = 0.0
        for x in xs:
            f, _ = m.forces(x, p)
            fd, _ = m.engine.frame_timing()
            (f.double() ** 2).sum().backward()
            _, bd = m.engine.frame_timing()
            fwd_dev += fd
            bwd_dev += bd
        for t in p.tensors():
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.no_grad():
            force_matching_loss(m, xs, targets, p)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss = force_matching_loss(m, xs, targets, p)
        loss.backward()
        torch.cuda.synchronize()
        t_fb = time.perf_counter() - t0
    return {"workload": f"polymer melt {w.n} beads (20,000 x 50), LJ + harmonic bonds, angles, "
                        f"dihedrals, exclusions to distance 2; 7 parameters; frames jittered "
                        f"by up to 0.01 nm",
            "frames": frames, "loss": float(loss.detach()),
            "frame_forward_device_us": fwd_dev / frames * 1e6,
            "frame_backward_device_us": bwd_dev / frames * 1e6,
            "kernel_backward_over_forward": bwd_dev / fwd_dev,
            "forward_s": t_fwd, "forward_backward_s": t_fb,
            "gradient_over_forward": t_fb / t_fwd}


def measured_traffic(workload: str, mean_neighbors: float):
    """HBM bytes per force-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy calibration) and the
    mean neighbour count that profile ran at. The bytes are reported only when
    that count is within 2 % of this run's (the traffic scales with it);
    otherwise None."""
    path = ROOT / "profiles" / f"force_kernel_{workload}_hbm_traffic.md"
    if not path.exists():
        return None, None
    if REAL_SIZE != 4:
        return None, {"file": str(path.relative_to(ROOT)), "note": "profiled on the binary32 build"}
    traffic, k = None, None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            traffic = int(line.split(":", 1)[1])
        if line.startswith("profiled_mean_neighbors:"):
            k = float(line.split(":", 1)[1])
    note = {"file": str(path.relative_to(ROOT)), "profiled_mean_neighbors": k}
    if traffic is None or k is None or abs(k - mean_neighbors) > 0.02 * mean_neighbors:
        return None, note
    return traffic, note


def other_workload_leg(name: str, steps: int, cpu_steps: int) -> dict:
    """A further benchmark configuration: atom-steps/s of `steps` steps after a
    32-step warm-up (which captures and replays the step graph) and the FP64
    oracle on `cpu_steps` steps of the same system."""
    from beads_amd.capi import Engine

    w, description = make_workload(name)
    print(f"bench: {description}", file=sys.stderr, flush=True)
    with Engine(w) as e:
        e.initialize(observables=False)
        e.profile(32)
        e.synchronize()
        t0 = time.perf_counter()
        e.profile(steps)
        e.synchronize()
        wall = time.perf_counter() - t0
        st = e.stats()
    out = {"workload": description, "n_beads": w.n, "steps": steps,
           "atom_steps_per_s": w.n * steps / wall, "ms_per_step": wall / steps * 1e3,
           "neighbor_builds": int(st.build_count), "reorders": int(st.reorder_count),
           "mean_neighbors": round(st.mean_neighbors, 2),
           "timing": "steps 33 to steps + 32 of a run from the initial state"}
    if cpu_steps:
        out["cpu_baseline"] = cpu_baseline(w, cpu_steps)
        out["gpu_over_cpu"] = out["atom_steps_per_s"] / out["cpu_baseline"]["value"]
    return out


def copy_calibration(mib: int) -> None:
    """A plain copy of known size: each launch reads and writes mib MiB."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": n * 4, "copy_bytes_written_per_launch": n * 4,
                      "launches": 5}))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="melt1m")
    ap.add_argument("--cpu-sample-steps", type=int, default=64)
    ap.add_argument("--equilibrate", type=int, default=300,
                    help="untimed steps run before the warm-up, so that the timed steps see "
                         "the steady rebuild rate rather than the relaxation of the lattice "
                         "start")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="skip the 4,000-bead FCC fluid and 4M-bead mixture lines")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--serial-steps", type=int, default=100)
    ap.add_argument("--replicas", type=int, default=512,
                    help="LJ replicas of the batched leg (0: skip it)")
    ap.add_argument("--melt-replicas", type=int, default=128,
                    help="bead-spring melt replicas of the batched-melt leg (0: skip it)")
    ap.add_argument("--adjoint-frames", type=int, default=64)
    ap.add_argument("--traj-steps", type=int, default=200)
    ap.add_argument("--half-list", action="store_true",
                    help="store each LJ pair once (reaction by FP64 atomics)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-phase-profile", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        raise SystemExit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    from beads_amd.capi import Engine

    w, description = make_workload(args.workload)
    if args.half_list:
        w = w.with_(half_list=True)
        description += ", half neighbour list"
    print(f"bench: {description}; building device state", file=sys.stderr, flush=True)
    t_setup = time.perf_counter()
    eng = Engine(w)  # host topology + host-to-device copies of the whole state
    eng.synchronize()
    setup_s = time.perf_counter() - t_setup
    eng.initialize(observables=False)
    eq = None
    if args.equilibrate:
        t_eq = time.perf_counter()
        eng.profile(args.equilibrate)
        eng.synchronize()
        t_eq = time.perf_counter() - t_eq
        eq = {"steps": args.equilibrate, "atom_steps_per_s": w.n * args.equilibrate / t_eq,
              "neighbor_builds": int(eng.stats().build_count)}
    eng.profile(args.warmup)
    eng.synchronize()
    s0 = eng.stats()

    t0 = time.perf_counter()
    times = eng.profile(args.steps, per_phase=False)
    eng.synchronize()
    wall = time.perf_counter() - t0
    s1 = eng.stats()
    value = w.n * args.steps / wall
    print(f"bench: {args.steps} steps in {wall:.3f} s (device {times.total_s:.3f} s)",
          file=sys.stderr, flush=True)

    roofline = None
    phases = None
    if not args.no_phase_profile:
        psteps = min(args.steps, 200)
        pt = eng.profile(psteps, per_phase=True)
        k_mean = eng.stats().mean_neighbors
        force_avg = pt.force_s / max(1, pt.force_launches)
        per_atom = force_bytes_per_atom(w, k_mean, REAL_SIZE)
        achieved = per_atom * w.n / force_avg / 1e9
        traffic, traffic_src = measured_traffic(args.workload, k_mean)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "traffic_profile": traffic_src,
                    "kernel": "force_kernel", "bytes_per_atom": round(per_atom, 1),
                    "algorithmic_bytes_per_launch": round(per_atom * w.n),
                    "mean_launch_us": round(force_avg * 1e6, 2)}
        phases = {"integrate_us": round(pt.integrate_s / psteps * 1e6, 2),
                  "rebuild_us": round(pt.rebuild_s / psteps * 1e6, 2),
                  "force_us": round(pt.force_s / psteps * 1e6, 2)}
    builds = int(s1.build_count - s0.build_count)
    eng.close()

    batched = None
    if not args.no_batch and args.replicas > 0:
        batched = batched_leg(args.batch_steps, args.serial_steps, args.replicas)
    batched_melts = None
    if not args.no_batch and args.melt_replicas > 0:
        batched_melts = batched_leg(args.batch_steps, args.serial_steps, args.melt_replicas,
                                    melts=True)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(args.adjoint_frames, args.traj_steps, 20)
        adjoint["melt"] = adjoint_melt_leg(8)

    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, args.cpu_sample_steps)

    others = None
    if not args.no_other_workloads and args.workload == "melt1m":
        cpu_on = not args.no_cpu_baseline
        others = {"fcc4000": other_workload_leg("fcc4000", 1000, 1000 if cpu_on else 0),
                  "mixture4m": other_workload_leg("mixture4m", 200, 8 if cpu_on else 0)}

    line = {
        "metric": "atom-steps/s at 1M beads; batched vs serial replicas/s; adjoint/forward cost",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if REAL_SIZE == 8 else "f32",
        "data": "synthetic",
        "config": {"workload": description, "n_beads": w.n, "steps": args.steps,
                   "neighbor_builds_in_timed_steps": builds,
                   "mean_neighbors": round(s1.mean_neighbors, 2), "phases": phases,
                   "equilibration": eq,
                   "setup_s": round(setup_s, 3),
                   "host_inclusive_atom_steps_per_s": w.n * args.steps / (wall + setup_s)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_melts": batched_melts,
        "adjoint": adjoint,
        "other_workloads": others,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
