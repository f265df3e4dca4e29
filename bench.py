"""Benchmark: atom-steps/s of the MI355X engine on the 1M-bead polymer melt.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload melt1m|fcc4000|mixture4m]

One step = one velocity-Verlet step of the whole system (pre-force half-kick,
drift, wrap, displacement check, predicated neighbour rebuild, fused LJ +
bonded forces, post-force half-kick, Berendsen factor). Inputs are resident in
HBM before the timed region; the timed region is bracketed by device
synchronisation. Prints one JSON line.

roofline: the fused force kernel (the dominant kernel) against HBM peak.
  algorithmic bytes per atom = 60 + 4 K + A_bonded (read x+type 16, count 4,
  indices 4 K, velocity 12, mass 4; write velocity 12, force 12; listed terms
  12 per bond + 16 per angle + 20 per dihedral, per atom), K the mean stored
  neighbour count; achieved = that x N / mean force-kernel duration, the
  duration measured with HIP events on the engine's stream.
cpu_baseline: the FP64 OpenMP host oracle on a bounded sample of the same
  workload (same system, fewer steps), on this host's cores.
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


def make_workload(name: str):
    from beads_amd import workloads as W

    if name == "melt1m":
        return W.polymer_melt(), "polymer melt, 20000 chains x 50 beads = 1,000,000 beads, nm_kjmol"
    if name == "fcc4000":
        return W.lj_fluid_fcc(), "LJ fluid, 4000 beads FCC, reduced units"
    if name == "mixture4m":
        return W.lj_mixture(), "4-type LJ mixture, 4,000,000 beads, nm_kjmol"
    raise SystemExit(f"unknown workload {name}")


def force_bytes_per_atom(w, mean_neighbors: float) -> float:
    bonded = (12.0 * len(w.bonds) + 16.0 * len(w.angles) + 20.0 * len(w.dihedrals)) / w.n
    return 60.0 + 4.0 * mean_neighbors + bonded


def cpu_baseline(w, sample_steps: int) -> dict:
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    _, st = O.run(w, steps=sample_steps, threads=threads)
    value = w.n * sample_steps / st.wall_seconds
    return {"value": value, "unit": "atom-steps/s", "cores": int(st.threads_used), "kind": "port",
            "sample": f"{w.name}: {sample_steps} steps of the same system, FP64 OpenMP oracle, "
                      f"{st.wall_seconds:.1f} s step loop"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="melt1m")
    ap.add_argument("--cpu-sample-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-phase-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        raise SystemExit(2)

    from beads_amd.capi import Engine

    w, description = make_workload(args.workload)
    print(f"bench: {description}; building device state", file=sys.stderr, flush=True)
    eng = Engine(w)
    eng.initialize(observables=False)
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
        per_atom = force_bytes_per_atom(w, k_mean)
        achieved = per_atom * w.n / force_avg / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "force_kernel", "bytes_per_atom": round(per_atom, 1),
                    "mean_launch_us": round(force_avg * 1e6, 2)}
        phases = {"integrate_us": round(pt.integrate_s / psteps * 1e6, 2),
                  "rebuild_us": round(pt.rebuild_s / psteps * 1e6, 2),
                  "force_us": round(pt.force_s / psteps * 1e6, 2)}
    builds = int(s1.build_count - s0.build_count)
    eng.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w, args.cpu_sample_steps)

    line = {
        "metric": "atom-steps/s at 1M beads",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": description, "n_beads": w.n, "steps": args.steps,
                   "neighbor_builds_in_timed_steps": builds,
                   "mean_neighbors": round(s1.mean_neighbors, 2), "phases": phases},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
