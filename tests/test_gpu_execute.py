"""The drop-in boundary end to end on the GPU: ``Simulation.execute()`` through
``_native.execute_simulation`` writes the thermo CSV, the BEADS trajectory and
the state snapshot, and every file agrees with the engine handle run on the
same workload (same kernels, same step sequence) and with the FP64 oracle.

Bars: thermo values are the reference's binary32 values written with 9
significant digits, so they must equal the engine's FP64 statistics to the
binary32 rounding (2e-7 relative); the step-0 trajectory frame is a binary32
copy of the device state and must equal it bit for bit, later frames and the
snapshot within a few binary32 ulps (see below); step-0 forces meet the 1e-5
force bar against the oracle.
"""

from __future__ import annotations

import csv
import ctypes

import numpy as np
import pytest

import beads_amd as beads
import oracle as O
from beads_amd import capi
from beads_amd.io import read_beads_trajectory
from gpu_cases import cases

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-5


def _force_error(f_gpu, f_ref):
    scale = np.sqrt(np.mean(f_ref ** 2))
    return float(np.max(np.abs(f_gpu - f_ref) / np.maximum(np.abs(f_ref), scale)))


def _rows(path):
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def test_execute_writes_thermo_trajectory_and_snapshot(tmp_path):
    w = cases()["melt"].with_(steps=50)
    sim = w.to_simulation()
    sim.set_thermo(every=10, prefix=str(tmp_path / "run"))
    sim.set_trajectory(every=25, prefix=str(tmp_path / "run"),
                       fields=["tag", "type", "position", "image", "velocity", "force"])
    sim.save_final_state(prefix=str(tmp_path / "run"))
    sim.set_logging(echo="log", prefix=str(tmp_path / "run"))
    sim.execute()

    thermo = _rows(tmp_path / "run.thermo.csv")
    assert [int(r["step"]) for r in thermo] == [0, 10, 20, 30, 40, 50]
    assert list(thermo[0]) == ["step", "pe", "pair_pe", "bond_pe", "angle_pe", "dihedral_pe",
                               "ke", "temp", "press"]

    # The same run through the engine handle.
    with capi.Engine(w) as e:
        e.initialize(observables=True)
        s0 = e.stats()
        st0 = e.state()
        e.advance(50, observables_on_last=True)
        s50 = e.stats()
        st50 = e.state()
    # Thermo rows come after the step's Berendsen rescale (simulation_run.cpp:
    # 212-235); the handle's stats report the stored velocities with the
    # factor still pending, so step 50 scales the kinetic energy by lambda^2.
    L = (w.box_bound[1] - w.box_bound[0]).astype(np.float64)
    volume = float(np.prod(L))
    for row, s, lam in ((thermo[0], s0, 1.0), (thermo[-1], s50, s50.lambda_)):
        assert float(row["pair_pe"]) == pytest.approx(s.pe[0], rel=2e-7)
        assert float(row["bond_pe"]) == pytest.approx(s.pe[1], rel=2e-7)
        assert float(row["pe"]) == pytest.approx(sum(s.pe), rel=2e-7)
        ke = s.ke * lam * lam
        assert float(row["ke"]) == pytest.approx(ke, rel=1e-6)
        assert float(row["temp"]) == pytest.approx(s.temperature * lam * lam, rel=1e-6)
        press = 16.605390666 * (2.0 * ke + sum(s.virial)) / (3.0 * volume)
        assert float(row["press"]) == pytest.approx(press, rel=1e-5, abs=1e-4)

    traj = read_beads_trajectory(tmp_path / "run.trajectory.beadsbin")
    assert traj["steps"].tolist() == [0, 25, 50]
    rec0, rec2 = traj["records"][0], traj["records"][-1]
    o0, oe = np.argsort(rec0["tag"]), np.argsort(st0["tags"])
    np.testing.assert_array_equal(rec0["position"][o0], st0["positions"][oe])
    np.testing.assert_array_equal(rec0["force"][o0], st0["forces"][oe])
    f_ref, _ = O.forces(w, positions=st0["positions"], tags=st0["tags"], types=st0["types"])
    assert _force_error(st0["forces"].astype(np.float64), f_ref) <= FORCE_RTOL
    # Output steps run the force kernel with observables (energy terms kept),
    # which may round a force by an ulp differently from the variant without
    # them, exactly as the reference's two kernel variants may; later frames
    # are therefore held to a few binary32 ulps, not bit for bit.
    o2, oe2 = np.argsort(rec2["tag"]), np.argsort(st50["tags"])
    np.testing.assert_allclose(rec2["position"][o2], st50["positions"][oe2], rtol=0, atol=1e-5)
    np.testing.assert_array_equal(rec2["image"][o2], st50["images"][oe2])

    snap = beads.System.from_state(tmp_path / "run.state.beadsbin")
    tags = np.asarray(snap.tags)
    os_, oe3 = np.argsort(tags), np.argsort(st50["tags"])
    np.testing.assert_array_equal(np.asarray(snap.positions)[os_], rec2["position"][o2])
    np.testing.assert_allclose(np.asarray(snap.positions)[os_], st50["positions"][oe3],
                               rtol=0, atol=1e-5)
    # The snapshot carries the rescaled velocities (the factor is applied at
    # the end of the step); the handle still holds it pending.
    vel = st50["velocities"][oe3].astype(np.float64) * s50.lambda_
    np.testing.assert_allclose(np.asarray(snap.velocities)[os_], vel, rtol=1e-5,
                               atol=1e-5 * float(np.abs(vel).max()))
    np.testing.assert_array_equal(np.asarray(snap.images)[os_], st50["images"][oe3])
    log = (tmp_path / "run.log").read_text()
    assert "[BEADS]" in log


def test_interrupt_callback_stops_the_run():
    w = cases()["lj_regular"].with_(steps=100000)
    spec, keep = capi.make_spec(w)
    calls = []

    @ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)
    def stop(_ctx):
        calls.append(1)
        return 1 if len(calls) >= 2 else 0

    spec.interrupt = ctypes.cast(stop, ctypes.c_void_p).value
    lib = capi.library()
    status = lib.beads_execute_simulation(ctypes.byref(spec))
    assert status == 6  # BEADS_E_INTERRUPTED
    assert len(calls) == 2


def test_execute_with_step_graphs_writes_the_same_files(tmp_path, monkeypatch):
    """The run loop replays captured step graphs between output steps; the
    files must be byte-identical to a run that launches every step eagerly."""
    w = cases()["lj_regular"].with_(steps=100)
    blobs = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("BEADS_AMD_GRAPHS", graphs)
        sim = w.to_simulation()
        prefix = str(tmp_path / f"g{graphs}")
        sim.set_thermo(every=32, prefix=prefix)
        sim.set_trajectory(every=48, prefix=prefix, fields=["tag", "position", "velocity"])
        sim.save_final_state(prefix=prefix)
        sim.set_logging(echo="none")
        sim.execute()
        blobs[graphs] = [(tmp_path / f"g{graphs}.{ext}").read_bytes()
                         for ext in ("thermo.csv", "trajectory.beadsbin", "state.beadsbin")]
    rows = blobs["1"][0].decode().splitlines()
    assert [r.split(",")[0] for r in rows[1:]] == ["0", "32", "64", "96"]
    assert blobs["1"] == blobs["0"]


def test_async_outputs_match_synchronous_outputs(tmp_path, monkeypatch):
    """Outputs staged through pinned slots and written by the writer thread
    (the default) are byte-identical to outputs read back and written at each
    output step (BEADS_AMD_SYNC_OUTPUT=1): thermo CSV with all listed-term
    columns, every trajectory field, the snapshot and the THERMO log lines.
    Output steps every 3 or 4 steps keep all staging slots in flight."""
    w = cases()["melt"].with_(steps=60)
    blobs = {}
    for sync in ("0", "1"):
        monkeypatch.setenv("BEADS_AMD_SYNC_OUTPUT", sync)
        sim = w.to_simulation()
        prefix = str(tmp_path / f"s{sync}")
        sim.set_thermo(every=4, prefix=prefix)
        sim.set_trajectory(every=3, prefix=prefix,
                           fields=["tag", "type", "position", "image", "velocity", "force"])
        sim.save_final_state(prefix=prefix)
        sim.set_logging(echo="log", prefix=prefix)
        sim.execute()
        files = [(tmp_path / f"s{sync}.{ext}").read_bytes()
                 for ext in ("thermo.csv", "trajectory.beadsbin", "state.beadsbin")]
        log = (tmp_path / f"s{sync}.log").read_text().splitlines()
        blobs[sync] = files + [[ln for ln in log if ln.startswith("[BEADS] THERMO")]]
    rows = blobs["0"][0].decode().splitlines()
    assert "dihedral_pe" in rows[0]
    assert [r.split(",")[0] for r in rows[1:]] == [str(s) for s in range(0, 61, 4)]
    assert len(blobs["0"][3]) == 17  # the THERMO header and 16 rows
    traj = read_beads_trajectory(tmp_path / "s0.trajectory.beadsbin")
    assert [int(s) for s in traj["steps"]] == list(range(0, 61, 3))
    for k in range(4):
        assert blobs["0"][k] == blobs["1"][k], k


def test_async_outputs_report_a_list_overflow(tmp_path, monkeypatch):
    """A list overflow found by the writer thread fails the run with the same
    error as the synchronous path, and no row of the failing step is written."""
    w = cases()["lj_regular"].with_(steps=20, max_neighbors=8)
    for sync in ("0", "1"):
        monkeypatch.setenv("BEADS_AMD_SYNC_OUTPUT", sync)
        sim = w.to_simulation()
        sim.set_thermo(every=5, prefix=str(tmp_path / f"o{sync}"))
        sim.set_logging(echo="none")
        with pytest.raises(RuntimeError, match="overflow"):
            sim.execute()
        assert not (tmp_path / f"o{sync}.thermo.csv").exists()


def test_lammps_data_import_runs_end_to_end(tmp_path):
    """A molecular LAMMPS data file imported with ``beads.io.read_lammps_data``
    runs through ``Simulation.execute()``; the step-0 trajectory frame holds
    the file's positions (binary32) and velocities in tag order, and the dump
    export writes one LAMMPS frame per trajectory frame."""
    from test_lammps_data import _text

    data = tmp_path / "in.data"
    data.write_text(_text())
    system = beads.io.read_lammps_data(data, units="nm_kjmol")
    ff = beads.ForceField()
    ff.pair_style("lj", cutoff=0.9)
    ff.pair_coeff("lj", type_i=1, type_j=1, epsilon=1.0, sigma=0.34)
    ff.pair_coeff("lj", type_i=1, type_j=2, epsilon=0.8, sigma=0.36)
    ff.pair_coeff("lj", type_i=2, type_j=2, epsilon=0.6, sigma=0.38)
    ff.bond_style("harmonic")
    ff.bond_coeff("harmonic", type=1, k=500.0, r0=0.35)
    ff.angle_style("harmonic")
    ff.angle_coeff("harmonic", type=1, k=50.0, theta0=120.0)
    ff.dihedral_style("harmonic")
    ff.dihedral_coeff("harmonic", type=1, k=1.0, d=1, n=1)
    sim = beads.Simulation(system=system, forcefield=ff,
                           dynamics=beads.Dynamics("velocity_verlet", dt=0.001))
    sim.set_runsteps(10)
    sim.set_neighbor(cutoff_buffer=0.25, rebuild_check_every=1, sort_every_rebuild=2,
                     max_neighbors=64)
    prefix = str(tmp_path / "imp")
    sim.set_thermo(every=5, prefix=prefix)
    sim.set_trajectory(every=5, prefix=prefix,
                       fields=["tag", "type", "position", "image", "velocity", "force"])
    sim.set_logging(echo="none")
    sim.execute()
    rows = _rows(tmp_path / "imp.thermo.csv")
    assert [r["step"] for r in rows] == ["0", "5", "10"]
    traj = read_beads_trajectory(tmp_path / "imp.trajectory.beadsbin")
    rec = traj["records"][0]
    order = np.argsort(rec["tag"])
    np.testing.assert_array_equal(rec["position"][order], system.positions)
    np.testing.assert_array_equal(rec["velocity"][order], system.velocities)
    np.testing.assert_array_equal(rec["type"][order], system.types)
    dump = tmp_path / "imp.lammpstrj"
    beads.io.write_lammps_dump(tmp_path / "imp.trajectory.beadsbin", dump)
    assert dump.read_text().count("ITEM: TIMESTEP") == 3


def test_log_file_layout(tmp_path):
    """The text log keeps the reference's layout (log_output.cpp:19-402): six
    INFO blocks at run start with labels padded so that ': ' starts at column
    36, a THERMO header written once, one THERMO line per thermo row with
    every CSV value right-aligned in 16 columns, and the execute-end block.
    Expected strings are formatted here from the workload and the CSV."""
    w = cases()["melt"].with_(steps=20)
    sim = w.to_simulation()
    prefix = str(tmp_path / "lg")
    sim.set_thermo(every=10, prefix=prefix)
    sim.set_logging(echo="log", prefix=prefix)
    sim.execute()
    text = (tmp_path / "lg.log").read_text()
    blocks = text.split("\n\n")

    def lv(label, value):
        return f"  {label}".ljust(36) + ": " + str(value)

    lo, hi = w.box_bound.astype(np.float32)
    g3 = lambda v: "(" + ",".join(f"{float(x):g}" for x in v) + ")"
    assert blocks[0].splitlines() == [
        "[BEADS] INFO execute start", lv("units", "nm_kjmol"), lv("particles", w.n),
        lv("box_lo", g3(lo)), lv("box_hi", g3(hi)), lv("box_lengths", g3(hi - lo)),
        lv("runsteps", 20)]
    assert blocks[1].splitlines() == [
        "[BEADS] INFO dynamics summary", lv("style", "velocity_verlet"),
        lv("dt", f"{float(np.float32(w.dt)):.9g}"), lv("thermostat", "berendsen")]
    assert blocks[2].splitlines()[0] == "[BEADS] INFO neighbor summary"
    assert blocks[2].splitlines()[1] == lv("cutoff_buffer", f"{float(np.float32(w.cutoff_buffer)):g}")
    assert blocks[2].splitlines()[2] == lv("path", "regular_cell")
    assert blocks[3].splitlines() == [
        "[BEADS] INFO forcefield summary", lv("pair", "lj"), lv("bond", "harmonic"),
        lv("angle", "harmonic"), lv("dihedral", "harmonic"),
        lv("bonded_exclusion_policy", "explicit"), lv("bonded_exclusion_distance", 2),
        lv("bonded_excluded_pairs", len(O.exclusion_pairs(w.n, w.bonds, 2)))]
    assert blocks[4].splitlines() == [
        "[BEADS] INFO topology summary", lv("bonds", len(w.bonds)), lv("angles", len(w.angles)),
        lv("dihedrals", len(w.dihedrals))]
    # THERMO lines follow the last INFO block without a blank line.
    out = blocks[5].splitlines()
    assert out[:5] == [
        "[BEADS] INFO output summary", lv("thermo", "enabled"), lv("final_state", "disabled"),
        lv("trajectory", "disabled"), lv("log", "log")]
    thermo = out[5:]
    header = ["Step", "PotEng", "PairEng", "BondEng", "AngleEng", "DihedEng", "KinEng", "Temp",
              "Press"]
    assert thermo[0] == "[BEADS] THERMO " + "".join(h.rjust(16) for h in header)
    rows = (tmp_path / "lg.thermo.csv").read_text().splitlines()[1:]
    assert thermo[1:] == ["[BEADS] THERMO " + "".join(v.rjust(16) for v in r.split(","))
                          for r in rows]
    end = blocks[6].splitlines()
    assert end[0] == "[BEADS] INFO execute end" and end[1] == lv("runsteps", 20)
    assert [ln.split(":")[0] for ln in end[2:]] == [
        "  run_wall_s".ljust(36), "  steps_per_second".ljust(36),
        "  neighbor_build_count".ljust(36), "  dangerous_rebuild_count".ljust(36)]
    assert len(blocks) == 7
