"""Python front end (CPU only): builder validation, the native spec dict, and
the snapshot / trajectory formats against independently written files.

Behaviour pinned here follows the reference front end (python/beads/*):
exception types for bad inputs, defaults, freezing on bind, the spec dict's
keys and dtypes, and the BEADS_SNAPSHOT / BEADS_TRAJECTORY v1 byte layouts
(final_state_output.cu:149-179, trajectory_output.cu:555-595).
"""

from __future__ import annotations

import struct

import numpy as np
import pytest

import beads_amd as beads
from beads_amd import workloads as W
from beads_amd._build_config import dtype_for

REAL = dtype_for("real_dtype")  # float32, or float64 with BEADS_AMD_PRECISION=double


def small_system(**kw):
    pos = np.array([[0.25, 0.25, 0.25], [0.75, 0.25, 0.25], [1.5, 1.5, 1.5]], np.float32)
    args = dict(positions=pos, types=[1, 1, 2], box_bound=[[0, 0, 0], [3, 3, 3]], units="nm_kjmol")
    args.update(kw)
    return beads.System(**args)


def test_public_surface():
    for name in ("System", "ForceField", "Dynamics", "Simulation", "build_info", "io"):
        assert hasattr(beads, name)
    assert beads.__version__
    for name in ("read_beads_trajectory", "read_lammps_data", "read_trajectory", "write_lammps_dump"):
        assert hasattr(beads.io, name)


def test_system_defaults_and_dtypes():
    s = small_system()
    assert s.n_particles == 3
    assert s.positions.dtype == REAL and s.types.dtype == np.int32
    assert s.tags.dtype == np.uint32 and list(s.tags) == [1, 2, 3]
    assert (s.masses == 1).all() and (s.velocities == 0).all() and (s.images == 0).all()
    assert list(s.molecule_ids) == [1, 1, 1]


@pytest.mark.parametrize("kw,exc", [
    ({"units": None}, TypeError),
    ({"units": "si"}, ValueError),
    ({"types": [1, 1, 3]}, ValueError),          # gap in dense types
    ({"types": [0, 1, 1]}, ValueError),
    ({"types": [1.0, 1.0, 1.0]}, TypeError),
    ({"box_bound": None}, TypeError),
    ({"box_bound": [[0, 0, 0], [3, 3, 0.1]]}, ValueError),   # positions outside
    ({"tags": [1, 1, 2]}, ValueError),
    ({"masses": [1.0, -1.0, 1.0]}, ValueError),
    ({"velocities": [[0, 0, 0]]}, ValueError),
    ({"molecule_ids": [1, 3, 3]}, ValueError),
])
def test_system_rejects_bad_inputs(kw, exc):
    with pytest.raises(exc):
        small_system(**kw)


def test_topology_canonical_forms_and_support():
    s = small_system()
    s.set_topology(bonds=[(2, 1, 1), (2, 3, 1)], angles=[(3, 2, 1, 1)])
    assert s.topology.bonds == ((1, 2, 1), (2, 3, 1))
    assert s.topology.angles == ((1, 2, 3, 1),)
    with pytest.raises(ValueError, match="supporting topology bonds"):
        small_system().set_topology(bonds=[(1, 2, 1)], angles=[(1, 2, 3, 1)])
    with pytest.raises(ValueError, match="duplicate"):
        small_system().set_topology(bonds=[(1, 2, 1), (2, 1, 1)])
    with pytest.raises(ValueError, match="unknown System tag"):
        small_system().set_topology(bonds=[(1, 9, 1)])


def forcefield(types=2):
    ff = beads.ForceField()
    ff.pair_style("lj", cutoff=0.9)
    for a in range(1, types + 1):
        for b in range(a, types + 1):
            ff.pair_coeff("lj", type_i=a, type_j=b, epsilon=1.0, sigma=0.34)
    return ff


def test_forcefield_coverage_and_freezing():
    s = small_system()
    ff = beads.ForceField()
    ff.pair_style("lj", cutoff=0.9)
    ff.pair_coeff("lj", type_i=1, type_j=1, epsilon=1.0, sigma=0.34)
    with pytest.raises(ValueError, match="missing"):
        beads.Simulation(system=s, forcefield=ff, dynamics=beads.Dynamics("velocity_verlet", dt=0.001))
    ff = forcefield()
    sim = beads.Simulation(system=small_system(), forcefield=ff,
                           dynamics=beads.Dynamics("velocity_verlet", dt=0.001))
    with pytest.raises(RuntimeError, match="frozen"):
        ff.pair_coeff("lj", type_i=1, type_j=2, epsilon=1.0, sigma=0.3)
    assert sim is not None
    with pytest.raises(ValueError, match="one active pair_style"):
        forcefield().pair_style("lj", cutoff=1.0)
    with pytest.raises(ValueError, match="already set"):
        forcefield().pair_coeff("lj", type_i=2, type_j=1, epsilon=1.0, sigma=0.3)


def test_dynamics_contract():
    with pytest.raises(ValueError):
        beads.Dynamics("langevin", dt=0.1)
    with pytest.raises(ValueError):
        beads.Dynamics("velocity_verlet", dt=0.0)
    with pytest.raises(ValueError):
        beads.Dynamics("velocity_verlet")
    with pytest.raises(TypeError):
        beads.Dynamics("velocity_verlet", dt=True)
    d = beads.Dynamics("velocity_verlet", dt=0.002)
    with pytest.raises(ValueError):
        d.set_thermostat("berendsen", temperature=300.0)
    with pytest.raises(ValueError):
        d.set_thermostat("berendsen", temperature=300.0, tau=0.0)
    d.set_thermostat("berendsen", temperature=300.0, tau=1.0)


def test_native_spec_dict_layout():
    w = W.polymer_melt(n_chains=3, beads_per_chain=8, seed=1)
    sim = w.to_simulation()
    sim.set_thermo(every=5, prefix="out")
    sim.set_trajectory(every=10, prefix="traj", fields=["tag", "position", "force"])
    sim.save_final_state(prefix="final")
    spec = sim._native_spec()
    assert set(spec) == {"system", "forcefield", "dynamics", "neighbor", "output", "runsteps"}
    sysd = spec["system"]
    assert sysd["positions"].dtype == REAL and sysd["positions"].flags.c_contiguous
    assert sysd["tags"].dtype == np.uint32 and sysd["images"].dtype == np.int32
    assert len(sysd["topology"]["bonds"]) == 3 * 7 and len(sysd["topology"]["dihedrals"]) == 3 * 5
    ff = spec["forcefield"]
    assert ff["pair_style"] == {"style": "lj", "params": {"cutoff": 1.2, "shift": False}}
    assert ff["bonded_exclusion_policy"] == "explicit" and ff["bonded_exclusion_distance"] == 2
    assert ff["angle_coeffs"][0]["params"] == {"k": 25.0, "theta0": 120.0}
    assert spec["dynamics"]["thermostat"]["style"] == "berendsen"
    assert spec["neighbor"] == {"cutoff_buffer": 0.3, "rebuild_check_every": 1,
                                "sort_every_rebuild": 10, "max_neighbors": 512}
    assert spec["output"]["trajectory"]["fields"] == ["tag", "position", "force"]
    assert spec["output"]["log"] == {"echo": "screen"}
    # The half-list option (no reference counterpart) only appears when chosen.
    sim.set_neighbor(half_list=True)
    assert sim._native_spec()["neighbor"]["half_list"] is True
    with pytest.raises(TypeError):
        sim.set_neighbor(half_list=1)


def test_default_exclusion_policy():
    s = small_system()
    s.set_topology(bonds=[(1, 2, 1)])
    ff = forcefield()
    ff.bond_style("harmonic")
    ff.bond_coeff("harmonic", type=1, k=100.0, r0=0.5)
    sim = beads.Simulation(system=s, forcefield=ff, dynamics=beads.Dynamics("velocity_verlet", dt=0.001))
    spec = sim._native_spec()["forcefield"]
    assert spec["bonded_exclusion_policy"] == "default" and spec["bonded_exclusion_distance"] == 2
    sim2 = beads.Simulation(system=small_system(), forcefield=forcefield(),
                            dynamics=beads.Dynamics("velocity_verlet", dt=0.001))
    assert sim2._native_spec()["forcefield"]["bonded_exclusion_policy"] == "none"


def test_output_settings_validation():
    sim = beads.Simulation(system=small_system(), forcefield=forcefield(),
                           dynamics=beads.Dynamics("velocity_verlet", dt=0.001))
    with pytest.raises(ValueError, match="canonical order"):
        sim.set_trajectory(every=1, prefix="t", fields=["tag", "force", "position"])
    with pytest.raises(ValueError, match="tag as the first"):
        sim.set_trajectory(every=1, prefix="t", fields=["position"])
    with pytest.raises(ValueError):
        sim.set_logging(echo="log")
    with pytest.raises(ValueError):
        sim.set_neighbor(cutoff_buffer=0.0, rebuild_check_every=5)
    with pytest.raises(ValueError):
        sim.set_thermo(every=0, prefix="x")
    sim.set_runsteps(0)


def _write_snapshot(path, units, step, lo, hi, tag, typ, mol, mass, pos, vel, img):
    with open(path, "wb") as f:
        f.write(b"BEADS_SNAPSHOT")
        f.write(struct.pack("=IIIIIII", 1, 4, 4, 4, 4, 4, 0x01020304))
        f.write(struct.pack("=Q", len(units)) + units.encode())
        f.write(struct.pack("=QQ", step, len(tag)))
        f.write(np.asarray(lo, np.float32).tobytes() + np.asarray(hi, np.float32).tobytes())
        for arr, dt in ((tag, np.uint32), (typ, np.int32), (mol, np.uint32), (mass, np.float32)):
            f.write(np.asarray(arr, dt).tobytes())
        for block, dt in ((pos, np.float32), (vel, np.float32), (img, np.int32)):
            for a in range(3):
                f.write(np.ascontiguousarray(np.asarray(block)[:, a], dt).tobytes())


def test_snapshot_reader_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 7
    pos = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    img = rng.integers(-3, 3, (n, 3)).astype(np.int32)
    tag = rng.permutation(n).astype(np.uint32) + 1
    p = tmp_path / "x.state.beadsbin"
    _write_snapshot(p, "reduced", 42, [0, 0, 0], [2, 2, 2], tag, np.ones(n), np.arange(1, n + 1),
                    np.full(n, 2.0), pos, vel, img)
    s = beads.System.from_state(p)
    assert s.units == "reduced"
    np.testing.assert_array_equal(s.positions, pos)
    np.testing.assert_array_equal(s.velocities, vel)
    np.testing.assert_array_equal(s.images, img)
    np.testing.assert_array_equal(s.tags, tag)
    bad = tmp_path / "bad.state.beadsbin"
    bad.write_bytes(p.read_bytes() + b"\0")
    with pytest.raises(ValueError, match="trailing"):
        beads.System.from_state(bad)


def _write_trajectory(p, endian="<", n=4, frames=3, marker=0x01020304, version=1):
    e = endian
    rec = np.dtype([("tag", e + "u4"), ("position", e + "f4", (3,)), ("image", e + "i4", (3,))])
    with open(p, "wb") as f:
        f.write(b"BEADS_TRAJECTORY")
        f.write(struct.pack(e + "7I", version, marker, 4, 4, 4, 4, 4))
        f.write(struct.pack(e + "Q", 8) + b"nm_kjmol")
        f.write(struct.pack(e + "IQQ", 1 | 4 | 8, n, rec.itemsize))
        for k in range(frames):
            f.write(struct.pack(e + "Q", 10 * k))
            f.write(np.array([0, 0, 0, 5, 5, 5], e + "f4").tobytes())
            r = np.zeros(n, rec)
            r["tag"] = [4, 3, 2, 1]
            r["position"] = k + np.arange(12, dtype=np.float32).reshape(4, 3) / 10
            r["image"][:, 0] = k
            f.write(r.tobytes())


def test_trajectory_reader_either_byte_order(tmp_path):
    """The marker read back in either order names the file's byte order
    (python/beads/io.py:795-802); both files give the same values."""
    out = {}
    for e in "<>":
        p = tmp_path / f"t{'le' if e == '<' else 'be'}.trajectory.beadsbin"
        _write_trajectory(p, e)
        out[e] = beads.io.read_beads_trajectory(p)
        assert out[e]["endian"] == e
    for key in ("steps", "box_bounds"):
        np.testing.assert_array_equal(out["<"][key], out[">"][key])
    for field in ("tag", "position", "image"):
        np.testing.assert_array_equal(out["<"]["records"][field], out[">"]["records"][field])
    bad = tmp_path / "bad.trajectory.beadsbin"
    _write_trajectory(bad, "<", marker=0x05060708)
    with pytest.raises(ValueError, match="Unsupported BEADS trajectory endian marker"):
        beads.io.read_beads_trajectory(bad)
    _write_trajectory(bad, ">", version=2)
    with pytest.raises(ValueError, match="Unsupported BEADS trajectory version: 2"):
        beads.io.read_beads_trajectory(bad)
    good = tmp_path / "tle.trajectory.beadsbin"
    bad.write_bytes(good.read_bytes()[:-3])
    with pytest.raises(ValueError, match="file size is inconsistent"):
        beads.io.read_beads_trajectory(bad)


def test_trajectory_reader_and_lammps_dump(tmp_path):
    n, frames = 4, 3
    p = tmp_path / "t.trajectory.beadsbin"
    _write_trajectory(p, "=", n, frames)
    t = beads.io.read_beads_trajectory(p)
    assert t["fields"] == ["tag", "position", "image"] and t["frame_count"] == frames
    assert list(t["steps"]) == [0, 10, 20] and t["units"] == "nm_kjmol"
    assert t["records"]["position"][2, 0, 0] == pytest.approx(2.0)
    out = tmp_path / "t.lammpstrj"
    beads.io.write_lammps_dump(p, out)
    lines = out.read_text().splitlines()
    assert lines[0] == "ITEM: TIMESTEP" and lines[8].startswith("ITEM: ATOMS id x y z ix iy iz")
    assert lines[9].split()[0] == "1"  # sorted by tag


def test_native_parser_accepts_a_full_spec(tmp_path):
    """Every field the builders emit parses in _native (explicit exclusion
    distance, thermostat, all outputs); without a device the run itself stops
    at the first HIP call, with RuntimeError, never a parse error."""
    w = W.polymer_melt(n_chains=40, beads_per_chain=50, seed=3).with_(steps=5)
    sim = w.to_simulation()
    sim.set_thermo(every=1, prefix=str(tmp_path / "run"))
    sim.set_trajectory(every=5, prefix=str(tmp_path / "run"),
                       fields=["tag", "type", "position", "image", "velocity", "force"])
    sim.save_final_state(prefix=str(tmp_path / "run"))
    sim.set_logging(echo="none")
    sim.set_neighbor(half_list=True)
    from beads_amd.simulation import load_native

    try:
        load_native().execute_simulation(sim._native_spec())
    except RuntimeError as exc:
        assert "HIP" in str(exc) or "device" in str(exc), exc


def test_force_matching_rejects_mismatched_or_empty_frames():
    import torch

    from beads_amd.adjoint import force_matching_loss, force_matching_value_and_grad

    x = torch.zeros((4, 3))
    for fn in (force_matching_loss, force_matching_value_and_grad):
        with pytest.raises(ValueError, match="one target per frame"):
            fn(None, [x, x], [x], None)
        with pytest.raises(ValueError, match="at least one frame"):
            fn(None, [], [], None)
