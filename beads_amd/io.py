"""File readers/writers: BEADS binary trajectories and LAMMPS custom dumps.

``read_beads_trajectory`` parses the BEADS_TRAJECTORY v1 layout written by the
engine (trajectory_output.cu:33-46, 148-250, 555-595; reference reader
python/beads/io.py:233-275, 755-940) into numpy structured arrays.
``write_lammps_dump`` converts one to LAMMPS ``dump custom`` text
(python/beads/io.py:278-298, 943-1028). ``read_lammps_data`` imports a
molecular LAMMPS data file as a ``System`` (python/beads/io.py:97-230).
"""

from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

from ._lammps_data import read_lammps_data

__all__ = ["read_beads_trajectory", "read_lammps_data", "read_trajectory", "write_lammps_dump"]

_MAGIC = b"BEADS_TRAJECTORY"
_ENDIAN = 0x01020304
_FIELD_BITS = (("tag", 1), ("type", 2), ("position", 4), ("image", 8), ("velocity", 16), ("force", 32))
_KNOWN = sum(bit for _, bit in _FIELD_BITS)


@dataclass(frozen=True)
class _Header:
    version: int
    endian: int
    real_size: int
    index_size: int
    image_size: int
    type_size: int
    tag_size: int
    units: str
    field_mask: int
    fields: tuple[str, ...]
    n_particles: int
    record_bytes: int
    frame_count: int
    record_dtype: np.dtype
    frame_dtype: np.dtype


def _int_dtype(size: int, signed: bool, label: str) -> np.dtype:
    table = {(4, True): np.int32, (8, True): np.int64, (4, False): np.uint32, (8, False): np.uint64}
    if (size, signed) not in table:
        raise ValueError(f"Trajectory {label} size {size} is not supported.")
    return np.dtype(table[(size, signed)])


def _read_header(handle) -> _Header:
    def take(n: int, label: str) -> bytes:
        data = handle.read(n)
        if len(data) != n:
            raise ValueError(f"Trajectory header ended while reading {label}.")
        return data

    if take(len(_MAGIC), "magic") != _MAGIC:
        raise ValueError("Not a BEADS trajectory file (magic mismatch).")
    version, endian, real_size, index_size, image_size, type_size, tag_size = struct.unpack(
        "=7I", take(28, "sizes"))
    if version != 1:
        raise ValueError(f"Trajectory version {version} is not supported.")
    if endian != _ENDIAN:
        raise ValueError("Trajectory endian marker is not supported.")
    (units_len,) = struct.unpack("=Q", take(8, "units length"))
    units = take(units_len, "units").decode("ascii")
    (mask,) = struct.unpack("=I", take(4, "field mask"))
    n, record_bytes = struct.unpack("=QQ", take(16, "counts"))
    if mask & ~_KNOWN or not mask & 1:
        raise ValueError("Trajectory field mask is not supported.")
    if real_size not in (4, 8):
        raise ValueError(f"Trajectory real size {real_size} is not supported.")
    real = np.dtype(np.float32 if real_size == 4 else np.float64)
    parts = [("tag", _int_dtype(tag_size, False, "tag"))]
    fields = ["tag"]
    if mask & 2:
        parts.append(("type", _int_dtype(type_size, True, "type")))
        fields.append("type")
    if mask & 4:
        parts.append(("position", real, (3,)))
        fields.append("position")
    if mask & 8:
        parts.append(("image", _int_dtype(image_size, True, "image"), (3,)))
        fields.append("image")
    if mask & 16:
        parts.append(("velocity", real, (3,)))
        fields.append("velocity")
    if mask & 32:
        parts.append(("force", real, (3,)))
        fields.append("force")
    record = np.dtype(parts)
    if record.itemsize != record_bytes:
        raise ValueError("Trajectory record size does not match its field mask.")
    frame = np.dtype([("step", np.uint64), ("box_bound", real, (2, 3)), ("records", record, (n,))])
    start = handle.tell()
    handle.seek(0, os.SEEK_END)
    payload = handle.tell() - start
    handle.seek(start)
    if frame.itemsize == 0 or payload % frame.itemsize:
        raise ValueError("Trajectory payload is not a whole number of frames.")
    return _Header(version, endian, real_size, index_size, image_size, type_size, tag_size, units,
                   mask, tuple(fields), int(n), int(record_bytes), payload // frame.itemsize,
                   record, frame)


def read_beads_trajectory(path: object) -> dict[str, object]:
    """Read a whole trajectory. ``records`` keeps the writer's slot order per
    frame; join frames on the ``tag`` field."""
    with open(os.fspath(path), "rb") as handle:
        h = _read_header(handle)
        frames = np.fromfile(handle, dtype=h.frame_dtype, count=h.frame_count)
    if frames.shape[0] != h.frame_count:
        raise ValueError("Trajectory payload ended before all frames were read.")
    boxes = np.array(frames["box_bound"], copy=True)
    return {
        "version": h.version, "endian": h.endian, "real_size": h.real_size,
        "index_size": h.index_size, "image_size": h.image_size, "type_size": h.type_size,
        "tag_size": h.tag_size, "units": h.units, "field_mask": h.field_mask,
        "fields": list(h.fields), "n_particles": h.n_particles, "particle_count": h.n_particles,
        "record_bytes": h.record_bytes, "frame_count": h.frame_count,
        "steps": np.array(frames["step"], copy=True), "box_bounds": boxes,
        "box_lo": boxes[:, 0, :], "box_hi": boxes[:, 1, :],
        "records": np.array(frames["records"], copy=True),
    }


def read_trajectory(path: object) -> dict[str, object]:
    """Alias of ``read_beads_trajectory``."""
    return read_beads_trajectory(path)


def write_lammps_dump(trajectory_path: object, dump_path: object, *, sort_by_tag: bool = True,
                      precision: int = 17) -> None:
    """Convert a BEADS trajectory to LAMMPS ``dump custom`` text."""
    if not isinstance(sort_by_tag, bool):
        raise TypeError("write_lammps_dump(...) sort_by_tag must be a bool.")
    if isinstance(precision, bool) or not isinstance(precision, int) or not 1 <= precision <= 17:
        raise ValueError("write_lammps_dump(...) precision must be an integer in 1..17.")
    traj = read_beads_trajectory(trajectory_path)
    fields = traj["fields"]
    columns = ["id"]
    if "type" in fields:
        columns.append("type")
    if "position" in fields:
        columns += ["x", "y", "z"]
    if "image" in fields:
        columns += ["ix", "iy", "iz"]
    if "velocity" in fields:
        columns += ["vx", "vy", "vz"]
    if "force" in fields:
        columns += ["fx", "fy", "fz"]
    fmt = f"{{:.{precision}g}}"
    with open(os.fspath(dump_path), "w", encoding="utf-8") as out:
        for k in range(traj["frame_count"]):
            rec = traj["records"][k]
            if sort_by_tag:
                rec = rec[np.argsort(rec["tag"], kind="stable")]
            lo, hi = traj["box_lo"][k], traj["box_hi"][k]
            out.write(f"ITEM: TIMESTEP\n{int(traj['steps'][k])}\n")
            out.write(f"ITEM: NUMBER OF ATOMS\n{traj['n_particles']}\n")
            out.write("ITEM: BOX BOUNDS pp pp pp\n")
            for a in range(3):
                out.write(f"{fmt.format(float(lo[a]))} {fmt.format(float(hi[a]))}\n")
            out.write("ITEM: ATOMS " + " ".join(columns) + "\n")
            for r in rec:
                row = [str(int(r["tag"]))]
                if "type" in fields:
                    row.append(str(int(r["type"])))
                for name in ("position", "image", "velocity", "force"):
                    if name in fields:
                        vals = r[name]
                        if name == "image":
                            row += [str(int(v)) for v in vals]
                        else:
                            row += [fmt.format(float(v)) for v in vals]
                out.write(" ".join(row) + "\n")
