"""File readers/writers: BEADS binary trajectories and LAMMPS custom dumps.

``read_beads_trajectory`` parses the BEADS_TRAJECTORY v1 layout written by the
engine (trajectory_output.cu:33-46, 148-250, 555-595; reference reader
python/beads/io.py:233-275, 755-940) into numpy structured arrays.
``write_lammps_dump`` converts one to LAMMPS ``dump custom`` text
(python/beads/io.py:278-298, 943-1028). ``read_lammps_data`` imports a
molecular LAMMPS data file as a ``System`` (python/beads/io.py:97-230).
"""

from __future__ import annotations

import operator
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
    endian: str  # "<" or ">", from the marker
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


# Sized numpy kinds of the header's element sizes; the messages are the
# reference reader's (python/beads/io.py:895-918).
def _sized(kind: str, size: int, endian: str, label: str) -> np.dtype:
    if size not in (4, 8):
        raise ValueError(f"Unsupported BEADS trajectory {label} size: {size}.")
    return np.dtype(f"{endian}{kind}{size}")


def _read_header(handle) -> _Header:
    """Header of a BEADS_TRAJECTORY v1 file in either byte order: the marker
    0x01020304 read back in little- or big-endian order names the order of
    every later field (the reference reader's rule, python/beads/io.py:788-893)."""
    def take(n: int, label: str) -> bytes:
        data = handle.read(n)
        if len(data) != n:
            raise ValueError(f"Trajectory file is truncated while reading {label}.")
        return data

    magic = take(len(_MAGIC), "magic")
    if magic != _MAGIC:
        raise ValueError(f"Unsupported BEADS trajectory magic: {magic!r}.")
    version_raw = take(4, "version")
    marker_raw = take(4, "endian marker")
    endian = next((e for e in "<>" if struct.unpack(e + "I", marker_raw)[0] == _ENDIAN), None)
    if endian is None:
        raise ValueError("Unsupported BEADS trajectory endian marker.")
    (version,) = struct.unpack(endian + "I", version_raw)
    if version != 1:
        raise ValueError(f"Unsupported BEADS trajectory version: {version}. "
                         f"This reader supports version 1.")
    sizes = []
    for label in ("real size", "index size", "image size", "type size", "tag size"):
        sizes.append(struct.unpack(endian + "I", take(4, label))[0])
    real_size, index_size, image_size, type_size, tag_size = sizes
    (units_len,) = struct.unpack(endian + "Q", take(8, "units length"))
    if units_len > 4096:
        raise ValueError("BEADS trajectory units string is unreasonably long.")
    units = take(units_len, "units").decode("ascii")
    (mask,) = struct.unpack(endian + "I", take(4, "field mask"))
    (n,) = struct.unpack(endian + "Q", take(8, "particle count"))
    (record_bytes,) = struct.unpack(endian + "Q", take(8, "record bytes"))
    if mask & ~_KNOWN:
        raise ValueError("BEADS trajectory field mask contains unsupported bits.")
    if not mask & 1:
        raise ValueError("BEADS trajectory field mask must include tag.")
    if n <= 0:
        raise ValueError("BEADS trajectory particle count must be positive.")
    real = _sized("f", real_size, endian, "real")
    parts = [("tag", _sized("u", tag_size, endian, "tag"))]
    image = _sized("i", image_size, endian, "image")
    kinds = {"type": _sized("i", type_size, endian, "type"), "position": (real, (3,)),
             "image": (image, (3,)), "velocity": (real, (3,)), "force": (real, (3,))}
    fields = tuple(name for name, bit in _FIELD_BITS if mask & bit)
    for name in fields[1:]:
        kind = kinds[name]
        parts.append((name, *kind) if isinstance(kind, tuple) else (name, kind))
    record = np.dtype(parts)
    if record.itemsize != record_bytes:
        raise ValueError("BEADS trajectory record byte count is inconsistent with its field mask.")
    frame = np.dtype([("step", np.dtype(endian + "u8")), ("box_bound", real, (2, 3)),
                      ("records", record, (n,))])
    start = handle.tell()
    handle.seek(0, os.SEEK_END)
    payload = handle.tell() - start
    handle.seek(start)
    if payload % frame.itemsize:
        raise ValueError("BEADS trajectory file size is inconsistent with its header.")
    return _Header(version, endian, real_size, index_size, image_size, type_size, tag_size, units,
                   mask, fields, int(n), int(record_bytes), payload // frame.itemsize,
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
    if isinstance(precision, (bool, np.bool_)):
        raise TypeError("write_lammps_dump(...) precision must be an integer, not bool.")
    try:
        precision = int(operator.index(precision))
    except TypeError as exc:
        raise TypeError("write_lammps_dump(...) precision must be an integer.") from exc
    if not 1 <= precision <= 17:
        raise ValueError("write_lammps_dump(...) precision must be between 1 and 17.")
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
