"""``System``: particle arrays, box, units and topology; snapshot restart.

Mirrors python/beads/_system.py:25-167 (constructor contract, defaults,
freezing on bind) and python/beads/_state_snapshot.py:20-120 (the
BEADS_SNAPSHOT v1 reader behind ``System.from_state``).
"""

from __future__ import annotations

import os
import struct
import sys

import numpy as np

from ._checks import IMAGE, INDEX, REAL, TYPE, integer_array, real_array, rows_of_three
from ._topology import EMPTY, Topology, make_topology, recheck

UNITS = ("reduced", "nm_kjmol")


def _positions(value: object) -> np.ndarray:
    arr = real_array("positions", value)
    if arr.ndim == 1 and arr.shape[0] == 0:
        raise ValueError("positions must contain at least one particle")
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise ValueError("positions must have shape (n_particles, 3)")
    if arr.shape[0] == 0:
        raise ValueError("positions must contain at least one particle")
    return arr


def _box(value: object | None) -> np.ndarray:
    if value is None:
        raise TypeError("System requires box_bound=")
    box = real_array("box_bound", value, (2, 3))
    if np.any(box[1] - box[0] <= 0.0):
        raise ValueError("box_bound upper row must be greater than lower row")
    return box


def _inside(positions: np.ndarray, box: np.ndarray) -> None:
    if np.any(positions < box[0]) or np.any(positions >= box[1]):
        raise ValueError("positions must be inside box_bound")


def _dense_types(arr: np.ndarray) -> None:
    if arr.size == 0:
        raise ValueError("types must be dense one-based ids without gaps")
    if np.any(arr < 1):
        raise ValueError("types must be one-based positive integers")
    present = np.unique(arr)
    if present[-1] != present.size:
        raise ValueError("types must be dense one-based ids without gaps")


def _permutation(name: str, arr: np.ndarray, n: int) -> None:
    if arr.size != n or np.any(arr < 1) or np.any(arr > n) or np.unique(arr).size != n:
        raise ValueError(f"{name} must be a permutation of 1..N")


def _dense_ids(name: str, arr: np.ndarray, n: int) -> None:
    if arr.size == 0:
        raise ValueError(f"{name} must form a contiguous 1-based set starting at 1")
    if np.any(arr < 1):
        raise ValueError(f"{name} must be one-based positive integers")
    present = np.unique(arr)
    if present[-1] > n or present[-1] != present.size:
        raise ValueError(f"{name} must form a contiguous 1-based set starting at 1")


def _owned(arr: np.ndarray, dtype, readonly: bool = False) -> np.ndarray:
    out = np.array(arr, dtype=dtype, copy=True, order="C")
    if readonly:
        out.setflags(write=False)
    return out


class System:
    """Particle state of one simulation. Frozen once bound to a Simulation."""

    __slots__ = ("positions", "box_bound", "types", "velocities", "masses", "tags",
                 "molecule_ids", "images", "units", "topology", "_frozen")

    def __init__(self, *, positions: object, types: object, box_bound: object | None = None,
                 velocities: object | None = None, masses: object | None = None,
                 tags: object | None = None, molecule_ids: object | None = None,
                 images: object | None = None, units: str | None = None) -> None:
        if units is None:
            raise TypeError('System requires explicit units="reduced" or units="nm_kjmol".')
        if not isinstance(units, str):
            raise TypeError("units must be a string")
        if units not in UNITS:
            raise ValueError('units must be "reduced" or "nm_kjmol"')
        pos = _positions(positions)
        n = pos.shape[0]
        box = _box(box_bound)
        _inside(pos, box)
        typ = integer_array("types", types, (n,), dtype=TYPE)
        if np.any(typ < 1):
            raise ValueError("types must be one-based positive integers")
        _dense_types(typ)
        vel = np.zeros_like(pos) if velocities is None else rows_of_three("velocities", velocities, n)
        mass = np.ones(n, dtype=REAL) if masses is None else real_array("masses", masses, (n,))
        if np.any(mass <= 0.0):
            raise ValueError("masses must be finite and positive")
        if tags is None:
            tag = np.arange(1, n + 1, dtype=INDEX)
        else:
            tag = integer_array("tags", tags, (n,), dtype=INDEX)
            _permutation("tags", tag, n)
        if molecule_ids is None:
            mol = np.ones(n, dtype=INDEX)
        else:
            mol = integer_array("molecule_ids", molecule_ids, (n,), dtype=INDEX)
            _dense_ids("molecule_ids", mol, n)
        img = (np.zeros((n, 3), dtype=IMAGE) if images is None
               else integer_array("images", images, (n, 3), dtype=IMAGE))
        object.__setattr__(self, "positions", _owned(pos, REAL))
        object.__setattr__(self, "box_bound", _owned(box, REAL))
        object.__setattr__(self, "types", _owned(typ, TYPE))
        object.__setattr__(self, "velocities", _owned(vel, REAL))
        object.__setattr__(self, "masses", _owned(mass, REAL))
        object.__setattr__(self, "tags", _owned(tag, INDEX))
        object.__setattr__(self, "molecule_ids", _owned(mol, INDEX))
        object.__setattr__(self, "images", _owned(img, IMAGE))
        object.__setattr__(self, "units", units)
        object.__setattr__(self, "topology", EMPTY)
        object.__setattr__(self, "_frozen", False)

    def __setattr__(self, name, value):
        raise AttributeError("System fields are read-only; build a new System instead.")

    @property
    def n_particles(self) -> int:
        return int(self.positions.shape[0])

    @classmethod
    def from_state(cls, path: object) -> "System":
        return cls(**read_snapshot(path))

    def set_topology(self, *, bonds: object | None = None, angles: object | None = None,
                     dihedrals: object | None = None) -> "System":
        self._require_mutable()
        topo = make_topology(bonds=bonds, angles=angles, dihedrals=dihedrals,
                             tags=self.tags.tolist())
        object.__setattr__(self, "topology", topo)
        return self

    def _require_mutable(self) -> None:
        if self._frozen:
            raise RuntimeError("System is frozen after binding to a Simulation.")

    def _freeze(self) -> "System":
        # Re-check the owned state (python/beads/_system.py:251-300) so arrays
        # mutated in place after construction are caught before the run.
        n = self.n_particles
        if not np.isfinite(self.positions).all():
            raise ValueError("positions must be finite")
        _inside(self.positions, self.box_bound)
        if not np.isfinite(self.velocities).all():
            raise ValueError("velocities must be finite")
        if not np.isfinite(self.masses).all() or np.any(self.masses <= 0.0):
            raise ValueError("masses must be finite and positive")
        _dense_types(self.types)
        _permutation("tags", self.tags, n)
        _dense_ids("molecule_ids", self.molecule_ids, n)
        for name, dtype in (("positions", REAL), ("box_bound", REAL), ("types", TYPE),
                            ("velocities", REAL), ("masses", REAL), ("tags", INDEX),
                            ("molecule_ids", INDEX), ("images", IMAGE)):
            object.__setattr__(self, name, _owned(getattr(self, name), dtype, readonly=True))
        topo: Topology = self.topology
        object.__setattr__(self, "topology", recheck(topo, self.tags))
        object.__setattr__(self, "_frozen", True)
        return self


# ---------------------------------------------------------------------------
# BEADS_SNAPSHOT v1 (final_state_output.cu:149-179; reader _state_snapshot.py:44-120)

_SNAP_MAGIC = b"BEADS_SNAPSHOT"
_ENDIAN = 0x01020304


class _Reader:
    def __init__(self, stream, what: str):
        self.stream = stream
        self.what = what

    def exact(self, count: int, label: str) -> bytes:
        data = self.stream.read(count)
        if len(data) != count:
            raise ValueError(f"{self.what} is truncated while reading {label}")
        return data

    def u32(self, label: str) -> int:
        return struct.unpack("=I", self.exact(4, label))[0]

    def u64(self, label: str) -> int:
        return struct.unpack("=Q", self.exact(8, label))[0]

    def array(self, dtype, count: int, label: str) -> np.ndarray:
        dt = np.dtype(dtype)
        return np.frombuffer(self.exact(count * dt.itemsize, label), dtype=dt).copy()


def _path_string(path: object, what: str) -> str:
    try:
        value = os.fspath(path)
    except TypeError as exc:
        raise TypeError(f"{what} path must be a string or path-like object") from exc
    if not isinstance(value, str):
        raise TypeError(f"{what} path must be a string or path-like object")
    if not value:
        raise ValueError(f"{what} path must not be empty")
    return value


def read_snapshot(path: object) -> dict:
    """Parse a BEADS_SNAPSHOT v1 file into ``System`` keyword arguments."""
    name = _path_string(path, "state snapshot")
    try:
        stream = open(name, "rb")
    except OSError as exc:
        raise OSError(f"failed to open state snapshot file: {name}") from exc
    with stream:
        r = _Reader(stream, "state snapshot")
        if r.exact(len(_SNAP_MAGIC), "magic") != _SNAP_MAGIC:
            raise ValueError("state snapshot magic is not supported")
        version = r.u32("version")
        if version != 1:
            raise ValueError(f"state snapshot version {version} is not supported")
        real_size = r.u32("real size")
        sizes = {label: r.u32(f"{label} size") for label in ("index", "image", "type", "tag")}
        if r.u32("endian marker") != _ENDIAN:
            raise ValueError("state snapshot endian marker is not supported")
        if real_size not in (4, 8):
            raise ValueError(f"state snapshot real size {real_size} is not supported; expected 4 or 8")
        for label, size in sizes.items():
            if size != 4:
                raise ValueError(f"state snapshot {label} size {size} is not supported; expected 4")
        units_len = r.u64("units length")
        if units_len > sys.maxsize:
            raise ValueError("state snapshot units exceeds supported host size")
        try:
            units = r.exact(units_len, "units").decode("ascii")
        except UnicodeDecodeError as exc:
            raise ValueError("state snapshot units are not ASCII") from exc
        if units not in UNITS:
            raise ValueError(f"state snapshot units {units!r} are not supported")
        r.u64("step")  # kept by the format, not part of System
        n = r.u64("particle count")
        if n == 0:
            raise ValueError("state snapshot particle count must be positive")
        real = np.float32 if real_size == 4 else np.float64
        lo = r.array(real, 3, "box lower")
        hi = r.array(real, 3, "box upper")
        tags = r.array(np.uint32, n, "tag")
        types = r.array(np.int32, n, "type")
        mol = r.array(np.uint32, n, "molecule_id")
        mass = r.array(real, n, "mass")
        pos = np.stack([r.array(real, n, f"position_{a}") for a in "xyz"], axis=1)
        vel = np.stack([r.array(real, n, f"velocity_{a}") for a in "xyz"], axis=1)
        img = np.stack([r.array(np.int32, n, f"image_{a}") for a in "xyz"], axis=1)
        if stream.read(1):
            raise ValueError("state snapshot contains trailing bytes")
    return {
        "units": units,
        "box_bound": np.stack([lo, hi]),
        "positions": pos,
        "velocities": vel,
        "masses": mass,
        "types": types,
        "tags": tags,
        "molecule_ids": mol,
        "images": img,
    }
