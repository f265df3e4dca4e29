"""Tag-based topology records and their canonical forms.

Canonical forms follow python/beads/_topology.py:53-129: bonds as (min, max),
angles with the outer tags ordered (i < k), dihedrals as the smaller of the
forward and reversed tag tuples; duplicates and repeated tags are rejected,
types must be dense 1-based, and every angle/dihedral edge must be a bond
(python/beads/_topology.py:203-224).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterable

import numpy as np

from ._checks import positive_index, positive_type_id


@dataclass(frozen=True)
class Topology:
    bonds: tuple[tuple[int, int, int], ...]
    angles: tuple[tuple[int, int, int, int], ...]
    dihedrals: tuple[tuple[int, int, int, int, int], ...]
    # The same records as int64 arrays when they were checked in bulk, so a
    # re-check (System freeze) needs no conversion back from tuples.
    arrays: tuple | None = field(default=None, compare=False, repr=False)


EMPTY = Topology((), (), ())


def _records(name: str, records: object | None) -> tuple:
    if records is None:
        return ()
    if isinstance(records, (str, bytes)):
        raise TypeError(f"{name} must be an iterable of topology records")
    try:
        return tuple(records)  # type: ignore[arg-type]
    except TypeError as exc:
        raise TypeError(f"{name} must be an iterable of topology records") from exc


def _fields(name: str, index: int, record: object, width: int) -> tuple:
    if isinstance(record, (str, bytes)):
        raise TypeError(f"{name}[{index}] must be a topology record")
    try:
        values = tuple(record)  # type: ignore[arg-type]
    except TypeError as exc:
        raise TypeError(f"{name}[{index}] must be a topology record") from exc
    if len(values) != width:
        raise ValueError(f"{name}[{index}] must contain exactly {width} values")
    return values


def _tag(name: str, index: int, pos: int, value: object, known: frozenset[int]) -> int:
    tag = positive_index(f"{name}[{index}][{pos}]", value)
    if tag not in known:
        raise ValueError(f"{name}[{index}] references unknown System tag {tag}")
    return tag


def _dense(label: str, types: Iterable[int]) -> None:
    present = sorted(set(types))
    if present and present != list(range(1, len(present) + 1)):
        raise ValueError(f"{label} topology types must be dense one-based ids without gaps")


def _canonical(name: str, records: object | None, width: int, known: frozenset[int]):
    out = []
    seen = set()
    n_tags = width - 1
    for index, record in enumerate(_records(name, records)):
        values = _fields(name, index, record, width)
        tags = tuple(_tag(name, index, k, values[k], known) for k in range(n_tags))
        type_id = positive_type_id(f"{name}[{index}] type", values[-1])
        if len(set(tags)) != n_tags:
            words = {2: "distinct tags", 3: "three distinct tags", 4: "four distinct tags"}
            raise ValueError(f"{name} entries must reference {words[n_tags]}")
        if n_tags == 2:
            key = (min(tags), max(tags))
        elif n_tags == 3:
            key = tags if tags[0] < tags[2] else (tags[2], tags[1], tags[0])
        else:
            key = min(tags, tuple(reversed(tags)))
        if key in seen:
            raise ValueError(f"{name} must not contain duplicate records")
        seen.add(key)
        out.append((*key, type_id))
    _dense(name[:-1], (r[-1] for r in out))
    return tuple(out)


def _canonical_bulk(records: object | None, width: int, n: int) -> np.ndarray | None:
    """_canonical for an integer array of valid records, in numpy, as an
    [m, width] int64 array; None when the input is not such an array or any
    record would be rejected (the per-record path then raises the exact
    error)."""
    if records is None or isinstance(records, (str, bytes)):
        return None
    try:
        arr = np.asarray(records)
    except (TypeError, ValueError):
        return None
    if arr.ndim != 2 or arr.shape[1] != width or arr.shape[0] == 0 or arr.dtype.kind not in "iu":
        return None
    if arr.dtype.kind == "u" and arr.max() > np.iinfo(np.int64).max:
        return None
    a = arr.astype(np.int64)
    tags, types = a[:, :-1], a[:, -1]
    if (tags < 1).any() or (tags > n).any() or (types < 1).any() or (types > np.iinfo(np.int32).max).any():
        return None
    k = width - 1
    for p in range(k):
        for q in range(p + 1, k):
            if (tags[:, p] == tags[:, q]).any():
                return None
    if k == 2:
        keys = np.sort(tags, axis=1)
    else:
        # angles: outer tags ordered; dihedrals: the smaller of the forward and
        # reversed tuples, decided by the first entries (distinct tags)
        flip = tags[:, 0] > tags[:, -1]
        keys = np.where(flip[:, None], tags[:, ::-1], tags)
    order = np.lexsort(keys.T[::-1])
    sk = keys[order]
    if (sk[1:] == sk[:-1]).all(axis=1).any():
        return None
    if types.max() > types.size:  # more ids than records: not dense
        return None
    present = np.flatnonzero(np.bincount(types))
    if not np.array_equal(present, np.arange(1, present.size + 1)):
        return None
    return np.concatenate([keys, types[:, None]], axis=1)


def _supported_bulk(records: np.ndarray, bond_keys: np.ndarray) -> bool:
    """Every consecutive tag pair of every record is a bond (keys lo << 32 | hi)."""
    if not len(records):
        return True
    t = np.asarray(records, dtype=np.uint64)[:, :-1]
    lo = np.minimum(t[:, :-1], t[:, 1:])
    hi = np.maximum(t[:, :-1], t[:, 1:])
    return bool(np.isin((lo << np.uint64(32)) | hi, bond_keys).all())


def _require_support(name: str, records, bond_pairs: frozenset) -> None:
    for record in records:
        tags = record[:-1]
        for a, b in zip(tags[:-1], tags[1:]):
            if (min(a, b), max(a, b)) not in bond_pairs:
                raise ValueError(f"{name} require supporting topology bonds for each edge")


def recheck(topo: Topology, tags: Iterable[int]) -> Topology:
    """The same topology re-validated against (possibly changed) tags."""
    if topo.arrays is not None:
        b, a, d = topo.arrays
        again = make_topology(bonds=b, angles=a, dihedrals=d, tags=tags, as_arrays=True)
        if again.arrays is not None and all(
                np.array_equal(x, y) for x, y in zip(again.arrays, topo.arrays)):
            return topo
    return make_topology(bonds=topo.bonds, angles=topo.angles, dihedrals=topo.dihedrals, tags=tags)


def make_topology(*, bonds=None, angles=None, dihedrals=None, tags: Iterable[int],
                  as_arrays: bool = False) -> Topology:
    """Canonical topology; as_arrays skips building the tuple records when
    every section was checked in bulk (the tuple fields are then empty and only
    `arrays` is meaningful)."""
    tag_arr = np.asarray(list(tags) if not isinstance(tags, np.ndarray) else tags)
    n = int(tag_arr.size)
    # System tags are a permutation of 1..n: records can then be checked in
    # bulk, falling back to the per-record path (and its messages) on any doubt.
    bulk = n > 0 and np.array_equal(np.sort(tag_arr.astype(np.int64)), np.arange(1, n + 1))
    known: frozenset[int] | None = None
    arrays: list = []

    def canonical(name: str, records, width: int) -> tuple:
        nonlocal known
        arr = _canonical_bulk(records, width, n) if bulk else None
        if arr is None:
            if known is None:
                known = frozenset(int(t) for t in tag_arr.tolist())
            out = _canonical(name, records, width, known)
            arrays.append(np.asarray(out, dtype=np.int64).reshape(-1, width))
            return out
        arrays.append(arr)
        return () if as_arrays else tuple(map(tuple, arr.tolist()))

    b = canonical("bonds", bonds, 3)
    bond_keys = None

    def support(name: str, recs: np.ndarray) -> None:
        nonlocal bond_keys
        if not len(recs):
            return
        if bond_keys is None:
            bt = arrays[0].astype(np.uint64)
            bond_keys = (bt[:, 0] << np.uint64(32)) | bt[:, 1]
        if not _supported_bulk(recs, bond_keys):
            bonds_t = b if b or not len(arrays[0]) else tuple(map(tuple, arrays[0].tolist()))
            recs_t = recs if isinstance(recs, tuple) else tuple(map(tuple, recs.tolist()))
            _require_support(name, recs_t, frozenset((r[0], r[1]) for r in bonds_t))

    a = canonical("angles", angles, 4)
    support("angles", arrays[1])
    d = canonical("dihedrals", dihedrals, 5)
    support("dihedrals", arrays[2])
    return Topology(b, a, d, arrays=tuple(arrays))
