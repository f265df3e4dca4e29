"""Tag-based topology records and their canonical forms.

Canonical forms follow python/beads/_topology.py:53-129: bonds as (min, max),
angles with the outer tags ordered (i < k), dihedrals as the smaller of the
forward and reversed tag tuples; duplicates and repeated tags are rejected,
types must be dense 1-based, and every angle/dihedral edge must be a bond
(python/beads/_topology.py:203-224).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable

from ._checks import positive_index, positive_type_id


@dataclass(frozen=True)
class Topology:
    bonds: tuple[tuple[int, int, int], ...]
    angles: tuple[tuple[int, int, int, int], ...]
    dihedrals: tuple[tuple[int, int, int, int, int], ...]


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


def _require_support(name: str, records, bond_pairs: frozenset) -> None:
    for record in records:
        tags = record[:-1]
        for a, b in zip(tags[:-1], tags[1:]):
            if (min(a, b), max(a, b)) not in bond_pairs:
                raise ValueError(f"{name} require supporting topology bonds for each edge")


def make_topology(*, bonds=None, angles=None, dihedrals=None, tags: Iterable[int]) -> Topology:
    known = frozenset(int(t) for t in tags)
    b = _canonical("bonds", bonds, 3, known)
    pairs = frozenset((r[0], r[1]) for r in b)
    a = _canonical("angles", angles, 4, known)
    _require_support("angles", a, pairs)
    d = _canonical("dihedrals", dihedrals, 5, known)
    _require_support("dihedrals", d, pairs)
    return Topology(b, a, d)
