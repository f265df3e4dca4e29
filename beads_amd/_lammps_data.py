"""LAMMPS data-file import (``beads.io.read_lammps_data``).

Behaviour of python/beads/io.py:97-230 with its helpers (:301-753): a minimal
``Atoms # molecular`` subset with optional Masses and Velocities, orthogonal
box bounds and Bonds/Angles/Dihedrals topology. Force-field sections are
skipped, type labels, Impropers, triclinic boxes and body/ellipsoid sections
are rejected, and every count, id and type must satisfy what
``beads.System`` accepts. The checks run in the reference's order and raise
its exception types and messages, so callers see the same failures.
"""

from __future__ import annotations

import math
import os

import numpy as np

_SYSTEM = ("Masses", "Atoms", "Velocities", "Bonds", "Angles", "Dihedrals", "Impropers")
_LABELS = ("Atom Type Labels", "Bond Type Labels", "Angle Type Labels",
           "Dihedral Type Labels", "Improper Type Labels")
_COEFFS = ("Pair Coeffs", "PairIJ Coeffs", "Bond Coeffs", "Angle Coeffs", "Dihedral Coeffs",
           "Improper Coeffs", "BondBond Coeffs", "BondAngle Coeffs",
           "MiddleBondTorsion Coeffs", "EndBondTorsion Coeffs", "AngleTorsion Coeffs",
           "AngleAngleTorsion Coeffs", "BondBond13 Coeffs", "AngleAngle Coeffs")
_UNSUPPORTED = ("Ellipsoids", "Lines", "Triangles", "Bodies")
_CANONICAL = {name.lower(): name for name in _SYSTEM + _LABELS + _COEFFS + _UNSUPPORTED}
_COUNTS = ("atoms", "bonds", "angles", "dihedrals", "impropers", "atom types", "bond types",
           "angle types", "dihedral types", "improper types")
_TERMS = (("Bonds", "bonds", "bond", 2), ("Angles", "angles", "angle", 3),
          ("Dihedrals", "dihedrals", "dihedral", 4))


def _err(message: str) -> ValueError:
    return ValueError(f"LAMMPS data {message}")


def _as_int(token: str, what: str, low: int | None = None) -> int:
    """Decimal integer; low = 1 (positive) or 0 (non-negative) adds a bound."""
    try:
        value = int(token, 10)
    except ValueError as exc:
        raise _err(f"{what} must be an integer.") from exc
    if low == 1 and value <= 0:
        raise _err(f"{what} must be positive.")
    if low == 0 and value < 0:
        raise _err(f"{what} must be non-negative.")
    return value


def _as_float(token: str, what: str) -> float:
    try:
        value = float(token)
    except ValueError as exc:
        raise _err(f"{what} must be numeric.") from exc
    if not math.isfinite(value):
        raise _err(f"{what} must be finite.")
    return value


def _columns(tokens: list[str], width: int, what: str, layout: str | None = None) -> None:
    if len(tokens) != width:
        hint = f" ({layout})" if layout else ""
        raise _err(f"{what} expected {width} columns{hint}, got {len(tokens)}.")


def _section_of(content: str) -> str | None:
    words = " ".join(content.split())
    name = _CANONICAL.get(words.lower())
    if name is None and words.lower().endswith((" coeffs", " type labels")):
        name = words
    return name


class _DataFile:
    """Header keywords and raw section rows of one data file."""

    def __init__(self, path: object) -> None:
        self.header: dict[str, object] = {}
        self.sections: dict[str, tuple[str, list[tuple[int, str]]]] = {}
        current: list[tuple[int, str]] | None = None
        with open(os.fspath(path), "r", encoding="utf-8") as fh:
            for number, raw in enumerate(fh, start=1):
                body, hash_mark, note = raw.partition("#")
                content = body.strip()
                if not content:
                    continue
                name = _section_of(content)
                if name is not None:
                    if name in self.sections:
                        raise _err(f"section {name!r} is duplicated.")
                    current = []
                    self.sections[name] = (note.strip() if hash_mark else "", current)
                elif current is not None:
                    current.append((number, content))
                else:
                    self._header_line(content)

    def _header_line(self, content: str) -> None:
        tokens = content.split()
        for axis in "xyz":
            if tokens[-2:] == [f"{axis}lo", f"{axis}hi"]:
                _columns(tokens, 4, f"{axis} bounds")
                self.header[f"{axis}lo"] = _as_float(tokens[0], f"{axis}lo")
                self.header[f"{axis}hi"] = _as_float(tokens[1], f"{axis}hi")
                return
        if tokens[-3:] == ["xy", "xz", "yz"]:
            raise _err("import supports only orthogonal box bounds; "
                       "triclinic tilt factors are unsupported.")
        if tokens[0] in ("avec", "bvec", "cvec") or tokens[:2] == ["abc", "origin"]:
            raise _err("import does not support general triclinic boxes.")
        key = " ".join(tokens[1:])
        if len(tokens) >= 2 and key in _COUNTS:
            self.header[key] = _as_int(tokens[0], key, low=0)
            return
        self.header.setdefault("title", content)

    def rows(self, name: str) -> list[tuple[int, str]]:
        return self.sections[name][1] if name in self.sections else []


def _declared(header: dict, key: str, ids, context: str) -> None:
    if key in header:
        highest = max((int(v) for v in ids), default=0)
        if highest > int(header[key]):
            raise _err(f"{context} type id {highest} exceeds declared {key} count "
                       f"{header[key]}.")


def _dense(label: str, ids) -> None:
    seen = sorted({int(v) for v in ids})
    if seen and seen != list(range(1, seen[-1] + 1)):
        raise _err(f"{label} type ids must be dense one-based ids without gaps.")


def _atoms(data: _DataFile) -> list[tuple[int, int, int, tuple, tuple]]:
    """(id, molecule, type, position, image) rows sorted by id."""
    style = data.sections["Atoms"][0]
    if style and style.split()[0].lower() != "molecular":
        raise _err("import supports only Atoms # molecular.")
    out, ids, with_images = [], set(), None
    for number, content in data.rows("Atoms"):
        t = content.split()
        if len(t) not in (6, 9):
            raise _err(f"Atoms line {number} must use molecular format: "
                       "atom-ID molecule-ID atom-type x y z [ix iy iz].")
        if with_images is None:
            with_images = len(t) == 9
        elif with_images != (len(t) == 9):
            raise _err("Atoms section must not mix image flags.")
        where = f"Atoms line {number}"
        atom = _as_int(t[0], f"{where} atom-ID", low=1)
        if atom in ids:
            raise _err(f"Atoms section duplicates atom ID {atom}.")
        ids.add(atom)
        mol = _as_int(t[1], f"{where} molecule-ID")
        if mol < 0:
            raise _err("molecule-ID values must be non-negative.")
        typ = _as_int(t[2], f"{where} atom-type", low=1)
        xyz = tuple(_as_float(v, f"{where} coordinate") for v in t[3:6])
        img = tuple(_as_int(v, f"{where} image flag") for v in t[6:9]) if with_images else (0, 0, 0)
        out.append((atom, mol, typ, xyz, img))
    out.sort(key=lambda row: row[0])
    return out


def _velocities(data: _DataFile, ids: set[int]) -> dict[int, tuple]:
    found: dict[int, tuple] = {}
    for number, content in data.rows("Velocities"):
        t = content.split()
        _columns(t, 4, f"Velocities line {number}")
        atom = _as_int(t[0], f"Velocities line {number} atom-ID", low=1)
        if atom not in ids:
            raise _err(f"Velocities section references unknown atom ID {atom}.")
        if atom in found:
            raise _err(f"Velocities section duplicates atom ID {atom}.")
        found[atom] = tuple(_as_float(v, f"Velocities line {number} component") for v in t[1:])
    if set(found) != ids:
        raise _err("Velocities section must include every atom exactly once.")
    return found


def _masses(data: _DataFile) -> dict[int, float]:
    found: dict[int, float] = {}
    for number, content in data.rows("Masses"):
        t = content.split()
        _columns(t, 2, f"Masses line {number}", "atom-type mass")
        typ = _as_int(t[0], f"Masses line {number} atom type", low=1)
        if typ in found:
            raise _err(f"Masses section duplicates atom type {typ}.")
        mass = _as_float(t[1], f"Masses line {number} mass")
        if mass <= 0.0:
            raise _err("masses must be positive.")
        found[typ] = mass
    return found


def _terms(data: _DataFile, section: str, width: int, ids: set[int]) -> tuple[tuple[int, ...], ...]:
    layout = "id type " + " ".join(f"atom{k}" for k in range(1, width + 1))
    out, term_ids = [], set()
    for number, content in data.rows(section):
        t = content.split()
        where = f"{section} line {number}"
        _columns(t, width + 2, where, layout)
        term = _as_int(t[0], f"{where} id", low=1)
        if term in term_ids:
            raise _err(f"{section} section duplicates id {term}.")
        term_ids.add(term)
        typ = _as_int(t[1], f"{where} type", low=1)
        members = tuple(_as_int(v, f"{where} atom id", low=1) for v in t[2:])
        unknown = [m for m in members if m not in ids]
        if unknown:
            raise _err(f"{section} section references unknown atom ID {unknown[0]}.")
        out.append(members + (typ,))
    return tuple(out)


def read_lammps_data(path: object, *, units: str):
    """Read a minimal LAMMPS data-file subset as a ``beads.System``.

    Atoms (molecular style, optional image flags), optional Masses and
    Velocities, orthogonal box bounds and Bonds/Angles/Dihedrals. Force-field
    coefficients are not read and no unit conversion happens: ``units`` goes
    to ``beads.System`` as given."""
    from .system import System

    if units is None:
        raise TypeError("read_lammps_data(...) requires the units keyword argument.")
    data = _DataFile(path)
    header, sections = data.header, data.sections
    labels = sorted(n for n in sections if n in _LABELS or n.endswith("Type Labels"))
    if labels:
        raise _err(f"type labels are not supported: {', '.join(labels)}.")
    unsupported = sorted(n for n in sections if n in _UNSUPPORTED)
    if unsupported:
        raise _err(f"import does not support sections: {', '.join(unsupported)}.")
    if any(k not in header for k in ("xlo", "xhi", "ylo", "yhi", "zlo", "zhi")):
        raise _err("file is missing required orthogonal box bounds.")
    for axis in "xyz":
        if float(header[f"{axis}hi"]) <= float(header[f"{axis}lo"]):
            raise _err(f"{axis}hi must be greater than {axis}lo.")
    if "atoms" not in header:
        raise _err("file is missing the atoms count.")
    if "Atoms" not in sections:
        raise _err("file is missing the Atoms section.")
    if int(header.get("impropers", 0)) > 0 or data.rows("Impropers"):
        raise _err("import does not support Impropers. Remove the Impropers section before "
                   "importing the file into BEADS.")

    atoms = _atoms(data)
    if len(atoms) != int(header["atoms"]):
        raise _err(f"atoms count is {header['atoms']}, but Atoms contains {len(atoms)} rows.")
    n = len(atoms)
    ids = [a[0] for a in atoms]
    if set(ids) != set(range(1, n + 1)):
        raise _err("atom-ID values must satisfy BEADS System.tags: a unique 1..N permutation.")
    mols = [a[1] for a in atoms]
    if any(m <= 0 for m in mols):
        raise _err("molecule-ID values must satisfy BEADS System.molecule_ids: "
                   "positive dense 1-based ids.")
    distinct = set(mols)
    if distinct != set(range(1, len(distinct) + 1)) or max(distinct, default=0) > n:
        raise _err("molecule-ID values must satisfy BEADS System.molecule_ids: "
                   "positive dense 1-based ids that do not exceed the particle count.")
    id_set = set(ids)
    types = [a[2] for a in atoms]
    _declared(header, "atom types", types, "atom")
    _dense("atom", types)

    velocities = _velocities(data, id_set) if "Velocities" in sections else None
    masses = None
    if "Masses" in sections:
        masses = _masses(data)
        _declared(header, "atom types", masses, "Masses atom")

    terms = {section: _terms(data, section, width, id_set) for section, _, _, width in _TERMS}
    for section, key, _, _ in _TERMS:
        if key in header and int(header[key]) != len(terms[section]):
            raise _err(f"{key} count is {header[key]}, but the section contains "
                       f"{len(terms[section])} rows.")
    for section, _, word, _ in _TERMS:
        _declared(header, f"{word} types", (e[-1] for e in terms[section]), word)
    for section, _, word, _ in _TERMS:
        _dense(f"{word} topology", (e[-1] for e in terms[section]))

    mass_array = None
    if masses is not None:
        missing = sorted(set(types) - set(masses))
        if missing:
            raise _err("Masses section is missing atom types: "
                       f"{', '.join(str(m) for m in missing)}.")
        mass_array = np.array([masses[t] for t in types], dtype=np.float64)
    system = System(
        box_bound=np.array([[header["xlo"], header["ylo"], header["zlo"]],
                            [header["xhi"], header["yhi"], header["zhi"]]], dtype=np.float64),
        positions=np.array([a[3] for a in atoms], dtype=np.float64),
        types=np.array(types, dtype=np.int64),
        velocities=(None if velocities is None
                    else np.array([velocities[i] for i in ids], dtype=np.float64)),
        masses=mass_array,
        tags=np.array(ids, dtype=np.uint64),
        molecule_ids=np.array(mols, dtype=np.uint64),
        images=np.array([a[4] for a in atoms], dtype=np.int64),
        units=units,
    )
    bonds, angles, dihedrals = (terms[s] for s, _, _, _ in _TERMS)
    if bonds or angles or dihedrals:
        system.set_topology(bonds=bonds, angles=angles, dihedrals=dihedrals)
    return system
