"""``ForceField`` and ``Dynamics`` builders.

ForceField follows python/beads/_forcefield.py:63-514: one pair style with
symmetric per-type-pair coefficients, optional harmonic bond/angle/dihedral
styles with per-type coefficients, and bonded LJ exclusions (default distance 2
whenever bonds exist, or explicit via ``exclude_bonded``). Coverage is checked
against the System when a Simulation binds it. Dynamics follows
python/beads/_dynamics.py:24-152 (velocity_verlet + optional berendsen).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Mapping

import numpy as np

from ._checks import freeze_params, frozen, positive_int, positive_type_id

DEFAULT_EXCLUSION_DISTANCE = 2


@dataclass(frozen=True)
class _Coeff:
    style: str
    key: tuple
    params: Mapping[str, Any]


@dataclass(frozen=True)
class ForceFieldConfig:
    pair_style: str | None
    pair_style_params: Mapping[str, Any]
    pair_coeffs: tuple[_Coeff, ...]          # key = (type_i, type_j), type_i <= type_j
    listed: Mapping[str, tuple[str | None, Mapping[str, Any], tuple[_Coeff, ...]]]
    bonded_exclusion_distance: int | None
    bonded_exclusion_policy: str


def _style(context: str, style: object) -> str:
    if not isinstance(style, str):
        raise TypeError(f"{context} name must be a string")
    if not style:
        raise ValueError(f"{context} name must not be empty")
    return style


_LISTED = ("bond", "angle", "dihedral")
_LISTED_WIDTH = {"bond": 1, "angle": 2, "dihedral": 3}  # index of type in a record - 1


class ForceField:
    def __init__(self) -> None:
        self._pair: tuple[str, dict] | None = None
        self._pair_coeffs: dict[tuple[int, int], tuple[str, dict]] = {}
        self._styles: dict[str, tuple[str, dict] | None] = {k: None for k in _LISTED}
        self._coeffs: dict[str, dict[int, tuple[str, dict]]] = {k: {} for k in _LISTED}
        self._exclusion_distance: int | None = None
        self._exclusion_policy: str | None = None
        self._frozen = False

    # -- pair ---------------------------------------------------------------
    def pair_style(self, style: str, /, **params: object) -> "ForceField":
        self._require_mutable()
        if self._pair is not None:
            raise ValueError("ForceField supports one active pair_style.")
        self._pair = (_style("pair_style", style),
                      freeze_params("pair parameter", params, allow_bool=True))
        return self

    def pair_coeff(self, style: str, /, *, type_i: object, type_j: object,
                   **params: object) -> "ForceField":
        self._require_mutable()
        if self._pair is None:
            raise ValueError("pair_style must be set before pair_coeff.")
        name = _style("pair_coeff style", style)
        if name != self._pair[0]:
            raise ValueError("pair_coeff style must match active pair_style.")
        a = positive_type_id("pair_coeff type_i", type_i)
        b = positive_type_id("pair_coeff type_j", type_j)
        key = (min(a, b), max(a, b))
        if key in self._pair_coeffs:
            raise ValueError(f"pair_coeff for types {key[0]} {key[1]} is already set.")
        self._pair_coeffs[key] = (name, freeze_params("pair parameter", params, allow_bool=True))
        return self

    # -- listed terms -------------------------------------------------------
    def _set_style(self, kind: str, style: str, params: dict) -> "ForceField":
        self._require_mutable()
        if self._styles[kind] is not None:
            raise ValueError(f"ForceField supports one active {kind}_style.")
        self._styles[kind] = (_style(f"{kind}_style", style),
                              freeze_params(f"{kind} parameter", params, allow_bool=True))
        return self

    def _set_coeff(self, kind: str, style: str, type_id: object, params: dict) -> "ForceField":
        self._require_mutable()
        active = self._styles[kind]
        if active is None:
            raise ValueError(f"{kind}_style must be set before {kind}_coeff.")
        name = _style(f"{kind}_coeff style", style)
        if name != active[0]:
            raise ValueError(f"{kind}_coeff style must match active {kind}_style.")
        t = positive_type_id(f"{kind}_coeff type", type_id)
        if t in self._coeffs[kind]:
            raise ValueError(f"{kind}_coeff for type {t} is already set.")
        self._coeffs[kind][t] = (name, freeze_params(f"{kind} parameter", params, allow_bool=True))
        return self

    def bond_style(self, style: str, /, **params: object) -> "ForceField":
        return self._set_style("bond", style, params)

    def bond_coeff(self, style: str, /, *, type: object, **params: object) -> "ForceField":
        return self._set_coeff("bond", style, type, params)

    def angle_style(self, style: str, /, **params: object) -> "ForceField":
        return self._set_style("angle", style, params)

    def angle_coeff(self, style: str, /, *, type: object, **params: object) -> "ForceField":
        return self._set_coeff("angle", style, type, params)

    def dihedral_style(self, style: str, /, **params: object) -> "ForceField":
        return self._set_style("dihedral", style, params)

    def dihedral_coeff(self, style: str, /, *, type: object, **params: object) -> "ForceField":
        return self._set_coeff("dihedral", style, type, params)

    def exclude_bonded(self, *, distance: object = DEFAULT_EXCLUSION_DISTANCE) -> "ForceField":
        self._require_mutable()
        if self._exclusion_policy is not None:
            raise ValueError("ForceField bonded exclusions are already configured.")
        d = positive_int("bonded exclusion distance", distance)
        if d not in (1, 2, 3):
            raise ValueError("bonded exclusion distance must be 1, 2, or 3.")
        self._exclusion_distance = d
        self._exclusion_policy = "explicit"
        return self

    # -- introspection ------------------------------------------------------
    @property
    def pair_style_name(self) -> str | None:
        return None if self._pair is None else self._pair[0]

    @property
    def pair_style_params(self) -> dict[str, Any]:
        return {} if self._pair is None else dict(self._pair[1])

    @property
    def pair_coeff_entries(self):
        return tuple((s, a, b, dict(p)) for (a, b), (s, p) in sorted(self._pair_coeffs.items()))

    def _require_mutable(self) -> None:
        if self._frozen:
            raise RuntimeError("ForceField is frozen after binding to a Simulation.")

    # -- binding ------------------------------------------------------------
    def _freeze(self, system=None) -> ForceFieldConfig:
        distance, policy = self._exclusion_distance, self._exclusion_policy
        if policy is None:
            if system is not None and system.topology.bonds:
                distance, policy = DEFAULT_EXCLUSION_DISTANCE, "default"
            else:
                policy = "none"
        listed = {}
        for kind in _LISTED:
            st = self._styles[kind]
            coeffs = tuple(_Coeff(s, (t,), frozen(p)) for t, (s, p) in sorted(self._coeffs[kind].items()))
            listed[kind] = (None if st is None else st[0], frozen({} if st is None else st[1]), coeffs)
        cfg = ForceFieldConfig(
            pair_style=None if self._pair is None else self._pair[0],
            pair_style_params=frozen({} if self._pair is None else self._pair[1]),
            pair_coeffs=tuple(_Coeff(s, k, frozen(p)) for k, (s, p) in sorted(self._pair_coeffs.items())),
            listed=frozen(listed),
            bonded_exclusion_distance=distance,
            bonded_exclusion_policy=policy,
        )
        if system is not None:
            _check_pair_coverage(cfg, system)
            for kind in _LISTED:
                _check_listed_coverage(kind, cfg, system)
            if cfg.bonded_exclusion_distance is not None and not system.topology.bonds:
                raise ValueError("ForceField.exclude_bonded(...) requires System topology bonds.")
        self._frozen = True
        return cfg


def _check_pair_coverage(cfg: ForceFieldConfig, system) -> None:
    if cfg.pair_style is None:
        raise ValueError("ForceField requires an active pair_style.")
    active = sorted(int(t) for t in np.unique(system.types))
    present = set(active)
    observed = {c.key for c in cfg.pair_coeffs}
    stray = sorted(k for k in observed if k[0] not in present or k[1] not in present)
    if stray:
        raise ValueError(f"ForceField pair_coeff type ids must be active System types; got {stray}.")
    wanted = {(a, b) for a in active for b in active if a <= b}
    missing = sorted(wanted - observed)
    if missing:
        raise ValueError(
            f"ForceField pair_coeffs must cover all active System type pairs; missing {missing}."
        )


def _check_listed_coverage(kind: str, cfg: ForceFieldConfig, system) -> None:
    style, _, coeffs = cfg.listed[kind]
    records = getattr(system.topology, kind + "s")
    plural = kind + "s"
    if style is None:
        if records:
            raise ValueError(f"System topology {plural} require an active {kind}_style.")
        if coeffs:
            raise ValueError(f"{kind}_coeff requires an active {kind}_style.")
        return
    if not records:
        raise ValueError(f"{kind}_style requires System topology {plural}.")
    wanted = {r[-1] for r in records}
    observed = {c.key[0] for c in coeffs}
    missing = sorted(wanted - observed)
    if missing:
        raise ValueError(
            f"{kind}_coeffs must cover all active {kind} topology types; missing {missing}."
        )
    extra = sorted(observed - wanted)
    if extra:
        raise ValueError(f"{kind}_coeff types must be active {kind} topology types; got {extra}.")


# ---------------------------------------------------------------- dynamics --

@dataclass(frozen=True)
class DynamicsConfig:
    style: str
    params: Mapping[str, Any]
    thermostat: tuple[str, Mapping[str, Any]] | None


class Dynamics:
    def __init__(self, style: str, /, **params: object) -> None:
        if not isinstance(style, str):
            raise TypeError("Dynamics style must be a string.")
        if not style:
            raise ValueError("Dynamics style must not be empty.")
        if style != "velocity_verlet":
            raise ValueError(f"Unknown Dynamics style {style!r}.")
        frozen_params = freeze_params("Dynamics parameter", params, allow_bool=False)
        if set(frozen_params) != {"dt"}:
            raise ValueError('Dynamics("velocity_verlet") requires exactly parameter "dt".')
        dt = frozen_params["dt"]
        if not isinstance(dt, (int, float)) or dt <= 0:
            raise ValueError("Dynamics dt must be positive.")
        self._style = style
        self._params = frozen_params
        self._thermostat: tuple[str, Mapping[str, Any]] | None = None
        self._frozen = False

    @property
    def style(self) -> str:
        return self._style

    @property
    def params(self) -> dict[str, Any]:
        return dict(self._params)

    def set_thermostat(self, style: str, /, **params: object) -> "Dynamics":
        if self._frozen:
            raise RuntimeError("Dynamics is frozen after binding to a Simulation.")
        if not isinstance(style, str):
            raise TypeError("Thermostat style must be a string.")
        if not style:
            raise ValueError("Thermostat style must not be empty.")
        if style != "berendsen":
            raise ValueError(f"Unknown Thermostat style {style!r}.")
        p = freeze_params("Thermostat parameter", params, allow_bool=False)
        if set(p) != {"temperature", "tau"}:
            raise ValueError(
                'Thermostat("berendsen") requires exactly parameters "temperature" and "tau".'
            )
        for key in ("temperature", "tau"):
            if not isinstance(p[key], (int, float)):
                raise TypeError(f"Thermostat parameter {key} must be a real number.")
        if float(p["temperature"]) < 0.0:
            raise ValueError("Thermostat temperature must be non-negative.")
        if float(p["tau"]) <= 0.0:
            raise ValueError("Thermostat tau must be positive.")
        self._thermostat = (style, frozen(p))
        return self

    def _freeze(self) -> DynamicsConfig:
        self._frozen = True
        return DynamicsConfig(self._style, frozen(self._params), self._thermostat)
