"""``Simulation``: binds System, ForceField and Dynamics with neighbour and
output settings and hands one spec dict to the native engine.

Mirrors python/beads/_simulation.py:22-291 (builder methods and the spec
dict layout), python/beads/_neighbor.py:10-43 (neighbour defaults),
python/beads/_output.py:11-128 (output prefixes and trajectory fields) and
python/beads/_native_bridge.py:11-31 (the call into the extension).
"""

from __future__ import annotations

import os
from dataclasses import dataclass
from importlib import import_module

from ._checks import nonnegative_real, nonnegative_uint64, positive_index, positive_uint64
from .forcefield import Dynamics, ForceField
from .system import System

TRAJECTORY_FIELDS = ("tag", "type", "position", "image", "velocity", "force")
DEFAULT_TRAJECTORY_FIELDS = ("tag", "type", "position", "image")


@dataclass(frozen=True)
class NeighborSettings:
    cutoff_buffer: float = 0.3
    rebuild_check_every: int = 1
    sort_every_rebuild: int = 10
    max_neighbors: int = 512
    # No reference counterpart: store each LJ pair once (j > i) and add the
    # reaction with atomics instead of keeping the reference's full list.
    half_list: bool = False
    # No reference counterpart: sort at the first rebuild at least this many
    # steps after the last sort, instead of every sort_every_rebuild-th build.
    sort_every_steps: int | None = None

    def __post_init__(self) -> None:
        if not isinstance(self.half_list, bool):
            raise TypeError("neighbor half_list must be a bool")
        if self.sort_every_steps is not None:
            object.__setattr__(self, "sort_every_steps",
                               positive_uint64("neighbor sort_every_steps", self.sort_every_steps))
        buf = nonnegative_real("neighbor cutoff_buffer", self.cutoff_buffer)
        check = positive_uint64("neighbor rebuild_check_every", self.rebuild_check_every)
        sort = positive_uint64("neighbor sort_every_rebuild", self.sort_every_rebuild)
        maxn = positive_index("neighbor max_neighbors", self.max_neighbors)
        if buf == 0.0 and check != 1:
            raise ValueError("neighbor cutoff_buffer=0 requires rebuild_check_every=1")
        object.__setattr__(self, "cutoff_buffer", buf)
        object.__setattr__(self, "rebuild_check_every", check)
        object.__setattr__(self, "sort_every_rebuild", sort)
        object.__setattr__(self, "max_neighbors", maxn)


def _prefix(label: str, prefix: object) -> str:
    try:
        value = os.fspath(prefix)
    except TypeError as exc:
        raise TypeError(f"{label} prefix must be a string or path-like object") from exc
    if not isinstance(value, str):
        raise TypeError(f"{label} prefix must be a string or path-like object")
    if not value:
        raise ValueError(f"{label} prefix must not be empty")
    return value


def _trajectory_fields(fields: object | None) -> tuple[str, ...]:
    if fields is None:
        return DEFAULT_TRAJECTORY_FIELDS
    if isinstance(fields, (str, bytes)):
        raise TypeError("trajectory fields must be an iterable of strings")
    try:
        values = tuple(fields)  # type: ignore[arg-type]
    except TypeError as exc:
        raise TypeError("trajectory fields must be an iterable of strings") from exc
    if not values:
        raise ValueError("trajectory fields must not be empty")
    for f in values:
        if not isinstance(f, str):
            raise TypeError("trajectory fields must contain only strings")
        if f not in TRAJECTORY_FIELDS:
            raise ValueError(f"trajectory field {f!r} is not supported")
    if values[0] != "tag":
        raise ValueError("trajectory fields must include tag as the first field")
    if len(set(values)) != len(values):
        raise ValueError("trajectory fields must not contain duplicates")
    order = [TRAJECTORY_FIELDS.index(f) for f in values]
    if order != sorted(order):
        raise ValueError(
            "trajectory fields must follow canonical order: tag,type,position,image,velocity,force"
        )
    return values


def _log_settings(echo: object, prefix: object | None) -> dict:
    if not isinstance(echo, str):
        raise TypeError("logging echo must be a string")
    if echo not in ("screen", "log", "both", "none"):
        raise ValueError("logging echo must be one of 'screen', 'log', 'both', or 'none'")
    if echo in ("log", "both"):
        if prefix is None:
            raise ValueError(f"logging echo={echo!r} requires a prefix")
        return {"echo": echo, "prefix": _prefix("log", prefix)}
    if prefix is not None:
        raise ValueError(f"logging echo={echo!r} does not accept a prefix")
    return {"echo": echo}


def load_native():
    """The compiled engine binding of the selected precision (BEADS_AMD_PRECISION);
    fails loudly when it is not built."""
    from ._build_config import native_module_name
    from ._runtime import share_hip_runtime

    share_hip_runtime()
    name = native_module_name()
    try:
        return import_module(name)
    except ModuleNotFoundError as exc:
        if exc.name != name:
            raise
        raise RuntimeError(
            "beads_amd native extension is not built. Run __graft_entry__.build() "
            "(or make -C beads_amd/csrc) before executing a simulation."
        ) from exc


class Simulation:
    def __init__(self, *, system: System, forcefield: ForceField, dynamics: Dynamics) -> None:
        if not isinstance(system, System):
            raise TypeError("Simulation requires a System.")
        if not isinstance(forcefield, ForceField):
            raise TypeError("Simulation requires a ForceField.")
        if not isinstance(dynamics, Dynamics):
            raise TypeError("Simulation requires a Dynamics.")
        self._system = system._freeze()
        self._forcefield = forcefield._freeze(system)
        self._dynamics = dynamics._freeze()
        self._neighbor = NeighborSettings()
        self._thermo: dict | None = None
        self._final_state: dict | None = None
        self._trajectory: dict | None = None
        self._log = {"echo": "screen"}
        self._runsteps = 0
        self._executed = False

    def _require_not_executed(self, action: str = "modified") -> None:
        if self._executed:
            if action == "execute":
                raise RuntimeError(
                    "Simulation.execute() cannot be called after execute(); "
                    "create a new Simulation for another run."
                )
            raise RuntimeError(
                "Simulation cannot be modified after execute(); "
                "create a new Simulation for another run."
            )

    def set_runsteps(self, runsteps: object) -> "Simulation":
        self._require_not_executed()
        self._runsteps = nonnegative_uint64("runsteps", runsteps)
        return self

    def set_neighbor(self, *, cutoff_buffer: object = 0.3, rebuild_check_every: object = 1,
                     sort_every_rebuild: object = 10, max_neighbors: object = 512,
                     half_list: bool = False,
                     sort_every_steps: object = None) -> "Simulation":
        self._require_not_executed()
        self._neighbor = NeighborSettings(cutoff_buffer, rebuild_check_every,
                                          sort_every_rebuild, max_neighbors, half_list,
                                          sort_every_steps)
        return self

    def set_thermo(self, *, every: object, prefix: object) -> "Simulation":
        self._require_not_executed()
        self._thermo = {"every": positive_uint64("thermo every", every),
                        "prefix": _prefix("thermo", prefix)}
        return self

    def save_final_state(self, *, prefix: object) -> "Simulation":
        self._require_not_executed()
        self._final_state = {"prefix": _prefix("final_state", prefix)}
        return self

    def set_trajectory(self, *, every: object, prefix: object,
                       fields: object | None = None) -> "Simulation":
        self._require_not_executed()
        self._trajectory = {"every": positive_uint64("trajectory every", every),
                            "prefix": _prefix("trajectory", prefix),
                            "fields": list(_trajectory_fields(fields))}
        return self

    def set_logging(self, *, echo: object = "screen", prefix: object | None = None) -> "Simulation":
        self._require_not_executed()
        self._log = _log_settings(echo, prefix)
        return self

    def _native_spec(self) -> dict[str, object]:
        s, ff, dyn = self._system, self._forcefield, self._dynamics
        output: dict[str, object] = {"log": dict(self._log)}
        if self._thermo is not None:
            output["thermo"] = dict(self._thermo)
        if self._final_state is not None:
            output["final_state"] = dict(self._final_state)
        if self._trajectory is not None:
            output["trajectory"] = dict(self._trajectory)
        dynamics: dict[str, object] = {"style": dyn.style, "params": dict(dyn.params)}
        if dyn.thermostat is not None:
            dynamics["thermostat"] = {"style": dyn.thermostat[0], "params": dict(dyn.thermostat[1])}
        forcefield: dict[str, object] = {
            "pair_style": {"style": ff.pair_style, "params": dict(ff.pair_style_params)},
            "pair_coeffs": [
                {"style": c.style, "type_i": c.key[0], "type_j": c.key[1], "params": dict(c.params)}
                for c in ff.pair_coeffs
            ],
            "bonded_exclusion_distance": ff.bonded_exclusion_distance,
            "bonded_exclusion_policy": ff.bonded_exclusion_policy,
        }
        for kind in ("bond", "angle", "dihedral"):
            style, params, coeffs = ff.listed[kind]
            forcefield[f"{kind}_style"] = None if style is None else {"style": style,
                                                                      "params": dict(params)}
            forcefield[f"{kind}_coeffs"] = [
                {"style": c.style, "type": c.key[0], "params": dict(c.params)} for c in coeffs
            ]
        topo = s.topology
        return {
            "system": {
                "units": s.units,
                "n_particles": s.n_particles,
                "box_bound": s.box_bound,
                "positions": s.positions,
                "velocities": s.velocities,
                "masses": s.masses,
                "types": s.types,
                "tags": s.tags,
                "molecule_ids": s.molecule_ids,
                "images": s.images,
                "topology": {
                    "bonds": [{"tag_i": a, "tag_j": b, "type": t} for a, b, t in topo.bonds],
                    "angles": [{"tag_i": a, "tag_j": b, "tag_k": c, "type": t}
                               for a, b, c, t in topo.angles],
                    "dihedrals": [{"tag_i": a, "tag_j": b, "tag_k": c, "tag_l": d, "type": t}
                                  for a, b, c, d, t in topo.dihedrals],
                },
            },
            "forcefield": forcefield,
            "dynamics": dynamics,
            "neighbor": {
                "cutoff_buffer": self._neighbor.cutoff_buffer,
                "rebuild_check_every": self._neighbor.rebuild_check_every,
                "sort_every_rebuild": self._neighbor.sort_every_rebuild,
                "max_neighbors": self._neighbor.max_neighbors,
                **({"half_list": True} if self._neighbor.half_list else {}),
                **({"sort_every_steps": self._neighbor.sort_every_steps}
                   if self._neighbor.sort_every_steps is not None else {}),
            },
            "output": output,
            "runsteps": self._runsteps,
        }

    def execute(self) -> None:
        self._require_not_executed("execute")
        self._executed = True
        spec = self._native_spec()
        native = load_native()
        from ._build_config import check_native_build_info
        check_native_build_info(native)
        native.execute_simulation(spec)
