"""Input checks shared by the public builders.

Behaviour follows the reference front end (python/beads/_validation.py:16-151,
python/beads/_params.py:20-83): the same accepted kinds of values, the same
exception types, and messages that name the offending input.
"""

from __future__ import annotations

from types import MappingProxyType
from typing import Any, Mapping

import numpy as np

from ._build_config import dtype_for

REAL = dtype_for("real_dtype")
TYPE = dtype_for("type_dtype")
INDEX = dtype_for("index_dtype")
IMAGE = dtype_for("image_dtype")
RUNSTEP = dtype_for("runstep_dtype")

_NON_REAL_KINDS = frozenset("bOSUc")
_NON_INT_KINDS = frozenset("bOSUfc")


def real_array(name: str, value: object, shape: tuple[int, ...] | None = None) -> np.ndarray:
    probe = np.asarray(value)
    if probe.dtype.kind in _NON_REAL_KINDS:
        raise TypeError(f"{name} must contain real numeric values")
    out = np.array(value, dtype=REAL, copy=True, order="C")
    if shape is not None and out.shape != shape:
        raise ValueError(f"{name} must have shape {shape}")
    if not np.isfinite(out).all():
        raise ValueError(f"{name} must be finite")
    return out


def rows_of_three(name: str, value: object, n_rows: int | None = None) -> np.ndarray:
    out = real_array(name, value)
    if out.ndim != 2 or out.shape[1] != 3:
        raise ValueError(f"{name} must have shape (n_particles, 3)")
    if n_rows is not None and out.shape[0] != n_rows:
        raise ValueError(f"{name} must have one row per particle")
    return out


def integer_array(name: str, value: object, shape: tuple[int, ...] | None = None,
                  *, dtype=TYPE) -> np.ndarray:
    probe = np.asarray(value)
    if probe.dtype.kind in _NON_INT_KINDS:
        raise TypeError(f"{name} must contain integer values")
    if shape is not None and probe.shape != shape:
        raise ValueError(f"{name} must have shape {shape}")
    info = np.iinfo(dtype)
    if probe.size and (int(probe.min()) < info.min or int(probe.max()) > info.max):
        raise ValueError(
            f"{name} contains values outside the supported {np.dtype(dtype).name} range"
        )
    return np.array(value, dtype=dtype, copy=True, order="C")


def _plain_int(name: str, value: object) -> int:
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{name} must be an integer")
    return int(value)


def positive_int(name: str, value: object) -> int:
    out = _plain_int(name, value)
    if out <= 0:
        raise ValueError(f"{name} must be positive")
    return out


def positive_index(name: str, value: object) -> int:
    out = positive_int(name, value)
    if out > np.iinfo(INDEX).max:
        raise ValueError(f"{name} contains values outside the supported uint32 range")
    return out


def positive_type_id(name: str, value: object) -> int:
    out = positive_int(name, value)
    if out > np.iinfo(TYPE).max:
        raise ValueError(f"{name} contains values outside the supported int32 range")
    return out


def nonnegative_uint64(name: str, value: object) -> int:
    out = _plain_int(name, value)
    if out < 0:
        raise ValueError(f"{name} must be non-negative")
    if out > np.iinfo(RUNSTEP).max:
        raise ValueError(f"{name} contains values outside the supported uint64 range")
    return out


def positive_uint64(name: str, value: object) -> int:
    out = nonnegative_uint64(name, value)
    if out == 0:
        raise ValueError(f"{name} must be positive")
    return out


def nonnegative_real(name: str, value: object) -> float:
    if isinstance(value, (bool, np.bool_)) or not isinstance(
        value, (int, float, np.integer, np.floating)
    ):
        raise TypeError(f"{name} must be a real number")
    out = float(value)
    if not np.isfinite(out):
        raise ValueError(f"{name} must be finite")
    if out < 0.0:
        raise ValueError(f"{name} must be non-negative")
    return out


def freeze_params(context: str, params: Mapping[str, object], *, allow_bool: bool) -> dict[str, Any]:
    """Scalar style parameters: bool (if allowed), int64, finite float representable
    in the build's real dtype, or str."""
    out: dict[str, Any] = {}
    for key, value in params.items():
        if not isinstance(key, str) or not key:
            raise TypeError(f"{context} names must be non-empty strings.")
        out[key] = _scalar(context, key, value, allow_bool)
    return out


def _scalar(context: str, key: str, value: object, allow_bool: bool) -> Any:
    if isinstance(value, (bool, np.bool_)):
        if not allow_bool:
            raise TypeError(f"{context} {key} must not be bool.")
        return bool(value)
    if isinstance(value, (int, np.integer)):
        iv = int(value)
        bounds = np.iinfo(np.int64)
        if iv < bounds.min or iv > bounds.max:
            raise ValueError(f"{context} {key} contains values outside the supported int64 range.")
        return iv
    if isinstance(value, (float, np.floating)):
        fv = float(value)
        if not np.isfinite(fv):
            raise ValueError(f"{context} {key} must be finite.")
        with np.errstate(over="ignore", under="ignore"):
            narrowed = float(REAL.type(fv))
        if not np.isfinite(narrowed):
            raise ValueError(f"{context} {key} must be representable as {REAL.name}.")
        return fv
    if isinstance(value, str):
        return value
    kinds = "bool, int, float, or string" if allow_bool else "int, float, or string"
    raise TypeError(f"{context} {key} must be {kinds}.")


def frozen(mapping: Mapping[str, Any]) -> Mapping[str, Any]:
    return MappingProxyType(dict(mapping))
