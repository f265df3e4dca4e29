"""Dtype configuration reported by the native engine.

Same contract as python/beads/_build_config.py:11-108: ``build_info()`` returns
a read-only mapping of the engine's dtypes, and a run refuses to start if the
extension reports a configuration different from the one the Python layer
validates arrays with.

The engine is built in both precisions (the reference's BEADS_REAL64 switch,
CMakeLists.txt:29-31): ``beads_amd._native`` over libbeads_amd.so stores reals
in binary32, ``beads_amd._native_real64`` over libbeads_amd_real64.so in
binary64. ``BEADS_AMD_PRECISION`` ("single", the default, or "double") chooses
the one this process uses; it is read once, when the package is imported.
"""

from __future__ import annotations

import os
from collections.abc import Mapping
from importlib import import_module
from types import MappingProxyType

import numpy as np

PRECISION_ENV = "BEADS_AMD_PRECISION"
_NATIVE_MODULES = {"single": "beads_amd._native", "double": "beads_amd._native_real64"}


def selected_precision() -> str:
    """The precision named by BEADS_AMD_PRECISION ("single" when unset)."""
    value = os.environ.get(PRECISION_ENV, "single").strip().lower() or "single"
    if value not in _NATIVE_MODULES:
        raise RuntimeError(f'{PRECISION_ENV} must be "single" or "double", not {value!r}.')
    return value


def native_module_name(precision: str | None = None) -> str:
    return _NATIVE_MODULES[precision or selected_precision()]

_SOURCE_DEFAULT = {
    "real_precision": "single",
    "real_dtype": "float32",
    "type_dtype": "int32",
    "index_dtype": "uint32",
    "image_dtype": "int32",
    "runstep_dtype": "uint64",
}
_KNOWN_DTYPES = {"float32", "float64", "int32", "uint32", "uint64"}


def _normalize(info: object) -> dict[str, str]:
    if not isinstance(info, Mapping):
        raise RuntimeError("beads_amd native build_info must be a mapping.")
    out = {}
    for key in _SOURCE_DEFAULT:
        value = info.get(key)
        if not isinstance(value, str) or not value:
            raise RuntimeError(f"beads_amd native build_info requires string field {key!r}.")
        out[key] = value
    if out["real_precision"] not in ("single", "double"):
        raise RuntimeError('beads_amd real_precision must be "single" or "double".')
    expected = "float32" if out["real_precision"] == "single" else "float64"
    if out["real_dtype"] != expected:
        raise RuntimeError(
            f"beads_amd real_dtype must match real_precision ({out['real_precision']} -> {expected})."
        )
    for key in ("real_dtype", "type_dtype", "index_dtype", "image_dtype", "runstep_dtype"):
        if out[key] not in _KNOWN_DTYPES:
            raise RuntimeError(f"beads_amd {key} is unsupported: {out[key]!r}.")
    return out


def _initial() -> dict[str, str]:
    precision = selected_precision()
    try:
        native = import_module(native_module_name(precision))
    except ImportError:
        out = dict(_SOURCE_DEFAULT)
        if precision == "double":
            out.update(real_precision="double", real_dtype="float64")
        return out
    return _normalize(native.build_info())


_BUILD_INFO = _initial()


def build_info() -> Mapping[str, str]:
    return MappingProxyType(dict(_BUILD_INFO))


def dtype_for(key: str) -> np.dtype:
    """numpy dtype of one build_info field (python/beads/_build_config.py:20-30)."""
    return np.dtype(_BUILD_INFO[key])


def check_native_build_info(native: object) -> None:
    fn = getattr(native, "build_info", None)
    if fn is None:
        raise RuntimeError("beads_amd native extension does not provide build_info.")
    reported = _normalize(fn())
    if reported != _BUILD_INFO:
        raise RuntimeError(
            "beads_amd Python/native dtype configuration mismatch: "
            f"Python uses {_BUILD_INFO}, native reports {reported}."
        )
