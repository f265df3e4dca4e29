"""The C-ABI library loads and exports every entry point include/beads_amd.h
declares; the pybind module reports the reference's build_info contract.
No device calls (CPU only)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import beads_amd
from beads_amd import capi

HEADER = Path(__file__).resolve().parents[1] / "include" / "beads_amd.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(beads_[a-z_]+)\s*\(", text)) - {"beads_interrupt_fn"}


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert {"beads_execute_simulation", "beads_build_info", "beads_engine_create",
            "beads_engine_advance", "beads_engine_destroy"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(capi._LIB_PATH))
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(capi.SIGNATURES) == declared_functions()


def test_abi_version_and_build_info():
    lib = capi.library()
    assert lib.beads_abi_version() == 3
    info = capi.BuildInfo()
    assert lib.beads_build_info(ctypes.byref(info)) == 0
    assert info.real_dtype == b"float32" and info.device_arch == b"gfx950"


def test_native_module_build_info_matches_reference_contract():
    from beads_amd import _native

    assert _native.build_info() == {
        "real_precision": "single", "real_dtype": "float32", "type_dtype": "int32",
        "index_dtype": "uint32", "image_dtype": "int32", "runstep_dtype": "uint64"}
    assert dict(beads_amd.build_info()) == _native.build_info()


def test_spec_struct_matches_header_field_order():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = text[text.index("typedef struct beads_spec {"):text.index("} beads_spec_t;")]
    fields = re.findall(r"\b([a-z_0-9]+)\s*;", body)
    assert fields == [f for f, _ in capi.Spec._fields_]


def test_engine_rejects_invalid_spec_without_device_work():
    from beads_amd import workloads as W

    w = W.lj_box(100, 0.8, 1.0, seed=1).with_(cutoff=50.0)  # r_s above L/2
    spec, keep = capi.make_spec(w)
    handle = ctypes.c_void_p()
    status = capi.library().beads_engine_create(ctypes.byref(spec), ctypes.byref(handle))
    assert status == 1  # BEADS_E_VALUE
    assert b"half shortest box length" in capi.library().beads_last_error()
    del keep


def test_engine_rejects_unregistered_styles_before_any_device_call():
    """Models come from per-kind registries keyed by style name
    (pair_factory.cpp:30-47): an unknown name is BEADS_E_VALUE with the
    reference's message, raised while the spec is validated (no device)."""
    from beads_amd import workloads as W

    w = W.polymer_melt(n_chains=2, beads_per_chain=10, seed=1)
    lib = capi.library()
    for field, kind in (("pair_style", "pair"), ("bond_style", "bond"), ("angle_style", "angle"),
                        ("dihedral_style", "dihedral"), ("integrator_style", "integrator"),
                        ("thermostat_style", "thermostat")):
        spec, keep = capi.make_spec(w)
        setattr(spec, field, b"nope")
        handle = ctypes.c_void_p()
        status = lib.beads_engine_create(ctypes.byref(spec), ctypes.byref(handle))
        assert status == 1, field  # BEADS_E_VALUE
        assert lib.beads_last_error().decode() == f'unsupported {kind} style "nope".'
        assert not handle.value
