"""The C-ABI libraries (binary32 and BEADS_REAL64 builds) load and export
every entry point include/beads_amd.h declares; the pybind modules report the
reference's build_info contract (module.cpp:15-27); the ctypes structs match
the C compiler's layout of the header in both precisions. No device calls
(CPU only)."""

from __future__ import annotations

import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import pytest

import beads_amd
from beads_amd import capi

PRECISIONS = ("single", "double")

HEADER = Path(__file__).resolve().parents[1] / "include" / "beads_amd.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(beads_[a-z_]+)\s*\(", text)) - {"beads_interrupt_fn"}


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert {"beads_execute_simulation", "beads_build_info", "beads_engine_create",
            "beads_engine_advance", "beads_engine_destroy"} <= names


@pytest.mark.parametrize("precision", PRECISIONS)
def test_library_exports_every_declared_symbol(precision):
    lib = ctypes.CDLL(str(capi._HERE / capi._LIB_FILES[precision]))
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(capi.signatures(precision)) == declared_functions()


@pytest.mark.parametrize("precision,dtype", [("single", b"float32"), ("double", b"float64")])
def test_abi_version_and_build_info(precision, dtype):
    lib = capi.library(precision)
    assert lib.beads_abi_version() == 6
    info = capi.BuildInfo()
    assert lib.beads_build_info(ctypes.byref(info)) == 0
    assert info.real_precision == precision.encode()
    assert info.real_dtype == dtype and info.device_arch == b"gfx950"


@pytest.mark.parametrize("module,precision,dtype", [("_native", "single", "float32"),
                                                    ("_native_real64", "double", "float64")])
def test_native_module_build_info_matches_reference_contract(module, precision, dtype):
    import importlib

    native = importlib.import_module(f"beads_amd.{module}")
    assert native.build_info() == {
        "real_precision": precision, "real_dtype": dtype, "type_dtype": "int32",
        "index_dtype": "uint32", "image_dtype": "int32", "runstep_dtype": "uint64"}


def test_package_build_info_is_the_selected_modules():
    import importlib

    from beads_amd._build_config import native_module_name

    native = importlib.import_module(native_module_name())
    assert dict(beads_amd.build_info()) == native.build_info()


def test_spec_struct_matches_header_field_order():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = text[text.index("typedef struct beads_spec {"):text.index("} beads_spec_t;")]
    fields = re.findall(r"\b([a-z_0-9]+)\s*;", body)
    assert fields == [f for f, _ in capi.Spec._fields_]
    assert fields == [f for f, _ in capi.Spec64._fields_]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
@pytest.mark.parametrize("precision", PRECISIONS)
def test_struct_layouts_match_the_c_compiler(precision, tmp_path):
    """sizeof and every field offset of beads_spec_t and beads_state_out_t as
    gcc lays out the header (with BEADS_REAL64 for the double build) equal the
    ctypes structs capi passes across the boundary."""
    spec_t, state_t = capi._STRUCTS[precision]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "beads_amd.h"',
             "int main(void) {"]
    for cname, st in (("beads_spec_t", spec_t), ("beads_state_out_t", state_t)):
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in st._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    flags = ["-DBEADS_REAL64"] if precision == "double" else []
    subprocess.run(["gcc", *flags, f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    expected = []
    for cname, st in (("beads_spec_t", spec_t), ("beads_state_out_t", state_t)):
        expected.append(f"{cname} size {ctypes.sizeof(st)}")
        expected += [f"{cname} {f} {getattr(st, f).offset}" for f, _ in st._fields_]
    assert [ln for ln in out if ln] == expected


@pytest.mark.parametrize("precision,limit", [("single", (1 << 32) // 16 - 1),
                                             ("double", (1 << 32) // 32 - 1)])
def test_particle_count_limit_of_the_gather_offsets(precision, limit):
    """The force pass gathers positions by a 32-bit byte offset: a spec above
    2^32 / sizeof(real4) - 1 particles is refused (BEADS_E_OVERFLOW) before
    any array is read or any device call is made."""
    from beads_amd import workloads as W

    w = W.lj_box(100, 0.8, 1.0, seed=1)
    lib = capi.library(precision)
    spec, keep = capi.make_spec(w, precision)
    spec.n_particles = limit + 1
    handle = ctypes.c_void_p()
    assert lib.beads_engine_create(ctypes.byref(spec), ctypes.byref(handle)) == 3
    assert lib.beads_last_error().decode() == (
        f"system.n_particles must be at most {limit} with {precision}-precision storage.")
    assert not handle.value
    del keep


@pytest.mark.parametrize("precision", PRECISIONS)
def test_engine_rejects_invalid_spec_without_device_work(precision):
    from beads_amd import workloads as W

    w = W.lj_box(100, 0.8, 1.0, seed=1).with_(cutoff=50.0)  # r_s above L/2
    spec, keep = capi.make_spec(w, precision)
    handle = ctypes.c_void_p()
    lib = capi.library(precision)
    status = lib.beads_engine_create(ctypes.byref(spec), ctypes.byref(handle))
    assert status == 1  # BEADS_E_VALUE
    assert b"half shortest box length" in lib.beads_last_error()
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


def test_batch_rejects_more_than_255_records_per_particle_before_any_device_call():
    """Batched replicas keep every listed-term record inline (no spill
    segments), at most 255 of a kind per particle: a replica whose hub bead
    has 300 bonds is BEADS_E_VALUE at create, raised while the specs are
    packed (no device)."""
    from gpu_cases import hub_lattice

    from beads_amd import workloads as W

    lib = capi.library()
    ws = [W.lj_box(300, 0.8, 1.0, seed=1), hub_lattice(300, 1, 2, False)]
    specs = (capi.Spec * len(ws))()
    keep = []
    for k, w in enumerate(ws):
        specs[k], kk = capi.make_spec(w)
        keep.append(kk)
    handle = ctypes.c_void_p()
    status = lib.beads_batch_create(specs, len(ws), ctypes.byref(handle))
    assert status == 1  # BEADS_E_VALUE
    assert lib.beads_last_error().decode() == \
        "batched replicas keep at most 255 bonds per particle; a particle has 300."
    assert not handle.value
