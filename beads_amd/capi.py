"""ctypes view of the C ABI in include/beads_amd.h.

``Engine`` drives one run through the engine-handle entry points
(create / initialize / advance / read back) so tests and the bench can step a
run and inspect neighbour lists, cells, forces and state between steps. The
library must be built; nothing here falls back to host code.

Both builds of the engine can be loaded side by side: ``precision="single"``
(libbeads_amd.so, beads_real_t = float) and ``precision="double"``
(libbeads_amd_real64.so, built with BEADS_REAL64, beads_real_t = double).
The default is BEADS_AMD_PRECISION's ("single" when unset).
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from ._build_config import selected_precision
from .workloads import Workload

_HERE = Path(__file__).resolve().parent
_LIB_FILES = {"single": "libbeads_amd.so", "double": "libbeads_amd_real64.so"}
_REAL_TYPES = {"single": (np.dtype(np.float32), C.c_float),
               "double": (np.dtype(np.float64), C.c_double)}

STATUS_NAMES = {0: "OK", 1: "VALUE", 2: "TYPE", 3: "OVERFLOW", 4: "INDEX", 5: "RUNTIME",
                6: "INTERRUPTED", 7: "DEVICE"}

_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


def _spec_fields(real):
    _rp = C.POINTER(real)
    return [
        ("units", C.c_int32), ("n_particles", C.c_uint32),
        ("box_bound", _rp), ("positions", _rp), ("velocities", _rp), ("masses", _rp),
        ("types", _i32p), ("tags", _u32p), ("molecule_ids", _u32p), ("images", _i32p),
        ("n_bonds", C.c_uint32), ("bonds", _u32p), ("n_angles", C.c_uint32), ("angles", _u32p),
        ("n_dihedrals", C.c_uint32), ("dihedrals", _u32p),
        ("pair_style", C.c_char_p), ("bond_style", C.c_char_p), ("angle_style", C.c_char_p),
        ("dihedral_style", C.c_char_p), ("integrator_style", C.c_char_p),
        ("thermostat_style", C.c_char_p),
        ("n_types", C.c_int32), ("lj_cutoff", real), ("lj_shift", C.c_int32),
        ("lj_epsilon", _rp), ("lj_sigma", _rp),
        ("n_bond_types", C.c_int32), ("bond_k", _rp), ("bond_r0", _rp),
        ("n_angle_types", C.c_int32), ("angle_k", _rp), ("angle_theta0_deg", _rp),
        ("n_dihedral_types", C.c_int32), ("dihedral_k", _rp), ("dihedral_d", _i32p),
        ("dihedral_n", _i32p), ("exclusion_distance", C.c_int32),
        ("exclusion_policy", C.c_int32),
        ("dt", real), ("berendsen", C.c_int32), ("berendsen_temperature", real),
        ("berendsen_tau", real),
        ("cutoff_buffer", real), ("rebuild_check_every", C.c_uint64),
        ("sort_every_rebuild", C.c_uint64), ("sort_every_steps", C.c_uint64),
        ("max_neighbors", C.c_uint32),
        ("half_list", C.c_int32),
        ("thermo_path", C.c_char_p), ("thermo_every", C.c_uint64),
        ("trajectory_path", C.c_char_p), ("trajectory_every", C.c_uint64),
        ("trajectory_fields", C.c_uint32), ("final_state_path", C.c_char_p),
        ("log_echo", C.c_int32), ("log_path", C.c_char_p),
        ("runsteps", C.c_uint64),
        ("interrupt", C.c_void_p), ("interrupt_context", C.c_void_p),
    ]


class Spec(C.Structure):
    """beads_spec_t of the binary32 build."""
    _fields_ = _spec_fields(C.c_float)


class Spec64(C.Structure):
    """beads_spec_t of the BEADS_REAL64 build."""
    _fields_ = _spec_fields(C.c_double)


BATCH_CONFIGS = 9  # BEADS_BATCH_CONFIGS


class BatchConfig(C.Structure):
    """beads_batch_config_t: mode 0 persistent (one workgroup per replica),
    1 phased (per-step batched kernels over 64-atom chunks)."""
    _fields_ = [("mode", C.c_int32), ("block", C.c_int32), ("team", C.c_int32),
                ("slots", C.c_int32)]

    def as_dict(self) -> dict:
        return {"mode": "persistent" if self.mode == 0 else "phased", "block": self.block,
                "team": self.team, "slots": self.slots}


class Stats(C.Structure):
    _fields_ = [
        ("step", C.c_uint64), ("build_count", C.c_uint64), ("dangerous_rebuild_count", C.c_uint64),
        ("reorder_count", C.c_uint64), ("overflow", C.c_int32), ("path", C.c_int32),
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("max_neighbors", C.c_uint32),
        ("mean_neighbors", C.c_double), ("ke", C.c_double), ("lambda_", C.c_double),
        ("pe", C.c_double * 4), ("virial", C.c_double * 4), ("temperature", C.c_double),
        ("pressure", C.c_double),
    ]


def _state_fields(real):
    rp = C.POINTER(real)
    return [("positions", rp), ("velocities", rp), ("forces", rp), ("images", _i32p),
            ("tags", _u32p), ("types", _i32p), ("molecule_ids", _u32p), ("masses", rp)]


class StateOut(C.Structure):
    _fields_ = _state_fields(C.c_float)


class StateOut64(C.Structure):
    _fields_ = _state_fields(C.c_double)


_STRUCTS = {"single": (Spec, StateOut), "double": (Spec64, StateOut64)}


class PhaseTimes(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("total_s", C.c_double), ("integrate_s", C.c_double),
                ("rebuild_s", C.c_double), ("force_s", C.c_double),
                ("force_launches", C.c_uint64)]


class BuildInfo(C.Structure):
    _fields_ = [(k, C.c_char_p) for k in ("real_precision", "real_dtype", "type_dtype",
                                          "index_dtype", "image_dtype", "runstep_dtype",
                                          "device_arch")]


def signatures(precision: str = "single") -> dict:
    """Every symbol include/beads_amd.h declares, with its ctypes signature in
    the given build."""
    spec_t, state_t = _STRUCTS[precision]
    real = _REAL_TYPES[precision][1]
    Spec, StateOut, _rp = spec_t, state_t, C.POINTER(real)  # noqa: N806
    return {
        "beads_last_error": (C.c_char_p, []),
        "beads_abi_version": (C.c_int, []),
        "beads_build_info": (C.c_int, [C.POINTER(BuildInfo)]),
        "beads_execute_simulation": (C.c_int, [C.POINTER(Spec)]),
        "beads_engine_create": (C.c_int, [C.POINTER(Spec), C.POINTER(C.c_void_p)]),
        "beads_engine_initialize": (C.c_int, [C.c_void_p, C.c_int32]),
        "beads_engine_advance": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32]),
        "beads_engine_synchronize": (C.c_int, [C.c_void_p]),
        "beads_engine_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
        "beads_engine_read_state": (C.c_int, [C.c_void_p, C.POINTER(StateOut)]),
        "beads_engine_read_neighbors": (C.c_int, [C.c_void_p, _u32p, _u32p, C.c_uint64]),
        "beads_engine_read_cells": (C.c_int, [C.c_void_p, _u32p, _u32p, _u32p, _u32p, C.c_uint64]),
        "beads_engine_profile_advance": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32,
                                                   C.POINTER(PhaseTimes)]),
        "beads_engine_set_lj_params": (C.c_int, [C.c_void_p, _rp, _rp]),
        "beads_engine_set_bond_params": (C.c_int, [C.c_void_p, _rp, _rp]),
        "beads_engine_set_angle_params": (C.c_int, [C.c_void_p, _rp, _rp]),
        "beads_engine_set_dihedral_params": (C.c_int, [C.c_void_p, _rp]),
        "beads_engine_adjoint_param_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
        "beads_engine_frame_forces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.POINTER(C.c_double), C.c_int32]),
        "beads_engine_frame_load": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
        "beads_engine_frame_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                                  C.c_void_p]),
        "beads_engine_frame_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double)]),
        "beads_engine_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
        "beads_engine_frame_forces_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                      C.c_int32]),
        "beads_engine_frame_load_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
        "beads_engine_frame_backward_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double,
                                                        C.c_void_p, C.c_void_p]),
        "beads_engine_frame_backward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double,
                                                          C.c_void_p, C.c_void_p]),
        "beads_engine_destroy": (None, [C.c_void_p]),
        "beads_batch_create": (C.c_int, [C.POINTER(Spec), C.c_uint32, C.POINTER(C.c_void_p)]),
        "beads_batch_config_info": (C.c_int, [C.c_uint32, C.POINTER(BatchConfig)]),
        "beads_batch_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_int32),
                                           C.POINTER(C.c_double)]),
        "beads_batch_set_config": (C.c_int, [C.c_void_p, C.c_uint32]),
        "beads_batch_get_config": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
        "beads_batch_set_block": (C.c_int, [C.c_void_p, C.c_int32]),
        "beads_batch_run": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32]),
        "beads_batch_profile": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
        "beads_batch_synchronize": (C.c_int, [C.c_void_p]),
        "beads_batch_replica_stats": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(Stats)]),
        "beads_batch_read_state": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(StateOut)]),
        "beads_batch_destroy": (None, [C.c_void_p]),
    }


SIGNATURES = signatures()

_libs: dict[str, C.CDLL] = {}


def library(precision: str | None = None) -> C.CDLL:
    """Load the engine library of one precision (raises if it has not been built)."""
    precision = precision or selected_precision()
    lib = _libs.get(precision)
    if lib is None:
        path = _HERE / _LIB_FILES[precision]
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build().")
        lib = C.CDLL(os.fspath(path))
        for name, (res, args) in signatures(precision).items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        info = BuildInfo()
        if lib.beads_build_info(C.byref(info)) != 0:
            raise RuntimeError(f"{path}: beads_build_info failed.")
        expected = _REAL_TYPES[precision][0].name
        if info.real_dtype.decode() != expected:
            raise RuntimeError(f"{path} reports real_dtype {info.real_dtype.decode()!r}, "
                               f"expected {expected!r}.")
        _libs[precision] = lib
    return lib


class BeadsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{STATUS_NAMES.get(status, status)}] {message}")
        self.status = status


def _check(status: int, lib: C.CDLL | None = None) -> None:
    if status != 0:
        raise BeadsError(status, (lib or library()).beads_last_error().decode())


def _ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def make_spec(w: Workload, precision: str | None = None):
    """Fill the beads_spec_t of one build from a Workload; the returned list
    keeps its arrays alive. Real arrays take the build's real dtype."""
    precision = precision or selected_precision()
    real_dtype, real_ct = _REAL_TYPES[precision]
    keep = []

    def arr(a, dtype):
        out = np.ascontiguousarray(a, dtype=real_dtype if dtype is np.float32 else dtype)
        keep.append(out)
        return out

    s = _STRUCTS[precision][0]()
    s.units = 0 if w.units == "reduced" else 1
    s.n_particles = w.n
    s.box_bound = _ptr(arr(w.box_bound, np.float32), real_ct)
    s.positions = _ptr(arr(w.positions, np.float32), real_ct)
    s.velocities = _ptr(arr(w.velocities, np.float32), real_ct)
    s.masses = _ptr(arr(w.masses, np.float32), real_ct)
    s.types = _ptr(arr(w.types, np.int32), C.c_int32)
    s.tags = _ptr(arr(w.tags, np.uint32), C.c_uint32)
    s.molecule_ids = _ptr(arr(w.molecule_ids, np.uint32), C.c_uint32)
    s.images = _ptr(arr(w.images, np.int32), C.c_int32)
    s.n_bonds = len(w.bonds)
    s.bonds = _ptr(arr(w.bonds.reshape(-1, 3), np.uint32), C.c_uint32)
    s.n_angles = len(w.angles)
    s.angles = _ptr(arr(w.angles.reshape(-1, 4), np.uint32), C.c_uint32)
    s.n_dihedrals = len(w.dihedrals)
    s.dihedrals = _ptr(arr(w.dihedrals.reshape(-1, 5), np.uint32), C.c_uint32)
    s.n_types = w.n_types
    s.lj_cutoff = w.cutoff
    s.lj_shift = 1 if w.shift else 0
    s.lj_epsilon = _ptr(arr(w.lj_epsilon, np.float32), real_ct)
    s.lj_sigma = _ptr(arr(w.lj_sigma, np.float32), real_ct)
    s.n_bond_types = len(w.bond_k) if len(w.bonds) else 0
    s.bond_k = _ptr(arr(w.bond_k, np.float32), real_ct)
    s.bond_r0 = _ptr(arr(w.bond_r0, np.float32), real_ct)
    s.n_angle_types = len(w.angle_k) if len(w.angles) else 0
    s.angle_k = _ptr(arr(w.angle_k, np.float32), real_ct)
    s.angle_theta0_deg = _ptr(arr(w.angle_theta0, np.float32), real_ct)
    s.n_dihedral_types = len(w.dihedral_k) if len(w.dihedrals) else 0
    s.dihedral_k = _ptr(arr(w.dihedral_k, np.float32), real_ct)
    s.dihedral_d = _ptr(arr(w.dihedral_d, np.int32), C.c_int32)
    s.dihedral_n = _ptr(arr(w.dihedral_n, np.int32), C.c_int32)
    s.exclusion_distance = int(w.exclusion_distance) if len(w.bonds) else 0
    s.exclusion_policy = 1 if s.exclusion_distance else 0  # the default policy's resolution
    s.dt = w.dt
    s.berendsen = 1 if w.berendsen else 0
    s.berendsen_temperature = w.temperature
    s.berendsen_tau = w.tau
    s.cutoff_buffer = w.cutoff_buffer
    s.rebuild_check_every = w.rebuild_check_every
    s.sort_every_rebuild = w.sort_every_rebuild
    s.sort_every_steps = w.sort_every_steps
    s.max_neighbors = w.max_neighbors
    s.half_list = 1 if w.half_list else 0
    s.log_echo = 3
    s.runsteps = w.steps
    return s, keep


class Engine:
    """One run on the device, stepped from Python."""

    def __init__(self, workload: Workload, precision: str | None = None):
        self.precision = precision or selected_precision()
        self.lib = library(self.precision)
        self.real, self._real_ct = _REAL_TYPES[self.precision]
        self.workload = workload
        self.n = workload.n
        self._spec, self._keep = make_spec(workload, self.precision)
        handle = C.c_void_p()
        self._check(self.lib.beads_engine_create(C.byref(self._spec), C.byref(handle)))
        self.handle = handle

    def _check(self, status: int) -> None:
        _check(status, self.lib)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.beads_engine_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self, observables: bool = True) -> None:
        self._check(self.lib.beads_engine_initialize(self.handle, 1 if observables else 0))

    def advance(self, steps: int, observables_on_last: bool = False) -> None:
        self._check(self.lib.beads_engine_advance(self.handle, steps, 1 if observables_on_last else 0))

    def synchronize(self) -> None:
        self._check(self.lib.beads_engine_synchronize(self.handle))

    def stats(self) -> Stats:
        s = Stats()
        self._check(self.lib.beads_engine_stats(self.handle, C.byref(s)))
        return s

    def state(self) -> dict[str, np.ndarray]:
        out, so = _state_buffers(self.n, self.precision)
        self._check(self.lib.beads_engine_read_state(self.handle, C.byref(so)))
        return out

    def neighbors(self) -> tuple[np.ndarray, np.ndarray]:
        """(counts[n], list[max_neighbors][n]) in the engine's slot-major layout."""
        m = self.workload.max_neighbors
        counts = np.zeros(self.n, np.uint32)
        lst = np.zeros((m, self.n), np.uint32)
        self._check(self.lib.beads_engine_read_neighbors(self.handle, _ptr(counts, C.c_uint32),
                                                    _ptr(lst, C.c_uint32), lst.size))
        return counts, lst

    def cells(self, n_cells: int) -> dict[str, np.ndarray]:
        out = {"cell_of": np.zeros(self.n, np.uint32), "cell_atoms": np.zeros(self.n, np.uint32),
               "cell_start": np.zeros(n_cells, np.uint32), "cell_count": np.zeros(n_cells, np.uint32)}
        self._check(self.lib.beads_engine_read_cells(
            self.handle, _ptr(out["cell_of"], C.c_uint32), _ptr(out["cell_atoms"], C.c_uint32),
            _ptr(out["cell_start"], C.c_uint32), _ptr(out["cell_count"], C.c_uint32), n_cells))
        return out

    def profile(self, steps: int, per_phase: bool = False) -> PhaseTimes:
        t = PhaseTimes()
        self._check(self.lib.beads_engine_profile_advance(self.handle, steps, 1 if per_phase else 0,
                                                     C.byref(t)))
        return t

    # ---- frame mode (device pointers; see beads_engine_frame_forces) ----

    def set_lj_params(self, epsilon: np.ndarray, sigma: np.ndarray) -> None:
        e = np.ascontiguousarray(epsilon, dtype=self.real)
        s = np.ascontiguousarray(sigma, dtype=self.real)
        self._check(self.lib.beads_engine_set_lj_params(self.handle, _ptr(e, self._real_ct),
                                                        _ptr(s, self._real_ct)))

    def set_bond_params(self, k: np.ndarray, r0: np.ndarray) -> None:
        kk = np.ascontiguousarray(k, dtype=self.real)
        rr = np.ascontiguousarray(r0, dtype=self.real)
        self._check(self.lib.beads_engine_set_bond_params(self.handle, _ptr(kk, self._real_ct),
                                                          _ptr(rr, self._real_ct)))

    def set_angle_params(self, k: np.ndarray, theta0_deg: np.ndarray) -> None:
        kk = np.ascontiguousarray(k, dtype=self.real)
        tt = np.ascontiguousarray(theta0_deg, dtype=self.real)
        self._check(self.lib.beads_engine_set_angle_params(self.handle, _ptr(kk, self._real_ct),
                                                           _ptr(tt, self._real_ct)))

    def set_dihedral_params(self, k: np.ndarray) -> None:
        kk = np.ascontiguousarray(k, dtype=self.real)
        self._check(self.lib.beads_engine_set_dihedral_params(self.handle,
                                                              _ptr(kk, self._real_ct)))

    def adjoint_param_count(self) -> int:
        c = C.c_uint64()
        self._check(self.lib.beads_engine_adjoint_param_count(self.handle, C.byref(c)))
        return c.value

    def frame_forces(self, positions_ptr: int, forces_ptr: int,
                     reuse_list: bool = False) -> np.ndarray:
        """Forces at device positions; returns [pair, bond, angle, dihedral] energies."""
        e = np.zeros(4, np.float64)
        self._check(self.lib.beads_engine_frame_forces(self.handle, C.c_void_p(positions_ptr),
                                                  C.c_void_p(forces_ptr),
                                                  _ptr(e, C.c_double), 1 if reuse_list else 0))
        return e

    def frame_load(self, positions_ptr: int, reuse_list: bool = True) -> None:
        self._check(self.lib.beads_engine_frame_load(self.handle, C.c_void_p(positions_ptr),
                                                1 if reuse_list else 0))

    def frame_backward(self, grad_forces_ptr: int, grad_energy: float, grad_positions_ptr: int,
                       grad_params_ptr: int) -> None:
        self._check(self.lib.beads_engine_frame_backward(
            self.handle, C.c_void_p(grad_forces_ptr), float(grad_energy),
            C.c_void_p(grad_positions_ptr), C.c_void_p(grad_params_ptr)))

    # Stream-ordered frame mode: the calls below enqueue on the stream given to
    # set_stream (a hipStream_t as an int; 0 = HIP's null stream) and return at
    # once; synchronize() waits and reports a list overflow.
    def set_stream(self, stream_handle: int) -> None:
        self._check(self.lib.beads_engine_set_stream(self.handle, C.c_void_p(stream_handle or None)))

    def frame_forces_async(self, positions_ptr: int, forces_ptr: int, energies_ptr: int,
                           reuse_list: bool = True) -> None:
        """energies_ptr: device memory for [pair, bond, angle, dihedral] (FP64)."""
        self._check(self.lib.beads_engine_frame_forces_async(
            self.handle, C.c_void_p(positions_ptr), C.c_void_p(forces_ptr),
            C.c_void_p(energies_ptr or None), 1 if reuse_list else 0))

    def frame_load_async(self, positions_ptr: int, reuse_list: bool = True) -> None:
        self._check(self.lib.beads_engine_frame_load_async(self.handle, C.c_void_p(positions_ptr),
                                                      1 if reuse_list else 0))

    def frame_backward_async(self, grad_forces_ptr: int, grad_energy: float,
                             grad_positions_ptr: int, grad_params_ptr: int) -> None:
        self._check(self.lib.beads_engine_frame_backward_async(
            self.handle, C.c_void_p(grad_forces_ptr), float(grad_energy),
            C.c_void_p(grad_positions_ptr), C.c_void_p(grad_params_ptr)))

    def frame_backward_noalloc(self, grad_forces_ptr: int, grad_energy: float,
                               grad_positions_ptr: int, grad_params_ptr: int) -> None:
        """frame_backward_async without any device allocation (static scratch)."""
        self._check(self.lib.beads_engine_frame_backward_noalloc(
            self.handle, C.c_void_p(grad_forces_ptr), float(grad_energy),
            C.c_void_p(grad_positions_ptr), C.c_void_p(grad_params_ptr)))

    def frame_timing(self) -> tuple[float, float]:
        f, b = C.c_double(), C.c_double()
        self._check(self.lib.beads_engine_frame_timing(self.handle, C.byref(f), C.byref(b)))
        return f.value, b.value


def _state_buffers(n: int, precision: str):
    real, real_ct = _REAL_TYPES[precision]
    out = {"positions": np.zeros((n, 3), real), "velocities": np.zeros((n, 3), real),
           "forces": np.zeros((n, 3), real), "images": np.zeros((n, 3), np.int32),
           "tags": np.zeros(n, np.uint32), "types": np.zeros(n, np.int32),
           "molecule_ids": np.zeros(n, np.uint32), "masses": np.zeros(n, real)}
    so = _STRUCTS[precision][1](
        _ptr(out["positions"], real_ct), _ptr(out["velocities"], real_ct),
        _ptr(out["forces"], real_ct), _ptr(out["images"], C.c_int32),
        _ptr(out["tags"], C.c_uint32), _ptr(out["types"], C.c_int32),
        _ptr(out["molecule_ids"], C.c_uint32), _ptr(out["masses"], real_ct))
    return out, so


class Batch:
    """Many systems (LJ pairs, harmonic bonds, angles and dihedrals, bonded
    exclusions) advanced together, either by one persistent
    launch (one workgroup per replica) or by per-step batched kernels over
    64-atom chunks of every replica (phased)."""

    def __init__(self, workloads: list[Workload], precision: str | None = None):
        self.precision = precision or selected_precision()
        self.lib = library(self.precision)
        self.workloads = list(workloads)
        specs = (_STRUCTS[self.precision][0] * len(self.workloads))()
        self._keep = []
        for k, w in enumerate(self.workloads):
            s, keep = make_spec(w, self.precision)
            specs[k] = s
            self._keep.append(keep)
        self._specs = specs
        handle = C.c_void_p()
        self._check(self.lib.beads_batch_create(specs, len(self.workloads), C.byref(handle)))
        self.handle = handle

    def _check(self, status: int) -> None:
        _check(status, self.lib)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.beads_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config_info(self, index: int) -> dict:
        c = BatchConfig()
        self._check(self.lib.beads_batch_config_info(index, C.byref(c)))
        return c.as_dict()

    def configs(self) -> list[dict]:
        return [self.config_info(k) for k in range(BATCH_CONFIGS)]

    def autotune(self, trial_steps: int = 20) -> tuple[int, list[float]]:
        """Times every configuration from the initial state, keeps the fastest
        (returned as its index) and restores the initial state."""
        chosen = C.c_int32()
        secs = (C.c_double * BATCH_CONFIGS)()
        self._check(self.lib.beads_batch_autotune(self.handle, trial_steps, C.byref(chosen), secs))
        return chosen.value, list(secs)

    def set_config(self, index: int) -> None:
        self._check(self.lib.beads_batch_set_config(self.handle, index))

    def config(self) -> int:
        k = C.c_uint32()
        self._check(self.lib.beads_batch_get_config(self.handle, C.byref(k)))
        return k.value

    def set_block(self, block: int) -> None:
        self._check(self.lib.beads_batch_set_block(self.handle, block))

    def run(self, steps: int, observables_on_last: bool = False) -> None:
        self._check(self.lib.beads_batch_run(self.handle, steps, 1 if observables_on_last else 0))

    def profile(self, steps: int) -> float:
        t = C.c_double()
        self._check(self.lib.beads_batch_profile(self.handle, steps, C.byref(t)))
        return t.value

    def synchronize(self) -> None:
        self._check(self.lib.beads_batch_synchronize(self.handle))

    def stats(self, r: int) -> Stats:
        s = Stats()
        self._check(self.lib.beads_batch_replica_stats(self.handle, r, C.byref(s)))
        return s

    def state(self, r: int) -> dict[str, np.ndarray]:
        out, so = _state_buffers(self.workloads[r].n, self.precision)
        self._check(self.lib.beads_batch_read_state(self.handle, r, C.byref(so)))
        return out
