"""ctypes wrapper of the FP64 host oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. It is the checker, never the thing measured or
shipped. Functions take a ``beads_amd.workloads.Workload`` (or its arrays) and
return numpy arrays; see oracle.h for the semantics of each call.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

PATH_DIRECT, PATH_REGULAR, PATH_SMALL = 0, 1, 2

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


class Plan(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("n_cells", C.c_uint32),
                ("cell_size", C.c_float * 3), ("inv_cell", C.c_float * 3),
                ("force_cutoff", C.c_float), ("search_cutoff", C.c_float),
                ("search_cutoff_sq", C.c_float), ("path", C.c_int32)]


class PlanF64(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("n_cells", C.c_uint32),
                ("cell_size", C.c_double * 3), ("inv_cell", C.c_double * 3),
                ("force_cutoff", C.c_double), ("search_cutoff", C.c_double),
                ("search_cutoff_sq", C.c_double), ("path", C.c_int32)]


class System(C.Structure):
    _fields_ = [("n", C.c_uint32), ("pos", _f32p), ("vel", _f32p), ("mass", _f32p), ("type", _i32p),
                ("tag", _u32p), ("image", _i32p), ("lo", C.c_float * 3), ("hi", C.c_float * 3),
                ("n_types", C.c_int32), ("lj_eps", _f32p), ("lj_sig", _f32p),
                ("lj_cutoff", C.c_float), ("lj_shift", C.c_int32),
                ("n_bonds", C.c_uint32), ("bonds", _u32p), ("bond_k", _f32p), ("bond_r0", _f32p),
                ("n_angles", C.c_uint32), ("angles", _u32p), ("angle_k", _f32p),
                ("angle_theta0_deg", _f32p), ("n_dihedrals", C.c_uint32), ("dihedrals", _u32p),
                ("dihedral_k", _f32p), ("dihedral_d", _i32p), ("dihedral_n", _i32p),
                ("exclusion_distance", C.c_int32)]


class Observables(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("pair_pe", "bond_pe", "angle_pe", "dihedral_pe",
                                          "pair_virial", "bond_virial", "angle_virial",
                                          "dihedral_virial")]

    def as_dict(self) -> dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


class RunParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("thermostat", C.c_int32), ("temperature", C.c_double),
                ("tau", C.c_double), ("kb", C.c_double), ("cutoff_buffer", C.c_float),
                ("rebuild_check_every", C.c_uint64), ("sort_every_rebuild", C.c_uint64),
                ("sort_every_steps", C.c_uint64),
                ("max_neighbors", C.c_uint32), ("steps", C.c_uint64), ("threads", C.c_int32)]


class RunStats(C.Structure):
    _fields_ = [("build_count", C.c_uint64), ("dangerous_rebuild_count", C.c_uint64),
                ("wall_seconds", C.c_double), ("mean_neighbors", C.c_double),
                ("threads_used", C.c_int32), ("final_ke", C.c_double),
                ("final_obs", Observables), ("reorder_count", C.c_uint64)]


_lib = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-C", os.fspath(HERE), "-s"], check=True)


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(os.fspath(LIB_PATH))
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_plan.argtypes = [_f32p, _f32p, C.c_float, C.c_float, C.c_uint32, C.POINTER(Plan)]
        lib.orc_cell_ids.argtypes = [C.c_uint32, _f32p, _f32p, C.POINTER(Plan), _u32p]
        lib.orc_stable_cell_order.argtypes = [C.c_uint32, _u32p, C.c_uint32, _f32p, _u32p, _u32p,
                                              _u32p]
        lib.orc_exclusion_pairs.restype = C.c_uint64
        lib.orc_exclusion_pairs.argtypes = [C.c_uint32, C.c_uint32, _u32p, C.c_int32, _u32p, C.c_uint64]
        lib.orc_neighbor_build.argtypes = [C.POINTER(System), C.c_float, C.c_uint32, C.c_int32, _u32p,
                                           _u32p, _i32p]
        lib.orc_forces.argtypes = [C.POINTER(System), _u32p, _u32p, C.c_uint32, _f64p,
                                   C.POINTER(Observables)]
        lib.orc_forces_f64.argtypes = [C.POINTER(System), _f64p, _f64p, C.POINTER(Observables)]
        lib.orc_vv_pre.argtypes = [C.c_uint32, _f32p, _f32p, _f32p, _f32p, _i32p, _f32p, _f32p,
                                   C.c_double, _f64p, _f64p, _i32p]
        lib.orc_vv_post.argtypes = [C.c_uint32, _f32p, _f32p, _f32p, C.c_double, _f64p]
        lib.orc_kinetic_energy.restype = C.c_double
        lib.orc_kinetic_energy.argtypes = [C.c_uint32, _f32p, _f32p]
        lib.orc_berendsen_lambda.restype = C.c_double
        lib.orc_berendsen_lambda.argtypes = [C.c_double, C.c_uint32, C.c_double, C.c_double,
                                             C.c_double, C.c_double]
        lib.orc_displacement_exceeds.argtypes = [C.c_uint32, _f32p, _i32p, _f32p, _i32p, _f32p,
                                                 _f32p, C.c_float]
        lib.orc_run.argtypes = [C.POINTER(System), C.POINTER(RunParams), _f64p, _f64p, _i32p, _u32p,
                                C.POINTER(RunStats)]
        lib.orc_plan_f64.argtypes = [_f32p, _f32p, C.c_double, C.c_double, C.c_uint32,
                                     C.POINTER(PlanF64)]
        lib.orc_cell_ids_f64.argtypes = [C.c_uint32, _f64p, _f32p, C.POINTER(PlanF64), _u32p]
        lib.orc_neighbor_build_f64.argtypes = [C.POINTER(System), _f64p, C.c_double, C.c_uint32,
                                               C.c_int32, _u32p, _u32p, _i32p]
        lib.orc_forces_f64_list.argtypes = [C.POINTER(System), _f64p, _u32p, _u32p, C.c_uint32,
                                            _f64p, C.POINTER(Observables)]
        lib.orc_run_tangent_f64.argtypes = [C.POINTER(System), C.POINTER(RunParams), _f64p, _f64p,
                                            _f64p, _f64p, _f64p, _f64p, _i32p, _u32p, _f64p,
                                            _f64p, C.POINTER(RunStats)]
        lib.orc_run_f64.argtypes = [C.POINTER(System), C.POINTER(RunParams), _f64p, _f64p, _f64p,
                                    _f64p, _i32p, _u32p, C.POINTER(RunStats)]
        _lib = lib
    return _lib


class OracleError(RuntimeError):
    pass


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


class OracleSystem:
    """Keeps the arrays of an orc_system_t alive."""

    def __init__(self, w, positions: np.ndarray | None = None, tags=None, types=None):
        self.w = w
        self.keep = []

        def a(x, dt):
            out = np.ascontiguousarray(x, dtype=dt)
            self.keep.append(out)
            return out

        s = System()
        pos = w.positions if positions is None else positions
        s.n = w.n
        s.pos = _p(a(pos, np.float32), C.c_float)
        s.vel = _p(a(w.velocities, np.float32), C.c_float)
        s.mass = _p(a(w.masses, np.float32), C.c_float)
        s.type = _p(a(w.types if types is None else types, np.int32), C.c_int32)
        s.tag = _p(a(w.tags if tags is None else tags, np.uint32), C.c_uint32)
        s.image = _p(a(w.images, np.int32), C.c_int32)
        for k in range(3):
            s.lo[k] = float(w.box_bound[0, k])
            s.hi[k] = float(w.box_bound[1, k])
        s.n_types = w.n_types
        s.lj_eps = _p(a(w.lj_epsilon, np.float32), C.c_float)
        s.lj_sig = _p(a(w.lj_sigma, np.float32), C.c_float)
        s.lj_cutoff = w.cutoff
        s.lj_shift = 1 if w.shift else 0
        s.n_bonds = len(w.bonds)
        s.bonds = _p(a(w.bonds.reshape(-1, 3), np.uint32), C.c_uint32)
        s.bond_k = _p(a(w.bond_k, np.float32), C.c_float)
        s.bond_r0 = _p(a(w.bond_r0, np.float32), C.c_float)
        s.n_angles = len(w.angles)
        s.angles = _p(a(w.angles.reshape(-1, 4), np.uint32), C.c_uint32)
        s.angle_k = _p(a(w.angle_k, np.float32), C.c_float)
        s.angle_theta0_deg = _p(a(w.angle_theta0, np.float32), C.c_float)
        s.n_dihedrals = len(w.dihedrals)
        s.dihedrals = _p(a(w.dihedrals.reshape(-1, 5), np.uint32), C.c_uint32)
        s.dihedral_k = _p(a(w.dihedral_k, np.float32), C.c_float)
        s.dihedral_d = _p(a(w.dihedral_d, np.int32), C.c_int32)
        s.dihedral_n = _p(a(w.dihedral_n, np.int32), C.c_int32)
        s.exclusion_distance = int(w.exclusion_distance) if len(w.bonds) else 0
        self.s = s


def plan(w) -> Plan:
    lib = library()
    p = Plan()
    lo = np.ascontiguousarray(w.box_bound[0], np.float32)
    hi = np.ascontiguousarray(w.box_bound[1], np.float32)
    if lib.orc_plan(_p(lo, C.c_float), _p(hi, C.c_float), w.cutoff, w.cutoff_buffer, w.n, C.byref(p)):
        raise OracleError(lib.orc_last_error().decode())
    return p


def cell_ids(w, positions: np.ndarray | None = None) -> np.ndarray:
    lib = library()
    p = plan(w)
    pos = np.ascontiguousarray(w.positions if positions is None else positions, np.float32)
    lo = np.ascontiguousarray(w.box_bound[0], np.float32)
    out = np.zeros(w.n, np.uint32)
    lib.orc_cell_ids(w.n, _p(pos, C.c_float), _p(lo, C.c_float), C.byref(p), _p(out, C.c_uint32))
    return out


def stable_cell_order(cells: np.ndarray, n_cells: int, positions: np.ndarray | None = None):
    """Counting sort by cell; members of a cell by (x, slot) when positions
    are given, else by slot. Returns (order, counts, offsets)."""
    lib = library()
    cells = np.ascontiguousarray(cells, np.uint32)
    order = np.zeros(cells.size, np.uint32)
    counts = np.zeros(n_cells, np.uint32)
    offsets = np.zeros(n_cells, np.uint32)
    pos = None
    if positions is not None:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    lib.orc_stable_cell_order(cells.size, _p(cells, C.c_uint32), n_cells,
                              _p(pos, C.c_float) if pos is not None else None,
                              _p(order, C.c_uint32), _p(counts, C.c_uint32),
                              _p(offsets, C.c_uint32))
    return order, counts, offsets


def exclusion_pairs(n: int, bonds: np.ndarray, distance: int) -> np.ndarray:
    lib = library()
    b = np.ascontiguousarray(np.asarray(bonds).reshape(-1, 3), np.uint32)
    total = lib.orc_exclusion_pairs(n, len(b), _p(b, C.c_uint32), distance, None, 0)
    out = np.zeros((max(1, total), 2), np.uint32)
    lib.orc_exclusion_pairs(n, len(b), _p(b, C.c_uint32), distance, _p(out, C.c_uint32), total)
    return out[:total]


def neighbor_lists(w, positions=None, tags=None, types=None, path: int = -1,
                   max_neighbors: int | None = None):
    """(counts[n], list[n][max]) atom-major, overflow flag."""
    lib = library()
    m = w.max_neighbors if max_neighbors is None else max_neighbors
    sys = OracleSystem(w, positions, tags, types)
    counts = np.zeros(w.n, np.uint32)
    lst = np.zeros((w.n, m), np.uint32)
    ovf = np.zeros(1, np.int32)
    if lib.orc_neighbor_build(C.byref(sys.s), w.cutoff_buffer, m, path, _p(counts, C.c_uint32),
                              _p(lst, C.c_uint32), _p(ovf, C.c_int32)):
        raise OracleError(lib.orc_last_error().decode())
    return counts, lst, bool(ovf[0])


def forces(w, positions=None, tags=None, types=None, counts=None, lst=None):
    """FP64 forces [n][3] and observables. counts/lst are atom-major if given;
    otherwise the lists are built as the engine builds them, with the
    workload's skin (r_s = r_c + cutoff_buffer): the force pass keeps pairs
    with r^2 < r_c^2 on the exact-intermediate minimum image, which can admit
    a pair at r_c that a list built with r_s = r_c (rounding its minimum image
    at the scale of the box) would have dropped."""
    lib = library()
    if counts is None:
        counts, lst, ovf = neighbor_lists(w, positions, tags, types)
        if ovf:
            raise OracleError("oracle neighbour list overflow; raise max_neighbors")
    sys = OracleSystem(w, positions, tags, types)
    f = np.zeros((w.n, 3), np.float64)
    obs = Observables()
    if counts is not None:
        c = np.ascontiguousarray(counts, np.uint32)
        l = np.ascontiguousarray(lst, np.uint32)
        rc = lib.orc_forces(C.byref(sys.s), _p(c, C.c_uint32), _p(l, C.c_uint32), l.shape[1],
                            _p(f, C.c_double), C.byref(obs))
    else:
        rc = lib.orc_forces(C.byref(sys.s), None, None, 0, _p(f, C.c_double), C.byref(obs))
    if rc:
        raise OracleError(lib.orc_last_error().decode())
    return f, obs.as_dict()


def forces_f64(w, pos64: np.ndarray):
    lib = library()
    sys = OracleSystem(w)
    p = np.ascontiguousarray(pos64, np.float64)
    f = np.zeros((w.n, 3), np.float64)
    obs = Observables()
    if lib.orc_forces_f64(C.byref(sys.s), _p(p, C.c_double), _p(f, C.c_double), C.byref(obs)):
        raise OracleError(lib.orc_last_error().decode())
    return f, obs.as_dict()


def vv_pre(w, pos, vel, force, images):
    lib = library()
    n = w.n
    args = [np.ascontiguousarray(x, dt) for x, dt in ((pos, np.float32), (vel, np.float32),
                                                      (force, np.float32), (w.masses, np.float32),
                                                      (images, np.int32))]
    lo = np.ascontiguousarray(w.box_bound[0], np.float32)
    hi = np.ascontiguousarray(w.box_bound[1], np.float32)
    po = np.zeros((n, 3)); vo = np.zeros((n, 3)); io = np.zeros((n, 3), np.int32)
    lib.orc_vv_pre(n, _p(args[0], C.c_float), _p(args[1], C.c_float), _p(args[2], C.c_float),
                   _p(args[3], C.c_float), _p(args[4], C.c_int32), _p(lo, C.c_float),
                   _p(hi, C.c_float), w.dt, _p(po, C.c_double), _p(vo, C.c_double),
                   _p(io, C.c_int32))
    return po, vo, io


def vv_post(w, vel, force, masses=None):
    lib = library()
    v = np.ascontiguousarray(vel, np.float32)
    f = np.ascontiguousarray(force, np.float32)
    m = np.ascontiguousarray(w.masses if masses is None else masses, np.float32)
    out = np.zeros((w.n, 3))
    lib.orc_vv_post(w.n, _p(v, C.c_float), _p(f, C.c_float), _p(m, C.c_float), w.dt, _p(out, C.c_double))
    return out


def kinetic_energy(vel, masses) -> float:
    v = np.ascontiguousarray(vel, np.float32)
    m = np.ascontiguousarray(masses, np.float32)
    return library().orc_kinetic_energy(m.size, _p(v, C.c_float), _p(m, C.c_float))


def berendsen_lambda(ke: float, n: int, kb: float, temperature: float, dt: float, tau: float) -> float:
    return library().orc_berendsen_lambda(ke, n, kb, temperature, dt, tau)


def displacement_exceeds(w, pos, images, ref_pos, ref_images) -> bool:
    """Binary32 displacement test of the rebuild tracker over all atoms."""
    lib = library()
    a = [np.ascontiguousarray(x, dt) for x, dt in ((pos, np.float32), (images, np.int32),
                                                   (ref_pos, np.float32), (ref_images, np.int32))]
    lo = np.ascontiguousarray(w.box_bound[0], np.float32)
    hi = np.ascontiguousarray(w.box_bound[1], np.float32)
    return bool(lib.orc_displacement_exceeds(a[0].shape[0], _p(a[0], C.c_float), _p(a[1], C.c_int32),
                                             _p(a[2], C.c_float), _p(a[3], C.c_int32),
                                             _p(lo, C.c_float), _p(hi, C.c_float),
                                             w.cutoff_buffer))


def _run_params(w, steps, threads) -> RunParams:
    return RunParams(dt=w.dt, thermostat=1 if w.berendsen else 0, temperature=w.temperature,
                     tau=w.tau, kb=w.kb, cutoff_buffer=w.cutoff_buffer,
                     rebuild_check_every=w.rebuild_check_every,
                     sort_every_rebuild=w.sort_every_rebuild,
                     sort_every_steps=w.sort_every_steps, max_neighbors=w.max_neighbors,
                     steps=w.steps if steps is None else steps, threads=threads)


def run(w, steps: int | None = None, threads: int = 0):
    """Whole FP64 run; returns final state (oracle slot order) and stats."""
    lib = library()
    sys = OracleSystem(w)
    p = _run_params(w, steps, threads)
    n = w.n
    pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); img = np.zeros((n, 3), np.int32)
    tags = np.zeros(n, np.uint32)
    st = RunStats()
    rc = lib.orc_run(C.byref(sys.s), C.byref(p), _p(pos, C.c_double), _p(vel, C.c_double),
                     _p(img, C.c_int32), _p(tags, C.c_uint32), C.byref(st))
    if rc:
        raise OracleError(lib.orc_last_error().decode())
    return {"positions": pos, "velocities": vel, "images": img, "tags": tags}, st


# ---- the BEADS_REAL64 build: binary64 positions, cell grid and decisions ----

def plan_f64(w) -> PlanF64:
    lib = library()
    p = PlanF64()
    lo = np.ascontiguousarray(w.box_bound[0], np.float32)
    hi = np.ascontiguousarray(w.box_bound[1], np.float32)
    if lib.orc_plan_f64(_p(lo, C.c_float), _p(hi, C.c_float), float(w.cutoff),
                        float(w.cutoff_buffer), w.n, C.byref(p)):
        raise OracleError(lib.orc_last_error().decode())
    return p


def cell_ids_f64(w, pos64: np.ndarray) -> np.ndarray:
    lib = library()
    p = plan_f64(w)
    pos = np.ascontiguousarray(pos64, np.float64)
    lo = np.ascontiguousarray(w.box_bound[0], np.float32)
    out = np.zeros(w.n, np.uint32)
    lib.orc_cell_ids_f64(w.n, _p(pos, C.c_double), _p(lo, C.c_float), C.byref(p),
                         _p(out, C.c_uint32))
    return out


def neighbor_lists_f64(w, pos64: np.ndarray, tags=None, types=None, path: int = -1,
                       max_neighbors: int | None = None):
    """(counts[n], list[n][max]) atom-major, overflow flag; binary64 decisions."""
    lib = library()
    m = w.max_neighbors if max_neighbors is None else max_neighbors
    sys = OracleSystem(w, tags=tags, types=types)
    pos = np.ascontiguousarray(pos64, np.float64)
    counts = np.zeros(w.n, np.uint32)
    lst = np.zeros((w.n, m), np.uint32)
    ovf = np.zeros(1, np.int32)
    if lib.orc_neighbor_build_f64(C.byref(sys.s), _p(pos, C.c_double), float(w.cutoff_buffer), m,
                                  path, _p(counts, C.c_uint32), _p(lst, C.c_uint32),
                                  _p(ovf, C.c_int32)):
        raise OracleError(lib.orc_last_error().decode())
    return counts, lst, bool(ovf[0])


def forces_real64(w, pos64: np.ndarray, tags=None, types=None, counts=None, lst=None):
    """FP64 forces [n][3] and observables at binary64 positions with the
    BEADS_REAL64 build's decisions: the lists (atom-major counts/lst, or built
    here with the workload's skin and binary64 decisions) and r^2 < r_c^2 in
    binary64."""
    lib = library()
    if counts is None:
        counts, lst, ovf = neighbor_lists_f64(w, pos64, tags=tags, types=types)
        if ovf:
            raise OracleError("oracle neighbour list overflow; raise max_neighbors")
    sys = OracleSystem(w, tags=tags, types=types)
    pos = np.ascontiguousarray(pos64, np.float64)
    c = np.ascontiguousarray(counts, np.uint32)
    l = np.ascontiguousarray(lst, np.uint32)
    f = np.zeros((w.n, 3), np.float64)
    obs = Observables()
    if lib.orc_forces_f64_list(C.byref(sys.s), _p(pos, C.c_double), _p(c, C.c_uint32),
                               _p(l, C.c_uint32), l.shape[1], _p(f, C.c_double), C.byref(obs)):
        raise OracleError(lib.orc_last_error().decode())
    return f, obs.as_dict()


def run_f64(w, pos64: np.ndarray, vel64: np.ndarray, steps: int | None = None, threads: int = 0):
    """Whole FP64 run from binary64 positions and velocities with the binary64
    cell grid and decisions; returns final state (oracle slot order) and stats."""
    lib = library()
    sys = OracleSystem(w)
    p = _run_params(w, steps, threads)
    n = w.n
    x0 = np.ascontiguousarray(pos64, np.float64)
    v0 = np.ascontiguousarray(vel64, np.float64)
    pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); img = np.zeros((n, 3), np.int32)
    tags = np.zeros(n, np.uint32)
    st = RunStats()
    rc = lib.orc_run_f64(C.byref(sys.s), C.byref(p), _p(x0, C.c_double), _p(v0, C.c_double),
                         _p(pos, C.c_double), _p(vel, C.c_double), _p(img, C.c_int32),
                         _p(tags, C.c_uint32), C.byref(st))
    if rc:
        raise OracleError(lib.orc_last_error().decode())
    return {"positions": pos, "velocities": vel, "images": img, "tags": tags}, st


def run_tangent_f64(w, pos64: np.ndarray, vel64: np.ndarray, dpos: np.ndarray, dvel: np.ndarray,
                    steps: int | None = None, threads: int = 0):
    """run_f64 with the tangent of the initial perturbation (dpos, dvel): returns
    the final state and {"dpos", "dvel"} in the same (oracle) slot order."""
    lib = library()
    sys = OracleSystem(w)
    p = _run_params(w, steps, threads)
    n = w.n
    arrs = [np.ascontiguousarray(a, np.float64) for a in (pos64, vel64, dpos, dvel)]
    pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); img = np.zeros((n, 3), np.int32)
    tags = np.zeros(n, np.uint32)
    dpo = np.zeros((n, 3)); dvo = np.zeros((n, 3))
    st = RunStats()
    rc = lib.orc_run_tangent_f64(C.byref(sys.s), C.byref(p), *[_p(a, C.c_double) for a in arrs],
                                 _p(pos, C.c_double), _p(vel, C.c_double), _p(img, C.c_int32),
                                 _p(tags, C.c_uint32), _p(dpo, C.c_double), _p(dvo, C.c_double),
                                 C.byref(st))
    if rc:
        raise OracleError(lib.orc_last_error().decode())
    return ({"positions": pos, "velocities": vel, "images": img, "tags": tags},
            {"dpos": dpo, "dvel": dvo}, st)
