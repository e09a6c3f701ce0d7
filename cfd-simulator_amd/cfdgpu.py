"""ctypes binding of libcfdgpu.so - the MI355X SIMPLE solver's C ABI (include/cfdgpu.h).

The object model mirrors the reference's C++ API (sidhanth-sridhar/CFD-SIMULATOR):
``Mesh`` <- mesh::generateCGrid / generateBox (include/cfd/mesh/CGrid.hpp, BoxGrid.hpp),
``SimpleSolver`` <- solver::SimpleSolver (include/cfd/solver/SimpleSolver.hpp:159-197),
``surface`` <- post::extractSurface + post::integrateForces (SurfaceData.hpp:159,
Forces.hpp:127-133). Errors raise ``CfdError`` carrying the reference's ErrorCode value.

There is no CPU fallback: if the library is missing, or no GPU is visible when a
solver is created, this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CFDGPU_LIBRARY", os.path.join(_HERE, "libcfdgpu.so"))

INTERIOR, WALL, FARFIELD, OUTLET, WAKECUT = 0, 1, 2, 3, 4
K_INTERNAL, K_INLET, K_OUTLET, K_FARFIELD, K_WALL = 0, 1, 2, 3, 4
MODEL_NONE, MODEL_LAMINAR, MODEL_SST = 0, 1, 2
FP64, MIXED = 0, 1
GS_LEX, GS_REDBLACK = 0, 1
PSOLVE_JACOBI, PSOLVE_LINE = 0, 1
ERROR_NAMES = {0: "Ok", 1: "InvalidArgument", 2: "OutOfRange", 3: "NotFound", 4: "IoFailure",
               5: "NotImplemented", 6: "InitializationFailure", 7: "NumericalFailure", 8: "Internal"}

EXPORTS = [
    "cfdgpu_version", "cfdgpu_last_error", "cfdgpu_device_count", "cfdgpu_cgrid_defaults",
    "cfdgpu_mesh_cgrid", "cfdgpu_mesh_box", "cfdgpu_mesh_structured", "cfdgpu_mesh_destroy",
    "cfdgpu_mesh_dims", "cfdgpu_mesh_nodes", "cfdgpu_mesh_cells", "cfdgpu_mesh_faces",
    "cfdgpu_mesh_quality", "cfdgpu_face_conditions", "cfdgpu_uniform_field", "cfdgpu_default_settings",
    "cfdgpu_default_sst", "cfdgpu_create", "cfdgpu_destroy", "cfdgpu_set_turbulence", "cfdgpu_initialise",
    "cfdgpu_iterate", "cfdgpu_iterate_timed", "cfdgpu_solve", "cfdgpu_batch_create", "cfdgpu_batch_destroy",
    "cfdgpu_batch_iterate", "cfdgpu_batch_iterate_timed", "cfdgpu_batch_solve", "cfdgpu_polar_sweep", "cfdgpu_polar_csv", "cfdgpu_get_settings", "cfdgpu_set_settings", "cfdgpu_get_field",
    "cfdgpu_set_field", "cfdgpu_get_massflux", "cfdgpu_set_massflux", "cfdgpu_get_turbulence",
    "cfdgpu_set_turbulence_fields", "cfdgpu_turbulence_info", "cfdgpu_divergence", "cfdgpu_divergence_noalloc",
    "cfdgpu_set_launch_tuning", "cfdgpu_get_launch_tuning", "cfdgpu_surface",
    "cfdgpu_sweep_angles", "cfdgpu_debug_array", "cfdgpu_server_defaults", "cfdgpu_server_create",
    "cfdgpu_server_destroy", "cfdgpu_server_submit", "cfdgpu_server_step", "cfdgpu_server_run",
    "cfdgpu_server_collect", "cfdgpu_server_fields", "cfdgpu_solver_surface",
]


class CfdError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class CGridOptions(C.Structure):
    _fields_ = [("upstream_chords", C.c_double), ("downstream_chords", C.c_double),
                ("vertical_chords", C.c_double), ("surface_points", C.c_int), ("wake_points", C.c_int),
                ("normal_points", C.c_int), ("first_layer_height", C.c_double),
                ("normal_blend_length", C.c_double), ("normal_smoothing_passes", C.c_int)]


class BoxOptions(C.Structure):
    _fields_ = [("origin_x", C.c_double), ("origin_y", C.c_double), ("length", C.c_double),
                ("height", C.c_double), ("cells_x", C.c_int), ("cells_y", C.c_int), ("grading", C.c_int),
                ("first_cell_height", C.c_double), ("left", C.c_int), ("right", C.c_int),
                ("lower", C.c_int), ("upper", C.c_int)]


class Settings(C.Structure):
    _fields_ = [("velocity_relaxation", C.c_double), ("pressure_relaxation", C.c_double),
                ("momentum_sweeps", C.c_int), ("pressure_iterations", C.c_int),
                ("pressure_tolerance", C.c_double), ("scheme", C.c_int),
                ("turbulence_relaxation", C.c_double), ("intensity", C.c_double),
                ("viscosity_ratio", C.c_double), ("length_scale", C.c_double),
                ("non_orthogonal_correctors", C.c_int), ("non_orthogonal_blend", C.c_double),
                ("precision", C.c_int), ("gs_ordering", C.c_int), ("pressure_solver", C.c_int),
                ("refinement_passes", C.c_int)]


class SstConstants(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("sigma_k1", "sigma_w1", "beta1", "sigma_k2", "sigma_w2", "beta2",
                                          "beta_star", "kappa", "a1", "production_limit")]


# CFDGPU_STAGE_* in order
STAGES = ("gradients", "momentum_assembly", "momentum_sweeps", "rhie_chow", "pressure_assembly",
          "pressure_solve", "correction", "sst_blending", "sst_assembly", "sst_sweeps",
          "sst_wall_viscosity", "monitor", "sweep_copy",
          "cg_iterations")  # the last is inside pressure_solve, not added to it


class Monitor(C.Structure):
    _fields_ = [("continuity", C.c_double), ("momentum_x", C.c_double), ("momentum_y", C.c_double),
                ("mass_imbalance", C.c_double), ("max_divergence", C.c_double),
                ("pressure_iterations", C.c_int), ("turb_first", C.c_double), ("turb_second", C.c_double),
                ("min_k", C.c_double), ("max_k", C.c_double), ("min_w", C.c_double), ("max_w", C.c_double),
                ("min_mut", C.c_double), ("max_mut", C.c_double), ("max_eddy_ratio", C.c_double)]

    def worst(self) -> float:
        return max(abs(self.continuity), abs(self.momentum_x), abs(self.momentum_y))


class StopRule(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int), ("check_every", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("last", Monitor), ("seconds", C.c_double)]


class PolarPoint(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("alpha_deg", "cl", "cd", "cd_pressure", "cd_friction", "cm", "l_over_d",
                                          "upper_separation", "lower_separation")] + \
               [("status", C.c_int), ("iterations", C.c_int), ("continuity", C.c_double), ("retried_cold", C.c_int)]


class PolarHeader(C.Structure):
    _fields_ = [("section", C.c_char_p), ("mesh", C.c_char_p), ("reynolds", C.c_double), ("speed", C.c_double),
                ("chord", C.c_double), ("moment_reference", C.c_double), ("continued", C.c_int)]


class Forces(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cl", "cd", "cm", "cd_pressure", "cd_friction", "lift", "drag",
                                          "pitching_moment", "fx_pressure", "fy_pressure", "fx_friction",
                                          "fy_friction", "upper_separation", "lower_separation",
                                          "stagnation_xc")] + \
               [("n_upper", C.c_int), ("n_lower", C.c_int), ("stations", C.c_int),
                ("upper_separation_face", C.c_int), ("lower_separation_face", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_slots", C.c_int), ("chunk_iterations", C.c_int),
                ("rule", StopRule), ("settings", Settings), ("model", C.c_int), ("warm_start", C.c_int),
                ("keep_fields", C.c_int), ("settle_window", C.c_int), ("settle_cl", C.c_double),
                ("settle_cd", C.c_double)]


class ServeRequest(C.Structure):
    _fields_ = [("m", C.c_int), ("p", C.c_int), ("t", C.c_int), ("chord", C.c_double), ("mesh", CGridOptions)] + \
               [(n, C.c_double) for n in ("alpha_deg", "reynolds", "speed", "density", "reference_pressure",
                                          "arrival_s")] + \
               [("settings", C.POINTER(Settings))]


class ServeResult(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ticket", "status", "iterations", "retried_cold", "warm_from", "cells")] + \
               [(n, C.c_double) for n in ("cl", "cd", "cd_pressure", "cd_friction", "cm", "l_over_d",
                                          "upper_separation", "lower_separation", "continuity", "arrival_s",
                                          "admitted_s", "finished_s", "device_bytes")]


class ServerStats(C.Structure):
    _fields_ = [("clock_s", C.c_double)] + \
               [(n, C.c_int) for n in ("queued", "active", "admitted", "finished", "converged", "peak_active",
                                       "rounds", "meshes_cached")] + \
               [(n, C.c_double) for n in ("resident_bytes", "peak_resident_bytes", "member_iterations")]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _opt(arr):
    """Optional double* argument."""
    return None if arr is None else arr.ctypes.data_as(C.POINTER(C.c_double))


def load(path: str = LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"the HIP library is not built: {path} is missing "
                           f"(run __graft_entry__.build() or make -C cfd-simulator_amd)")
    lib = C.CDLL(path)
    vp, pp = C.c_void_p, C.POINTER(C.c_void_p)
    D = C.POINTER(C.c_double)
    sig = {
        "cfdgpu_version": (C.c_char_p, []),
        "cfdgpu_last_error": (C.c_char_p, []),
        "cfdgpu_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "cfdgpu_cgrid_defaults": (C.c_int, [C.c_int, C.POINTER(CGridOptions)]),
        "cfdgpu_mesh_cgrid": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(CGridOptions), pp]),
        "cfdgpu_mesh_box": (C.c_int, [C.POINTER(BoxOptions), pp]),
        "cfdgpu_mesh_structured": (C.c_int, [C.c_int, C.c_int, _dp, _ip, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                             pp]),
        "cfdgpu_mesh_destroy": (None, [vp]),
        "cfdgpu_mesh_dims": (C.c_int, [vp] + [C.POINTER(C.c_int)] * 4),
        "cfdgpu_mesh_nodes": (C.c_int, [vp, _dp]),
        "cfdgpu_mesh_cells": (C.c_int, [vp, _dp, _dp]),
        "cfdgpu_mesh_faces": (C.c_int, [vp, _dp, _dp, _dp, _ip, _ip, _ip, _ip]),
        "cfdgpu_mesh_quality": (C.c_int, [vp, _dp]),
        "cfdgpu_face_conditions": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, _ip, _dp, _dp]),
        "cfdgpu_uniform_field": (C.c_int, [C.c_int] + [C.c_double] * 6 + [_dp] * 5),
        "cfdgpu_default_settings": (None, [C.POINTER(Settings)]),
        "cfdgpu_default_sst": (None, [C.POINTER(SstConstants)]),
        "cfdgpu_create": (C.c_int, [pp, vp, _ip, _dp, _dp, C.POINTER(Settings)]),
        "cfdgpu_destroy": (None, [vp]),
        "cfdgpu_set_turbulence": (C.c_int, [vp, C.c_int, C.c_void_p]),
        "cfdgpu_initialise": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "cfdgpu_iterate": (C.c_int, [vp, C.c_int, C.POINTER(Monitor)]),
        "cfdgpu_iterate_timed": (C.c_int, [vp, C.c_int, C.POINTER(Monitor), D, C.POINTER(C.c_int)]),
        "cfdgpu_solve": (C.c_int, [vp, C.POINTER(StopRule), C.POINTER(SolveReport)]),
        "cfdgpu_batch_create": (C.c_int, [pp, pp, C.c_int]),
        "cfdgpu_batch_destroy": (None, [vp]),
        "cfdgpu_batch_iterate": (C.c_int, [vp, C.c_int, C.POINTER(Monitor)]),
        "cfdgpu_batch_iterate_timed": (C.c_int, [vp, C.c_int, C.POINTER(Monitor), D, C.POINTER(C.c_int)]),
        "cfdgpu_batch_solve": (C.c_int, [vp, C.POINTER(StopRule), C.POINTER(SolveReport)]),
        "cfdgpu_get_settings": (C.c_int, [vp, C.POINTER(Settings)]),
        "cfdgpu_set_settings": (C.c_int, [vp, C.POINTER(Settings)]),
        "cfdgpu_get_field": (C.c_int, [vp, D, D, D, D]),
        "cfdgpu_set_field": (C.c_int, [vp, D, D, D]),
        "cfdgpu_get_massflux": (C.c_int, [vp, _dp]),
        "cfdgpu_set_massflux": (C.c_int, [vp, _dp]),
        "cfdgpu_get_turbulence": (C.c_int, [vp, D, D, D, D]),
        "cfdgpu_set_turbulence_fields": (C.c_int, [vp, D, D, D]),
        "cfdgpu_turbulence_info": (C.c_int, [vp, D, D, D]),
        "cfdgpu_divergence": (C.c_int, [vp, _dp]),
        "cfdgpu_divergence_noalloc": (C.c_int, [vp, _dp]),
        "cfdgpu_set_launch_tuning": (C.c_int, [C.c_int, C.c_int]),
        "cfdgpu_get_launch_tuning": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "cfdgpu_surface": (C.c_int, [vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp, C.c_int, C.POINTER(C.c_int), C.POINTER(Forces)]),
        "cfdgpu_sweep_angles": (C.c_int, [C.c_double, C.c_double, C.c_double, _dp, C.c_int]),
        "cfdgpu_polar_sweep": (C.c_int, [vp, C.POINTER(Settings), C.c_int, C.c_void_p] + [C.c_double] * 5 +
                               [_dp, C.c_int, C.c_int, C.POINTER(StopRule), C.POINTER(PolarPoint)]),
        "cfdgpu_polar_csv": (C.c_int, [C.POINTER(PolarPoint), C.c_int, C.POINTER(PolarHeader), C.c_char_p,
                                       C.c_int]),
        "cfdgpu_debug_array": (C.c_int, [vp, C.c_char_p, _dp, C.c_int]),
        "cfdgpu_server_defaults": (None, [C.POINTER(ServerOptions)]),
        "cfdgpu_server_create": (C.c_int, [C.POINTER(ServerOptions), pp]),
        "cfdgpu_server_destroy": (None, [vp]),
        "cfdgpu_server_submit": (C.c_int, [vp, C.POINTER(ServeRequest), C.POINTER(C.c_int)]),
        "cfdgpu_server_step": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "cfdgpu_server_run": (C.c_int, [vp, C.c_double, C.POINTER(ServerStats)]),
        "cfdgpu_server_collect": (C.c_int, [vp, C.POINTER(ServeResult), C.c_int, C.POINTER(C.c_int)]),
        "cfdgpu_server_fields": (C.c_int, [vp, C.c_int, D, D, D, D]),
        "cfdgpu_solver_surface": (C.c_int, [vp] + [C.c_double] * 5 + [_dp, C.c_int, C.POINTER(C.c_int),
                                                                      C.POINTER(Forces)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def _check(rc: int):
    if rc != 0:
        raise CfdError(rc, lib().cfdgpu_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    lib().cfdgpu_device_count(C.byref(n))
    return n.value


def set_launch_tuning(cells_per_thread: int, column_rows: int) -> None:
    """Launch shapes of the line-preconditioned CG (process-wide; results agree to rounding)."""
    _check(lib().cfdgpu_set_launch_tuning(int(cells_per_thread), int(column_rows)))


def launch_tuning() -> tuple[int, int]:
    cells, rows = C.c_int(), C.c_int()
    _check(lib().cfdgpu_get_launch_tuning(C.byref(cells), C.byref(rows)))
    return cells.value, rows.value


def default_settings(**overrides) -> Settings:
    s = Settings()
    lib().cfdgpu_default_settings(C.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def default_sst() -> SstConstants:
    c = SstConstants()
    lib().cfdgpu_default_sst(C.byref(c))
    return c


def cgrid_options(resolution: str = "coarse", **overrides) -> CGridOptions:
    o = CGridOptions()
    _check(lib().cfdgpu_cgrid_defaults({"coarse": 0, "medium": 1, "fine": 2}[resolution], C.byref(o)))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


class Mesh:
    """A structured mesh in the reference's numbering (include/cfd/mesh/Mesh.hpp)."""

    def __init__(self, handle):
        self.h = handle
        ni, nj, nc, nf = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().cfdgpu_mesh_dims(handle, C.byref(ni), C.byref(nj), C.byref(nc), C.byref(nf)))
        self.nodes_i, self.nodes_j, self.cells, self.faces = ni.value, nj.value, nc.value, nf.value
        self.ni, self.nj = self.nodes_i - 1, self.nodes_j - 1
        self.centroid = np.zeros((self.cells, 2))
        self.area = np.zeros(self.cells)
        _check(lib().cfdgpu_mesh_cells(handle, self.centroid.reshape(-1), self.area))
        self.fcentre = np.zeros((self.faces, 2))
        self.fnormal = np.zeros((self.faces, 2))
        self.farea = np.zeros(self.faces)
        self.owner = np.zeros(self.faces, dtype=np.int32)
        self.neighbour = np.zeros(self.faces, dtype=np.int32)
        self.boundary = np.zeros(self.faces, dtype=np.int32)
        self.partner = np.zeros(self.faces, dtype=np.int32)
        _check(lib().cfdgpu_mesh_faces(handle, self.fcentre.reshape(-1), self.fnormal.reshape(-1), self.farea,
                                       self.owner, self.neighbour, self.boundary, self.partner))
        self.node_xy = np.zeros((self.nodes_i * self.nodes_j, 2))
        _check(lib().cfdgpu_mesh_nodes(handle, self.node_xy.reshape(-1)))
        q = np.zeros(7)
        _check(lib().cfdgpu_mesh_quality(handle, q))
        self.min_area, self.max_area, self.max_aspect, self.max_nonortho_deg = q[:4]
        self.inverted, self.min_wall_spacing, self.total_area = int(q[4]), q[5], q[6]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cfdgpu_mesh_destroy(self.h)
            self.h = None

    def cell_index(self, i, j):
        return j * self.ni + i


def cgrid(m: int, p: int, t: int, chord: float = 1.0, options: CGridOptions | None = None, **kw) -> Mesh:
    o = options if options is not None else cgrid_options(kw.pop("resolution", "coarse"), **kw)
    h = C.c_void_p()
    _check(lib().cfdgpu_mesh_cgrid(m, p, t, chord, C.byref(o), C.byref(h)))
    return Mesh(h)


def box(**kw) -> Mesh:
    o = BoxOptions(origin_x=0.0, origin_y=0.0, length=1.0, height=1.0, cells_x=40, cells_y=20, grading=0,
                   first_cell_height=0.01, left=FARFIELD, right=OUTLET, lower=WALL, upper=WALL)
    for k, v in kw.items():
        setattr(o, k, v)
    h = C.c_void_p()
    _check(lib().cfdgpu_mesh_box(C.byref(o), C.byref(h)))
    return Mesh(h)


def face_conditions(mesh: Mesh, speed=50.0, alpha_deg=0.0, reference_pressure=0.0):
    kind = np.zeros(mesh.faces, dtype=np.int32)
    vel = np.zeros((mesh.faces, 2))
    pres = np.zeros(mesh.faces)
    _check(lib().cfdgpu_face_conditions(mesh.h, speed, alpha_deg, reference_pressure, kind, vel.reshape(-1), pres))
    return kind, vel, pres


def uniform_field(cells: int, speed=50.0, alpha_deg=0.0, density=1.225, reynolds=1e6, chord=1.0,
                  reference_pressure=0.0):
    arrs = [np.zeros(cells) for _ in range(5)]
    _check(lib().cfdgpu_uniform_field(cells, speed, alpha_deg, density, reynolds, chord, reference_pressure, *arrs))
    return tuple(arrs)


class SimpleSolver:
    """solver::SimpleSolver on the GPU (include/cfd/solver/SimpleSolver.hpp:159-197)."""

    def __init__(self, mesh: Mesh, kind, velocity, pressure, settings: Settings | None = None):
        self.mesh = mesh
        self._settings = settings or default_settings()
        h = C.c_void_p()
        _check(lib().cfdgpu_create(C.byref(h), mesh.h, np.ascontiguousarray(kind, dtype=np.int32),
                                   np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1),
                                   np.ascontiguousarray(pressure, dtype=np.float64), C.byref(self._settings)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cfdgpu_destroy(self.h)
            self.h = None

    def set_turbulence_model(self, model: int, constants: SstConstants | None = None):
        ptr = C.cast(C.pointer(constants), C.c_void_p) if constants is not None else None
        _check(lib().cfdgpu_set_turbulence(self.h, model, ptr))

    def initialise(self, u, v, p, rho, mu):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().cfdgpu_initialise(self.h, f(u), f(v), f(p), f(rho), f(mu)))

    def iterate(self, count: int = 1) -> Monitor:
        m = Monitor()
        _check(lib().cfdgpu_iterate(self.h, count, C.byref(m)))
        return m

    def iterate_timed(self, count: int = 1):
        """iterate() with HIP events around every stage: (monitor, {stage: ms}, {stage: calls})."""
        m = Monitor()
        ms = (C.c_double * len(STAGES))()
        calls = (C.c_int * len(STAGES))()
        _check(lib().cfdgpu_iterate_timed(self.h, count, C.byref(m), ms, calls))
        return m, dict(zip(STAGES, list(ms))), dict(zip(STAGES, list(calls)))

    def solve(self, tolerance: float, max_iterations: int, check_every: int = 25) -> SolveReport:
        rule = StopRule(tolerance, max_iterations, check_every)
        rep = SolveReport()
        _check(lib().cfdgpu_solve(self.h, C.byref(rule), C.byref(rep)))
        return rep

    @property
    def settings(self) -> Settings:
        s = Settings()
        _check(lib().cfdgpu_get_settings(self.h, C.byref(s)))
        return s

    def set_settings(self, s: Settings):
        _check(lib().cfdgpu_set_settings(self.h, C.byref(s)))

    def field(self):
        n = self.mesh.cells
        u, v, p, mu = (np.zeros(n) for _ in range(4))
        _check(lib().cfdgpu_get_field(self.h, _opt(u), _opt(v), _opt(p), _opt(mu)))
        return u, v, p, mu

    def set_field(self, u=None, v=None, p=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        u, v, p = f(u), f(v), f(p)
        _check(lib().cfdgpu_set_field(self.h, _opt(u), _opt(v), _opt(p)))

    def mass_flux(self):
        out = np.zeros(self.mesh.faces)
        _check(lib().cfdgpu_get_massflux(self.h, out))
        return out

    def set_mass_flux(self, flux):
        _check(lib().cfdgpu_set_massflux(self.h, np.ascontiguousarray(flux, dtype=np.float64)))

    def turbulence(self):
        n = self.mesh.cells
        k, w, mut, d = (np.zeros(n) for _ in range(4))
        _check(lib().cfdgpu_get_turbulence(self.h, _opt(k), _opt(w), _opt(mut), _opt(d)))
        return dict(k=k, omega=w, mut=mut, wall_distance=d)

    def set_turbulence_fields(self, k=None, omega=None, mut=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        k, omega, mut = f(k), f(omega), f(mut)
        _check(lib().cfdgpu_set_turbulence_fields(self.h, _opt(k), _opt(omega), _opt(mut)))

    def turbulence_info(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().cfdgpu_turbulence_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def divergence(self):
        out = np.zeros(self.mesh.cells)
        _check(lib().cfdgpu_divergence(self.h, out))
        return out

    def surface(self, speed=50.0, alpha_deg=0.0, density=1.225, reference_pressure=0.0, chord=1.0):
        """The surface distribution and forces of the current field, on the
        device (cfdgpu_solver_surface). Returns (Forces, rows[n,12])."""
        cap = 12 * self.mesh.ni
        rows = np.zeros(cap)
        f = Forces()
        n = C.c_int()
        _check(lib().cfdgpu_solver_surface(self.h, speed, alpha_deg, density, reference_pressure, chord, rows, cap,
                                           C.byref(n), C.byref(f)))
        return f, rows[:12 * n.value].reshape(n.value, 12)

    def debug_array(self, name: str):
        cap = 2 * self.mesh.cells
        out = np.zeros(cap)
        n = lib().cfdgpu_debug_array(self.h, name.encode(), out, cap)
        if n < 0:
            raise CfdError(-n, name)
        return out[:n].copy()


class Batch:
    """Several SimpleSolvers advanced together (cfdgpu_batch_*): one launch per
    stage for all members. The solvers must stay alive while the batch is used."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        _check(lib().cfdgpu_batch_create(C.byref(h), arr, len(self.solvers)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cfdgpu_batch_destroy(self.h)
            self.h = None

    def iterate(self, count: int = 1):
        mons = (Monitor * len(self.solvers))()
        _check(lib().cfdgpu_batch_iterate(self.h, count, mons))
        return list(mons)

    def iterate_timed(self, count: int = 1):
        """(monitors, {stage: ms summed over the iterations}, {stage: calls})."""
        mons = (Monitor * len(self.solvers))()
        ms = (C.c_double * len(STAGES))()
        calls = (C.c_int * len(STAGES))()
        _check(lib().cfdgpu_batch_iterate_timed(self.h, count, mons, ms, calls))
        return list(mons), dict(zip(STAGES, list(ms))), dict(zip(STAGES, list(calls)))

    def solve(self, tolerance: float, max_iterations: int, check_every: int = 25):
        rule = StopRule(tolerance, max_iterations, check_every)
        reps = (SolveReport * len(self.solvers))()
        _check(lib().cfdgpu_batch_solve(self.h, C.byref(rule), reps))
        return list(reps)


STATUS_NAMES = {0: "converged", 1: "iteration-limit", 2: "diverged", 3: "rejected"}


def server_options(**overrides) -> ServerOptions:
    o = ServerOptions()
    lib().cfdgpu_server_defaults(C.byref(o))
    for k, v in overrides.items():
        if k in ("tolerance", "max_iterations"):
            setattr(o.rule, k, v)
        elif k == "settings":
            o.settings = v
        else:
            setattr(o, k, v)
    return o


def serve_request(m: int, p: int, t: int, alpha_deg: float, reynolds: float, mesh: CGridOptions,
                  arrival_s: float = 0.0, chord=1.0, speed=1.0, density=1.0, reference_pressure=0.0,
                  settings: Settings | None = None) -> ServeRequest:
    r = ServeRequest(m, p, t, chord, mesh, alpha_deg, reynolds, speed, density, reference_pressure, arrival_s)
    if settings is not None:
        r._settings = settings  # kept alive with the request
        r.settings = C.pointer(settings)
    return r


class Server:
    """The request loop of polar-point solves with continuous batching
    (cfdgpu_server_*): requests are admitted as they arrive, by memory budget
    and slot count, and advanced together on the device in chunks."""

    def __init__(self, options: ServerOptions | None = None, **overrides):
        self.options = options if options is not None else server_options(**overrides)
        h = C.c_void_p()
        _check(lib().cfdgpu_server_create(C.byref(self.options), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cfdgpu_server_destroy(self.h)
            self.h = None

    def submit(self, request: ServeRequest) -> int:
        t = C.c_int()
        _check(lib().cfdgpu_server_submit(self.h, C.byref(request), C.byref(t)))
        return t.value

    def step(self) -> ServerStats:
        st = ServerStats()
        _check(lib().cfdgpu_server_step(self.h, C.byref(st)))
        return st

    def run(self, wall_limit_s: float = 0.0) -> ServerStats:
        st = ServerStats()
        _check(lib().cfdgpu_server_run(self.h, wall_limit_s, C.byref(st)))
        return st

    def collect(self, capacity: int = 4096):
        out = (ServeResult * capacity)()
        n = C.c_int()
        _check(lib().cfdgpu_server_collect(self.h, out, capacity, C.byref(n)))
        return list(out[:n.value])

    def fields(self, ticket: int, cells: int):
        arrs = [np.zeros(cells) for _ in range(4)]
        _check(lib().cfdgpu_server_fields(self.h, ticket, *[_opt(a) for a in arrs]))
        return tuple(arrs)


def surface(mesh: Mesh, u, v, p, mu, speed=50.0, alpha_deg=0.0, density=1.225, reference_pressure=0.0, chord=1.0):
    """extractSurface + integrateForces in FP64. Returns (Forces, rows[n,12])."""
    cap = 12 * mesh.ni
    rows = np.zeros(cap)
    f = Forces()
    n = C.c_int()
    g = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    _check(lib().cfdgpu_surface(mesh.h, g(u), g(v), g(p), g(mu), speed, alpha_deg, density, reference_pressure,
                                chord, rows, cap, C.byref(n), C.byref(f)))
    return f, rows[:12 * n.value].reshape(n.value, 12)


def polar_sweep(mesh: Mesh, alphas, settings: Settings | None = None, model: int = 0, speed=50.0, density=1.225,
                reynolds=1e6, chord=1.0, reference_pressure=0.0, continued=True, tolerance=1e-6,
                max_iterations=10000, check_every=25):
    """The reference's continued polar sweep on the device (cfdgpu_polar_sweep)."""
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    out = (PolarPoint * len(a))()
    rule = StopRule(tolerance, max_iterations, check_every)
    _check(lib().cfdgpu_polar_sweep(mesh.h, C.byref(settings or default_settings()), model, None, speed, density,
                                    reynolds, chord, reference_pressure, a, len(a), 1 if continued else 0,
                                    C.byref(rule), out))
    return list(out)


def polar_csv(points, section="unknown", mesh="unknown", reynolds=1e6, speed=50.0, chord=1.0,
              moment_reference=0.25, continued=True) -> str:
    """post::toCsv (Polar.cpp:126-163)."""
    arr = (PolarPoint * len(points))(*points)
    h = PolarHeader(section.encode(), mesh.encode(), reynolds, speed, chord, moment_reference, 1 if continued else 0)
    n = lib().cfdgpu_polar_csv(arr, len(points), C.byref(h), None, 0)
    if n < 0:
        raise CfdError(-n, lib().cfdgpu_last_error().decode())
    buf = C.create_string_buffer(n + 1)
    lib().cfdgpu_polar_csv(arr, len(points), C.byref(h), buf, n + 1)
    return buf.value.decode()


def sweep_angles(start: float, end: float, step: float):
    out = np.zeros(401)
    n = lib().cfdgpu_sweep_angles(start, end, step, out, 401)
    if n < 0:
        raise CfdError(-n, lib().cfdgpu_last_error().decode())
    return out[:n].copy()
