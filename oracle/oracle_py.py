"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Only tests/,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import it.
The oracle restates the reference solver (sidhanth-sridhar/CFD-SIMULATOR,
src/solver/SimpleSolver.cpp et al.); see oracle.cpp for per-routine citations.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

INTERIOR, WALL, FARFIELD, OUTLET, WAKECUT = 0, 1, 2, 3, 4
K_INTERNAL, K_INLET, K_OUTLET, K_FARFIELD, K_WALL = 0, 1, 2, 3, 4

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class CGridOptions(C.Structure):
    _fields_ = [("upstream", C.c_double), ("downstream", C.c_double), ("vertical", C.c_double),
                ("surface_points", C.c_int), ("wake_points", C.c_int), ("normal_points", C.c_int),
                ("first_layer", C.c_double), ("normal_blend_length", C.c_double),
                ("normal_smoothing_passes", C.c_int)]


class BoxOptions(C.Structure):
    _fields_ = [("origin_x", C.c_double), ("origin_y", C.c_double), ("length", C.c_double),
                ("height", C.c_double), ("cells_x", C.c_int), ("cells_y", C.c_int),
                ("grading", C.c_int), ("first_cell_height", C.c_double), ("left", C.c_int),
                ("right", C.c_int), ("lower", C.c_int), ("upper", C.c_int)]


class Settings(C.Structure):
    _fields_ = [("velocity_relaxation", C.c_double), ("pressure_relaxation", C.c_double),
                ("momentum_sweeps", C.c_int), ("pressure_iterations", C.c_int),
                ("pressure_tolerance", C.c_double), ("scheme", C.c_int),
                ("turbulence_relaxation", C.c_double), ("intensity", C.c_double),
                ("viscosity_ratio", C.c_double), ("length_scale", C.c_double),
                ("non_orthogonal_correctors", C.c_int), ("non_orthogonal_blend", C.c_double)]


class SstConstants(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("sigma_k1", "sigma_w1", "beta1", "sigma_k2", "sigma_w2",
                                          "beta2", "beta_star", "kappa", "a1", "production_limit")]


class Monitor(C.Structure):
    _fields_ = [("continuity", C.c_double), ("momentum_x", C.c_double), ("momentum_y", C.c_double),
                ("mass_imbalance", C.c_double), ("max_divergence", C.c_double),
                ("pressure_iterations", C.c_int), ("turb_first", C.c_double),
                ("turb_second", C.c_double), ("min_k", C.c_double), ("max_k", C.c_double),
                ("min_w", C.c_double), ("max_w", C.c_double), ("min_mut", C.c_double),
                ("max_mut", C.c_double), ("max_eddy_ratio", C.c_double)]

    def worst(self) -> float:
        return max(abs(self.continuity), abs(self.momentum_x), abs(self.momentum_y))


class Forces(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cl", "cd", "cm", "cd_pressure", "cd_friction", "lift", "drag",
                                          "pitching_moment", "fx_pressure", "fy_pressure",
                                          "fx_friction", "fy_friction", "upper_separation",
                                          "lower_separation", "stagnation_xc")] + \
               [("n_upper", C.c_int), ("n_lower", C.c_int), ("stations", C.c_int)]


def _load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_naca_contour": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.c_int]),
        "orc_thickness": (C.c_double, [C.c_double, C.c_double, C.c_int]),
        "orc_camber": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "orc_camber_slope": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "orc_geometric_distribution": (C.c_int, [C.c_double, C.c_double, C.c_int, _dp]),
        "orc_cgrid_defaults": (C.c_int, [C.c_int, C.POINTER(CGridOptions)]),
        "orc_cgrid": (vp, [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(CGridOptions)]),
        "orc_box": (vp, [C.POINTER(BoxOptions)]),
        "orc_mesh_free": (None, [vp]),
        "orc_mesh_dims": (None, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
        "orc_mesh_quality": (None, [vp, _dp]),
        "orc_mesh_nodes": (None, [vp, _dp]),
        "orc_mesh_cells": (None, [vp, _dp, _dp, _ip]),
        "orc_mesh_faces": (None, [vp, _dp, _dp, _dp, _ip, _ip, _ip, _ip]),
        "orc_face_conditions": (None, [vp, C.c_double, C.c_double, C.c_double, _ip, _dp, _dp]),
        "orc_set_sum_block": (None, [C.c_int]),
        "orc_set_tanh_variant": (None, [C.c_int]),
        "orc_default_settings": (None, [C.POINTER(Settings)]),
        "orc_default_sst": (None, [C.POINTER(SstConstants)]),
        "orc_solver_create": (vp, [vp, _ip, _dp, _dp, C.POINTER(Settings)]),
        "orc_solver_free": (None, [vp]),
        "orc_solver_set_model": (C.c_int, [vp, C.c_int, C.c_void_p]),
        "orc_solver_initialise": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "orc_solver_set_settings": (None, [vp, C.POINTER(Settings)]),
        "orc_solver_iterate": (C.c_int, [vp, C.c_int, C.POINTER(Monitor)]),
        "orc_solver_field": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_solver_set_field": (None, [vp, _dp, _dp, _dp]),
        "orc_solver_flux": (None, [vp, _dp]),
        "orc_solver_set_flux": (None, [vp, _dp]),
        "orc_solver_turbulence": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_solver_set_turbulence": (None, [vp, _dp, _dp, _dp]),
        "orc_solver_divergence": (None, [vp, _dp]),
        "orc_solver_freestream_k": (C.c_double, [vp]),
        "orc_solver_freestream_omega": (C.c_double, [vp]),
        "orc_solver_max_yplus": (C.c_double, [vp]),
        "orc_solver_array": (C.c_int, [vp, C.c_char_p, _dp, C.c_int]),
        "orc_stage_interpolate": (None, [vp]),
        "orc_stage_gradients": (None, [vp]),
        "orc_stage_momentum": (None, [vp, C.c_int]),
        "orc_stage_rhie_chow": (None, [vp]),
        "orc_stage_pressure": (None, [vp]),
        "orc_stage_cg": (C.c_int, [vp, _dp]),
        "orc_stage_correct": (None, [vp, _dp]),
        "orc_set_momentum_diagonal": (None, [vp, _dp]),
        "orc_ldu_multiply": (None, [vp, _dp, _dp, _dp, _dp, _dp]),
        "orc_ldu_gauss_seidel": (None, [vp, _dp, _dp, _dp, _dp, _dp, C.c_int]),
        "orc_ldu_cg": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_green_gauss": (C.c_int, [vp, _dp, _dp]),
        "orc_surface": (C.c_int, [vp, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, _dp, C.c_int, C.POINTER(Forces)]),
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
        _lib = _load()
    return _lib


def set_sum_block(block: int) -> None:
    """0: global sums in the reference's sequential order (the default);
    B > 0: blocks of B cells, then the block totals; -1: pairwise
    (process-wide)."""
    lib().orc_set_sum_block(int(block))


def set_tanh_variant(variant: int) -> None:
    """0: std::tanh in the SST blending functions (the default); 1: (1 - e^-2x)
    / (1 + e^-2x), a few ulps away (process-wide)."""
    lib().orc_set_tanh_variant(int(variant))


def default_settings() -> Settings:
    s = Settings()
    lib().orc_default_settings(C.byref(s))
    return s


def default_sst() -> SstConstants:
    c = SstConstants()
    lib().orc_default_sst(C.byref(c))
    return c


def cgrid_options(resolution: str = "coarse", **overrides) -> CGridOptions:
    o = CGridOptions()
    idx = {"coarse": 0, "medium": 1, "fine": 2}[resolution]
    if lib().orc_cgrid_defaults(idx, C.byref(o)) != 0:
        raise ValueError(resolution)
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


class Mesh:
    """Reference-numbered structured mesh (src/mesh/Mesh.cpp:89-339)."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle rejected the mesh options")
        self.h = handle
        ni, nj, nc, nf = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().orc_mesh_dims(handle, C.byref(ni), C.byref(nj), C.byref(nc), C.byref(nf))
        self.nodes_i, self.nodes_j, self.cells, self.faces = ni.value, nj.value, nc.value, nf.value
        self.ni, self.nj = self.nodes_i - 1, self.nodes_j - 1  # cells per direction
        self.centroid = np.zeros((self.cells, 2))
        self.area = np.zeros(self.cells)
        self.cell_faces = np.zeros((self.cells, 4), dtype=np.int32)
        lib().orc_mesh_cells(handle, self.centroid.reshape(-1), self.area, self.cell_faces.reshape(-1))
        self.fcentre = np.zeros((self.faces, 2))
        self.fnormal = np.zeros((self.faces, 2))
        self.farea = np.zeros(self.faces)
        self.owner = np.zeros(self.faces, dtype=np.int32)
        self.neighbour = np.zeros(self.faces, dtype=np.int32)
        self.boundary = np.zeros(self.faces, dtype=np.int32)
        self.partner = np.zeros(self.faces, dtype=np.int32)
        lib().orc_mesh_faces(handle, self.fcentre.reshape(-1), self.fnormal.reshape(-1), self.farea,
                             self.owner, self.neighbour, self.boundary, self.partner)
        self.node_xy = np.zeros((self.nodes_i * self.nodes_j, 2))
        lib().orc_mesh_nodes(handle, self.node_xy.reshape(-1))
        q = np.zeros(7)
        lib().orc_mesh_quality(handle, q)
        self.min_area, self.max_area, self.max_aspect, self.max_nonortho_deg = q[:4]
        self.inverted, self.min_wall_spacing, self.total_area = int(q[4]), q[5], q[6]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mesh_free(self.h)
            self.h = None

    def cell_index(self, i, j):
        return j * self.ni + i

    @property
    def j_face_count(self):
        return self.ni * self.nodes_j


def cgrid(m: int, p: int, t: int, chord: float = 1.0, options: CGridOptions | None = None, **kw) -> Mesh:
    o = options if options is not None else cgrid_options(kw.pop("resolution", "coarse"), **kw)
    return Mesh(lib().orc_cgrid(m, p, t, chord, C.byref(o)))


def box(**kw) -> Mesh:
    o = BoxOptions(origin_x=0.0, origin_y=0.0, length=1.0, height=1.0, cells_x=40, cells_y=20, grading=0,
                   first_cell_height=0.01, left=FARFIELD, right=OUTLET, lower=WALL, upper=WALL)
    for k, v in kw.items():
        setattr(o, k, v)
    return Mesh(lib().orc_box(C.byref(o)))


def face_conditions(mesh: Mesh, speed=50.0, alpha_deg=0.0, reference_pressure=0.0):
    kind = np.zeros(mesh.faces, dtype=np.int32)
    vel = np.zeros((mesh.faces, 2))
    pres = np.zeros(mesh.faces)
    lib().orc_face_conditions(mesh.h, speed, alpha_deg, reference_pressure, kind, vel.reshape(-1), pres)
    return kind, vel, pres


class Solver:
    """src/solver/SimpleSolver.hpp:159-197 restated on the CPU."""

    def __init__(self, mesh: Mesh, kind, velocity, pressure, settings: Settings | None = None):
        self.mesh = mesh
        self.settings = settings or default_settings()
        self.h = lib().orc_solver_create(mesh.h, np.ascontiguousarray(kind, dtype=np.int32),
                                         np.ascontiguousarray(velocity, dtype=np.float64).reshape(-1),
                                         np.ascontiguousarray(pressure, dtype=np.float64),
                                         C.byref(self.settings))
        if not self.h:
            raise ValueError("oracle rejected the solver settings")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_solver_free(self.h)
            self.h = None

    def set_model(self, model: int, constants: SstConstants | None = None):
        ptr = C.cast(C.pointer(constants), C.c_void_p) if constants is not None else None
        if lib().orc_solver_set_model(self.h, model, ptr) != 0:
            raise ValueError("model initialise failed")

    def initialise(self, u, v, p, rho, mu):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        if lib().orc_solver_initialise(self.h, f(u), f(v), f(p), f(rho), f(mu)) != 0:
            raise ValueError("initialise failed")

    def set_settings(self, settings: Settings):
        self.settings = settings
        lib().orc_solver_set_settings(self.h, C.byref(settings))

    def iterate(self, count: int = 1) -> Monitor:
        m = Monitor()
        lib().orc_solver_iterate(self.h, count, C.byref(m))
        return m

    def field(self):
        n = self.mesh.cells
        u, v, p, mu = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        lib().orc_solver_field(self.h, u, v, p, mu)
        return u, v, p, mu

    def set_field(self, u, v, p):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        lib().orc_solver_set_field(self.h, f(u), f(v), f(p))

    def flux(self):
        out = np.zeros(self.mesh.faces)
        lib().orc_solver_flux(self.h, out)
        return out

    def set_flux(self, flux):
        lib().orc_solver_set_flux(self.h, np.ascontiguousarray(flux, dtype=np.float64))

    def turbulence(self):
        n = self.mesh.cells
        arrs = [np.zeros(n) for _ in range(6)]
        if lib().orc_solver_turbulence(self.h, *arrs) != 0:
            return None
        return dict(zip(("k", "omega", "mut", "wall_distance", "f1", "cross_diffusion"), arrs))

    def set_turbulence(self, k, omega, mut):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        lib().orc_solver_set_turbulence(self.h, f(k), f(omega), f(mut))

    def divergence(self):
        out = np.zeros(self.mesh.cells)
        lib().orc_solver_divergence(self.h, out)
        return out

    def array(self, name: str):
        cap = 2 * max(self.mesh.faces, self.mesh.cells)
        out = np.zeros(cap)
        n = lib().orc_solver_array(self.h, name.encode(), out, cap)
        if n < 0:
            raise KeyError(name)
        return out[:n].copy()

    def freestream_k(self):
        return lib().orc_solver_freestream_k(self.h)

    def freestream_omega(self):
        return lib().orc_solver_freestream_omega(self.h)

    def max_yplus(self):
        return lib().orc_solver_max_yplus(self.h)


def uniform_field(cells: int, speed=50.0, alpha_deg=0.0, density=1.225, reynolds=1e6, chord=1.0,
                  reference_pressure=0.0):
    """FlowField::uniform (src/flow/FlowField.cpp:21-47) with Freestream.cpp:9-40."""
    a = alpha_deg * np.pi / 180.0
    mu = density * speed * chord / reynolds
    return (np.full(cells, speed * np.cos(a)), np.full(cells, speed * np.sin(a)),
            np.full(cells, reference_pressure), np.full(cells, density), np.full(cells, mu))


def surface(mesh: Mesh, u, v, p, mu, speed=50.0, alpha_deg=0.0, density=1.225, reference_pressure=0.0,
            chord=1.0):
    """extractSurface + integrateForces (src/post/SurfaceData.cpp:208-348, Forces.cpp:60-130).

    Returns (forces, rows) with rows = [side, face, x, y, x/c, arc, p, Cp, tau_w, Cf, nearStag, ds].
    """
    cap = 12 * mesh.ni
    rows = np.zeros(cap)
    f = Forces()
    g = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    n = lib().orc_surface(mesh.h, g(u), g(v), g(p), g(mu), speed, alpha_deg, density, reference_pressure,
                          chord, rows, cap, C.byref(f))
    if n < 0:
        raise ValueError("no aerofoil surface")
    return f, rows[:12 * n].reshape(n, 12)
