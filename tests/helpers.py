"""Shared builders for parity tests: the same mesh, conditions and starting field
handed to the HIP path (cfdgpu) and to the FP64 oracle (oracle_py)."""
from __future__ import annotations

import numpy as np

# Small C-grid used throughout the reference's tests: NACA 0012, coarse preset
# with 40 surface / 20 wake / 28 normal points -> 118 x 27 cells
# (tests/TurbulenceTests.cpp:50-65, ValidationTests.cpp:458-505).
SMALL = dict(surface_points=40, wake_points=20, normal_points=28)


def cgrid_pair(G, O, m=0, p=0, t=12, resolution="coarse", **kw):
    gkw = {("first_layer_height" if k == "first_layer" else
            k + "_chords" if k in ("upstream", "downstream", "vertical") else k): v for k, v in kw.items()}
    gm = G.cgrid(m, p, t, options=G.cgrid_options(resolution, **gkw))
    om = O.cgrid(m, p, t, options=O.cgrid_options(resolution, **kw))
    return gm, om


def box_pair(G, O, **kw):
    return G.box(**kw), O.box(**kw)


def solver_pair(G, O, gm, om, kind, vel, pres, field, gsettings=None, osettings=None, model=0):
    gs = G.SimpleSolver(gm, kind, vel, pres, gsettings)
    os_ = O.Solver(om, kind, vel, pres, osettings)
    if model:
        gs.set_turbulence_model(model)
        os_.set_model(model)
    gs.initialise(*field)
    os_.initialise(*field)
    return gs, os_


def oracle_settings_like(O, gset):
    s = O.default_settings()
    for name, _ in O.Settings._fields_:
        setattr(s, name, getattr(gset, name))
    return s


def ldu_to_rows(om, upper, lower):
    """Reference LDU (per face) -> cell rows (aS, aN, aW, aE) in the device layout."""
    ni, nj = om.ni, om.nj
    nJf = ni * (nj + 1)
    aS = np.zeros(om.cells)
    aN = np.zeros(om.cells)
    aW = np.zeros(om.cells)
    aE = np.zeros(om.cells)
    for j in range(nj):
        for i in range(ni):
            c = j * ni + i
            below = j * ni + i
            if j > 0:
                aS[c] = lower[below]
            elif om.partner[below] >= 0:
                aS[c] = upper[below]
            if j < nj - 1:
                aN[c] = upper[(j + 1) * ni + i]
            if i > 0:
                aW[c] = lower[nJf + i * nj + j]
            if i < ni - 1:
                aE[c] = upper[nJf + (i + 1) * nj + j]
    return aS, aN, aW, aE


def rel_err(a, b):
    scale = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()) / scale


# ---------------------------------------------------------------------------
# The reference's exact-solution validation cases (tests/ValidationTests.cpp),
# built the same way for the HIP path (M = cfdgpu) and the oracle (M = oracle_py).
# ---------------------------------------------------------------------------

def _uniform(cells, u, rho, mu):
    return (np.full(cells, float(u[0])), np.full(cells, float(u[1])), np.zeros(cells), np.full(cells, rho),
            np.full(cells, mu))


def _settings(M, **kw):
    """cartesianSettings (ValidationTests.cpp:66-71): relaxation 0.7 / 0.3."""
    if hasattr(M, "Solver"):  # oracle
        s = M.default_settings()
        for k, v in kw.items():
            setattr(s, k, v)
    else:
        s = M.default_settings(**kw)
    s.velocity_relaxation = 0.7
    s.pressure_relaxation = 0.3
    return s


def validation_case(M, name):
    """Returns (mesh, kind, velocity, pressure, field, settings, exact) for one of
    uniform, couette, poiseuille, blasius, mass_balance. `exact` holds the
    case's constants."""
    if name == "uniform":  # ValidationTests.cpp:98-139
        mesh = M.box(length=2.0, height=1.0, cells_x=20, cells_y=15, grading=0, left=M.FARFIELD, right=M.OUTLET,
                     lower=M.FARFIELD, upper=M.FARFIELD)
        kind = np.full(mesh.faces, M.K_INTERNAL, dtype=np.int32)
        vel = np.zeros((mesh.faces, 2))
        bnd = mesh.neighbour < 0
        out = bnd & (mesh.boundary == M.OUTLET)
        kind[bnd] = M.K_INLET
        vel[bnd] = (1.0, 0.0)
        kind[out] = M.K_OUTLET
        vel[out] = 0.0
        return mesh, kind, vel, np.zeros(mesh.faces), _uniform(mesh.cells, (0.4, 0.0), 1.0, 0.05), \
            _settings(M), {"speed": 1.0}
    if name == "couette":  # :145-197
        mesh = M.box(length=2.0, height=1.0, cells_x=20, cells_y=20, grading=0, left=M.FARFIELD, right=M.OUTLET,
                     lower=M.WALL, upper=M.FARFIELD)
        kind = np.full(mesh.faces, M.K_INTERNAL, dtype=np.int32)
        vel = np.zeros((mesh.faces, 2))
        bnd = mesh.neighbour < 0
        for f in np.nonzero(bnd)[0]:
            if mesh.boundary[f] == M.WALL:
                kind[f] = M.K_WALL
            elif mesh.boundary[f] == M.OUTLET:
                kind[f] = M.K_OUTLET
            else:
                kind[f] = M.K_INLET
                vel[f] = (1.0, 0.0) if abs(mesh.fnormal[f, 1]) > 0.5 else (mesh.fcentre[f, 1], 0.0)
        return mesh, kind, vel, np.zeros(mesh.faces), _uniform(mesh.cells, (0.5, 0.0), 1.0, 0.05), \
            _settings(M), {"speed": 1.0, "height": 1.0}
    if name == "poiseuille":  # :203-275
        mesh = M.box(length=5.0, height=1.0, cells_x=50, cells_y=32, grading=0, left=M.FARFIELD, right=M.OUTLET,
                     lower=M.WALL, upper=M.WALL)
        exact = lambda y: 4.0 * y * (1.0 - y)
        kind = np.full(mesh.faces, M.K_INTERNAL, dtype=np.int32)
        vel = np.zeros((mesh.faces, 2))
        bnd = mesh.neighbour < 0
        for f in np.nonzero(bnd)[0]:
            if mesh.boundary[f] == M.FARFIELD:
                kind[f] = M.K_INLET
                vel[f] = (exact(mesh.fcentre[f, 1]), 0.0)
            elif mesh.boundary[f] == M.OUTLET:
                kind[f] = M.K_OUTLET
            else:
                kind[f] = M.K_WALL
        return mesh, kind, vel, np.zeros(mesh.faces), _uniform(mesh.cells, (0.6, 0.0), 1.0, 0.1), \
            _settings(M), {"exact": exact, "mu": 0.1}
    if name == "blasius":  # :281-377
        mesh = M.box(origin_x=-0.5, length=1.5, height=1.0, cells_x=110, cells_y=60, grading=1,
                     first_cell_height=1.0e-4, left=M.FARFIELD, right=M.OUTLET, lower=M.WALL, upper=M.OUTLET)
        kind = np.full(mesh.faces, M.K_INTERNAL, dtype=np.int32)
        vel = np.zeros((mesh.faces, 2))
        bnd = mesh.neighbour < 0
        for f in np.nonzero(bnd)[0]:
            if mesh.boundary[f] == M.FARFIELD:
                kind[f] = M.K_INLET
                vel[f] = (1.0, 0.0)
            elif mesh.boundary[f] == M.OUTLET:
                kind[f] = M.K_OUTLET
            elif mesh.fcentre[f, 0] < 0.0:
                kind[f] = M.K_INLET
                vel[f] = (1.0, 0.0)
            else:
                kind[f] = M.K_WALL
        return mesh, kind, vel, np.zeros(mesh.faces), _uniform(mesh.cells, (1.0, 0.0), 1.0, 1.0e-4), \
            _settings(M, scheme=1), {"mu": 1.0e-4}
    if name == "mass_balance":  # :383-431
        mesh = M.box(length=3.0, height=1.0, cells_x=30, cells_y=20, grading=0, left=M.FARFIELD, right=M.OUTLET,
                     lower=M.WALL, upper=M.WALL)
        kind = np.full(mesh.faces, M.K_INTERNAL, dtype=np.int32)
        vel = np.zeros((mesh.faces, 2))
        bnd = mesh.neighbour < 0
        for f in np.nonzero(bnd)[0]:
            if mesh.boundary[f] == M.FARFIELD:
                kind[f] = M.K_INLET
                vel[f] = (1.0, 0.0)
            elif mesh.boundary[f] == M.OUTLET:
                kind[f] = M.K_OUTLET
            else:
                kind[f] = M.K_WALL
        return mesh, kind, vel, np.zeros(mesh.faces), _uniform(mesh.cells, (1.0, 0.0), 1.0, 0.02), \
            _settings(M), {}
    raise KeyError(name)


def converge(solver, max_iterations, tol):
    """ValidationTests.cpp:75-90: iterate until all three residuals < tol. Returns
    (monitor, iterations taken)."""
    if hasattr(solver, "solve"):  # HIP path: the device-side stop rule is the same test
        rep = solver.solve(tol, max_iterations, check_every=50)
        return rep.last, rep.iterations
    for it in range(max_iterations):
        mon = solver.iterate(1)
        if mon.continuity < tol and mon.momentum_x < tol and mon.momentum_y < tol:
            return mon, it + 1
    return mon, max_iterations


def check_validation(name, mesh, u, v, p, mon, flux, exact, taken, max_iterations):
    """The assertions of ValidationTests.cpp for one case; returns a dict of the
    measured quantities (for messages)."""
    ni, nj = mesh.ni, mesh.nj
    cx, cy = mesh.centroid[:, 0], mesh.centroid[:, 1]
    if name != "mass_balance":
        assert taken < max_iterations, f"{name}: did not converge ({mon.continuity}, {mon.momentum_x})"
    if name == "uniform":
        assert np.max(np.abs(u - 1.0)) < 1e-10 and np.max(np.abs(v)) < 1e-10
        assert mon.max_divergence < 1e-10
    elif name == "couette":
        col = np.arange(nj) * ni + (ni - 2)
        assert np.max(np.abs(u[col] - cy[col])) < 1e-9
    elif name == "poiseuille":
        col = np.arange(nj) * ni + (ni - 2)
        assert np.max(np.abs(u[col] - exact["exact"](cy[col]))) < 0.005
        mid = nj // 2
        a, b = mid * ni + 4, mid * ni + (ni - 6)
        grad = (p[b] - p[a]) / (cx[b] - cx[a])
        expected = -8.0 * exact["mu"]
        assert abs(grad - expected) < 0.02 * abs(expected)
        assert mon.max_divergence < 1e-6
    elif name == "blasius":
        row = np.arange(ni)
        x, y = cx[row], cy[row]
        sel = (x >= 0.15) & (x <= 0.95)
        cf = exact["mu"] * u[row][sel] / y[sel] / 0.5
        blasius = 0.664 / np.sqrt(x[sel] / exact["mu"])
        assert sel.sum() > 10
        assert np.max(np.abs(cf - blasius) / blasius) < 0.08
        column = (ni) // 2
        assert np.max(u[np.arange(nj) * ni + column]) < 1.03
        assert mon.max_divergence < 1e-5
    elif name == "mass_balance":
        net = np.sum(flux[mesh.neighbour < 0])
        assert abs(net) < 1e-10


# ---------------------------------------------------------------------------
# k-omega SST cases of the reference's tests/TurbulenceTests.cpp
# ---------------------------------------------------------------------------

def sst_small_cgrid(M, reynolds=1.0e6, alpha=0.0, intensity=0.001, viscosity_ratio=1.0, model=2):
    """sharedMesh() + stream() of TurbulenceTests.cpp:50-90: NACA 0012, the
    small coarse C-grid, U = rho = 1."""
    mesh = M.cgrid(0, 0, 12, options=M.cgrid_options("coarse", **SMALL))
    kind, vel, pres = M.face_conditions(mesh, 1.0, alpha)
    field = M.uniform_field(mesh.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=reynolds)
    if hasattr(M, "Solver"):
        st = M.default_settings()
        st.intensity, st.viscosity_ratio = intensity, viscosity_ratio
        s = M.Solver(mesh, kind, vel, pres, st)
        if model:
            s.set_model(model)
    else:
        s = M.SimpleSolver(mesh, kind, vel, pres, M.default_settings(intensity=intensity,
                                                                      viscosity_ratio=viscosity_ratio))
        if model:
            s.set_turbulence_model(model)
    s.initialise(*field)
    return mesh, kind, field, s


def decaying_turbulence_case(M):
    """TurbulenceTests.cpp:484-581: a 10 x 1 box, 200 x 4 cells, far field on
    three sides and an outlet; U = 10, rho = 1, Re 1e5; I = 0.1 (k0 = 1.5),
    mu_t/mu = 100; 3,000 iterations."""
    mesh = M.box(length=10.0, height=1.0, cells_x=200, cells_y=4, grading=0, left=M.FARFIELD, right=M.OUTLET,
                 lower=M.FARFIELD, upper=M.FARFIELD)
    kind, vel, pres = M.face_conditions(mesh, 10.0, 0.0)
    field = M.uniform_field(mesh.cells, speed=10.0, alpha_deg=0.0, density=1.0, reynolds=1e5)
    if hasattr(M, "Solver"):
        st = M.default_settings()
        st.intensity, st.viscosity_ratio = 0.1, 100.0
        s = M.Solver(mesh, kind, vel, pres, st)
        s.set_model(2)
    else:
        s = M.SimpleSolver(mesh, kind, vel, pres, M.default_settings(intensity=0.1, viscosity_ratio=100.0))
        s.set_turbulence_model(2)
    s.initialise(*field)
    return mesh, s


def check_decaying_turbulence(mesh, k, omega, f1, k0, omega0):
    """The closed form of TurbulenceTests.cpp:459-581: with F1 = 0 (no wall)
    omega = omega0 / (1 + beta2 omega0 x / U), k = k0 (...)^(-beta*/beta2);
    3 % on omega and 3.5 % on k for 1 < x < 9."""
    beta2, beta_star, U = 0.0828, 0.09, 10.0
    assert np.max(f1) < 1e-6
    x = mesh.centroid[:, 0]
    sel = (x >= 1.0) & (x <= 9.0)
    growth = 1.0 + beta2 * omega0 * x[sel] / U
    w_exact = omega0 / growth
    k_exact = k0 * growth ** (-beta_star / beta2)
    assert sel.sum() > 50
    assert np.max(np.abs(omega[sel] - w_exact) / w_exact) < 0.03
    assert np.max(np.abs(k[sel] - k_exact) / k_exact) < 0.035
