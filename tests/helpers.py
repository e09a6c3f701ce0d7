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
