"""cfdgpu_divergence_noalloc, called through the C ABI, against
cfdgpu_divergence on the same state: bit-identical, on two calls in a row
(the second on a smaller mesh, so the shared staging array still holds the
first call's longer field), and on the first solver again. (The staging
array holds 1,048,576 cells, the most cells a solver accepts, so no solver
reaches its size check.)"""
import ctypes as C

import numpy as np
import pytest

from helpers import SMALL

pytestmark = pytest.mark.gpu


def _solver(G, iterations, **mesh_kw):
    mesh = G.cgrid(0, 0, 12, options=G.cgrid_options("coarse", **mesh_kw))
    kind, vel, pres = G.face_conditions(mesh, 1.0, 3.0)
    s = G.SimpleSolver(mesh, kind, vel, pres, G.default_settings())
    s.set_turbulence_model(G.MODEL_SST)
    s.initialise(*G.uniform_field(mesh.cells, speed=1.0, alpha_deg=3.0, density=1.0, reynolds=1e5))
    if iterations:
        s.iterate(iterations)
    return mesh, s


def _both(G, mesh, s):
    lib = G.lib()
    ref = np.zeros(mesh.cells)
    got = np.full(mesh.cells, np.nan)
    assert lib.cfdgpu_divergence(s.h, ref) == 0
    assert lib.cfdgpu_divergence_noalloc(s.h, got) == 0
    return ref, got


def test_noalloc_divergence_matches_the_allocating_entry_point(G, gpu_available):
    print("case: divergence_noalloc on 512x128 then 118x27 (SST, after 20 iterations)", flush=True)
    big = _solver(G, 20, surface_points=193, wake_points=64, normal_points=129, first_layer_height=2.5e-5)
    ref, got = _both(G, *big)
    assert np.array_equal(ref, got) and np.abs(ref).max() > 0
    small = _solver(G, 20, **SMALL)
    ref2, got2 = _both(G, *small)
    assert got2.shape == ref2.shape and np.array_equal(ref2, got2)
    # the first solver again, after the smaller call
    ref3, got3 = _both(G, *big)
    assert np.array_equal(ref3, got3)

