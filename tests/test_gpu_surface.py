"""Surface extraction and force integration on the device
(cfdgpu_solver_surface): the same rows and forces, bit for bit, as the host
routine (SurfaceData.cpp:208-348, Forces.cpp:60-130, restated in post.cpp)
applied to the downloaded field, on converging SST fields and on a random
field whose wall shear changes sign on both sides (the separation scans of
SurfaceData.cpp:179-204)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(G, m, p, t, alpha, **mesh):
    grid = G.cgrid(m, p, t, options=G.cgrid_options("coarse", **mesh))
    kind, vel, pres = G.face_conditions(grid, 1.0, alpha)
    s = G.SimpleSolver(grid, kind, vel, pres, G.default_settings(viscosity_ratio=10.0))
    s.set_turbulence_model(G.MODEL_SST)
    s.initialise(*G.uniform_field(grid.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=1e5))
    return grid, s


def _same(G, grid, s, alpha):
    fh, rh = G.surface(grid, *s.field(), speed=1.0, alpha_deg=alpha, density=1.0, reference_pressure=0.0)
    fd, rd = s.surface(speed=1.0, alpha_deg=alpha, density=1.0, reference_pressure=0.0)
    assert rh.shape == rd.shape and np.array_equal(rh, rd)
    for name, _ in G.Forces._fields_:
        assert getattr(fh, name) == getattr(fd, name), name
    return fh


@pytest.mark.parametrize("section,alpha", [((0, 0, 12), 0.0), ((2, 4, 12), 4.0), ((4, 4, 15), -6.0)])
def test_device_surface_matches_the_host_routine(G, gpu_available, section, alpha):
    print(f"case: NACA {section} alpha {alpha}, SST, 300 iterations", flush=True)
    grid, s = _solver(G, *section, alpha, surface_points=56, wake_points=24, normal_points=36)
    s.iterate(300)
    f = _same(G, grid, s, alpha)
    assert np.isfinite(f.cl) and f.stations == f.n_upper + f.n_lower


def test_device_surface_separation_scans(G, gpu_available):
    print("case: random field on NACA 2412 (sign changes of the wall shear)", flush=True)
    grid, s = _solver(G, 2, 4, 12, 6.0, surface_points=56, wake_points=24, normal_points=36)
    rng = np.random.default_rng(3)
    s.set_field(rng.standard_normal(grid.cells), rng.standard_normal(grid.cells), rng.standard_normal(grid.cells))
    f = _same(G, grid, s, 6.0)
    assert f.upper_separation >= 0.0 and f.lower_separation >= 0.0
    assert f.upper_separation_face >= 0 and f.lower_separation_face >= 0
