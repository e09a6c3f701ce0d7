"""The reference's k-omega SST closed forms (tests/TurbulenceTests.cpp) on the
HIP path: freestream values, mu_t = rho k / omega, the wall-cell omega, the
decaying-turbulence solution and the laminar control closure."""
import numpy as np
import pytest

from helpers import check_decaying_turbulence, decaying_turbulence_case, sst_small_cgrid

pytestmark = pytest.mark.gpu


def test_gpu_freestream_values_and_eddy_viscosity(G, gpu_available):
    print("case: SST freestream values, 118x27 C-grid, I=0.01, ratio 5", flush=True)
    mesh, _, field, s = sst_small_cgrid(G, intensity=0.01, viscosity_ratio=5.0)
    mu = field[4][0]
    k_in, w_in, _ = s.turbulence_info()
    assert abs(k_in - 1.5 * 0.01 ** 2) < 1e-12
    assert abs(w_in - k_in / (5.0 * mu)) < 1e-9 * w_in
    t = s.turbulence()
    far = t["wall_distance"] > 1.0
    assert far.sum() > 100
    assert np.max(np.abs(t["mut"][far] - 5.0 * mu)) < 1e-12 * mu
    np.testing.assert_allclose(t["mut"], field[3] * t["k"] / t["omega"], rtol=1e-9, atol=0)


def test_gpu_wall_cell_omega(G, gpu_available):
    print("case: SST wall-cell omega after one iteration, 118x27 C-grid", flush=True)
    mesh, kind, field, s = sst_small_cgrid(G)
    s.iterate(1)
    t = s.turbulence()
    c = G.default_sst()
    walls = np.nonzero(kind == G.K_WALL)[0]
    own = mesh.owner[walls]
    nu = field[4][own] / field[3][own]
    d = t["wall_distance"][own]
    visc = 6.0 * nu / (c.beta1 * d * d)
    logl = np.sqrt(np.maximum(t["k"][own], 0.0)) / (c.beta_star ** 0.25 * c.kappa * d)
    np.testing.assert_allclose(t["omega"][own], np.sqrt(visc ** 2 + logl ** 2), rtol=1e-6, atol=0)


def test_gpu_decaying_turbulence(G, gpu_available):
    print("case: decaying turbulence, 200x4 box, 3000 iterations", flush=True)
    mesh, s = decaying_turbulence_case(G)
    for _ in range(30):
        mon = s.iterate(100)
        assert np.isfinite(mon.continuity)
    t = s.turbulence()
    k_in, w_in, _ = s.turbulence_info()
    check_decaying_turbulence(mesh, t["k"], t["omega"], s.debug_array("f1"), k_in, w_in)


def test_gpu_laminar_closure_is_indistinguishable_from_no_closure(G, gpu_available):
    print("case: laminar closure vs none, 40 iterations", flush=True)
    a = sst_small_cgrid(G, reynolds=500.0, model=0)[3]
    b = sst_small_cgrid(G, reynolds=500.0, model=1)[3]
    a.iterate(40)
    b.iterate(40)
    for x, y in zip(a.field(), b.field()):
        assert np.array_equal(x, y)
