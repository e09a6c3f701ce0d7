"""Pins of the oracle's k-omega SST (CPU), from the reference's
tests/TurbulenceTests.cpp: published constants, freestream values, the
defining mu_t relation, the wall-cell omega, the decaying-turbulence closed
form and the laminar control closure."""
import numpy as np
import pytest

from helpers import check_decaying_turbulence, decaying_turbulence_case, sst_small_cgrid

O = pytest.importorskip("oracle_py")


def test_blended_constants_match_the_published_values():
    """TurbulenceTests.cpp:209-213: gamma_i = beta_i / beta* - sigma_w_i kappa^2 / sqrt(beta*)."""
    c = O.default_sst()
    g1 = c.beta1 / c.beta_star - c.sigma_w1 * c.kappa ** 2 / np.sqrt(c.beta_star)
    g2 = c.beta2 / c.beta_star - c.sigma_w2 * c.kappa ** 2 / np.sqrt(c.beta_star)
    assert abs(g1 - 0.5532) < 1e-4 and abs(g2 - 0.4403) < 1e-4


def test_freestream_values_follow_from_intensity_and_viscosity_ratio():
    """TurbulenceTests.cpp:227-266: k = 1.5 (I U)^2, omega = rho k / (r mu); mu_t is
    r mu in the freestream (d > 1) and at most that, and positive, nearer the wall."""
    mesh, _, field, s = sst_small_cgrid(O, intensity=0.01, viscosity_ratio=5.0)
    mu = field[4][0]
    assert abs(s.freestream_k() - 1.5 * 0.01 ** 2) < 1e-12
    w_in = s.freestream_omega()
    assert abs(w_in - 1.0 * s.freestream_k() / (5.0 * mu)) < 1e-9 * w_in
    t = s.turbulence()
    far = t["wall_distance"] > 1.0
    assert far.sum() > 100
    assert np.max(np.abs(t["mut"][far] - 5.0 * mu)) < 1e-12 * mu
    assert np.all(t["mut"][~far] <= 5.0 * mu * (1 + 1e-12)) and np.all(t["mut"][~far] > 0)


def test_eddy_viscosity_is_density_times_energy_over_dissipation():
    """TurbulenceTests.cpp:270-286, at initialisation."""
    mesh, _, field, s = sst_small_cgrid(O)
    t = s.turbulence()
    np.testing.assert_allclose(t["mut"], field[3] * t["k"] / t["omega"], rtol=1e-9, atol=0)


def test_wall_cell_omega_follows_the_sublayer_solution():
    """TurbulenceTests.cpp:290-334: after one iteration the wall cells hold
    sqrt(w_vis^2 + w_log^2), w_vis = 6 nu / (beta1 d^2), w_log = sqrt(k) / (beta*^1/4 kappa d)."""
    mesh, kind, field, s = sst_small_cgrid(O)
    s.iterate(1)
    t = s.turbulence()
    c = O.default_sst()
    walls = np.nonzero(kind == O.K_WALL)[0]
    assert len(walls) > 10
    own = mesh.owner[walls]
    nu = field[4][own] / field[3][own]
    d = t["wall_distance"][own]
    visc = 6.0 * nu / (c.beta1 * d * d)
    logl = np.sqrt(np.maximum(t["k"][own], 0.0)) / (c.beta_star ** 0.25 * c.kappa * d)
    np.testing.assert_allclose(t["omega"][own], np.sqrt(visc ** 2 + logl ** 2), rtol=1e-6, atol=0)
    assert np.all(t["k"][own] >= 0) and np.all(np.isfinite(t["k"][own]))


def test_decaying_turbulence_follows_the_analytic_solution():
    """TurbulenceTests.cpp:459-581."""
    mesh, s = decaying_turbulence_case(O)
    for _ in range(30):
        mon = s.iterate(100)
        assert np.isfinite(mon.continuity)
    t = s.turbulence()
    check_decaying_turbulence(mesh, t["k"], t["omega"], t["f1"], s.freestream_k(), s.freestream_omega())


def test_laminar_closure_is_indistinguishable_from_no_closure():
    """TurbulenceTests.cpp:96-124: 40 iterations, bit-identical fields."""
    a = sst_small_cgrid(O, reynolds=500.0, model=0)[3]
    b = sst_small_cgrid(O, reynolds=500.0, model=1)[3]
    a.iterate(40)
    b.iterate(40)
    for x, y in zip(a.field(), b.field()):
        assert np.array_equal(x, y)


def test_wall_distance_is_the_nearest_wall_face_centre():
    """TurbulenceModel.cpp:28-65: each cell's distance to the nearest wall-face
    centre, by brute force over all pairs (NumPy, independent of the oracle's
    loop): equal to the last bit, since both take the minimum of squared
    distances and the square root last."""
    mesh, _, _, s = sst_small_cgrid(O)
    t = s.turbulence()
    walls = mesh.fcentre[mesh.boundary == O.WALL]
    assert len(walls) > 10
    dx = mesh.centroid[:, None, 0] - walls[None, :, 0]
    dy = mesh.centroid[:, None, 1] - walls[None, :, 1]
    brute = np.sqrt(np.min(dx * dx + dy * dy, axis=1))
    assert np.array_equal(t["wall_distance"], brute)
