"""The oracle's surface extraction and force integration (orc_surface, the
restatement of SurfaceData.cpp:27-348 and Forces.cpp:16-130) pinned by
checks that share no code with it:

* the closed-contour identities of ForceTests.cpp:93-158: a uniform pressure
  exerts no force and no moment, still fluid no friction, the parts add up to
  the whole, and the wind axes keep the body force's magnitude;
* a pressure linear in x and y, imposed at the wall face centres: the
  midpoint rule is exact on each straight wall segment, so the pressure force
  is -(a, b) times the area of the wall polygon (the divergence theorem,
  polygon area and centroid by the shoelace formula), and the moment about the
  quarter chord follows from the polygon's centroid within the midpoint
  rule's bound for the quadratic moment integrand;
* Cp and Cf by their definitions (SurfaceDataTests.cpp:131-162, 276-281) and
  the wall shear of a profile linear in wall distance exact, of a quadratic
  one within 1 % (SurfaceDataTests.cpp:186-274);
* separation: a linear profile whose slope changes sign between two upper
  stations separates at their mid-chord (equal and opposite shear), and the
  stagnation point is the wall face whose cell holds the highest pressure;
* solved flows (ForceTests.cpp:253-306): a symmetric section at zero
  incidence gives no lift or moment and positive drag, lift follows the sign
  of the incidence and grows with it."""
import numpy as np
import pytest

import oracle_py as O
from helpers import SMALL

Q_STREAM = dict(speed=1.0, density=2.0)  # q = 1 exactly, so coefficients are forces


@pytest.fixture(scope="module")
def mesh():
    """sharedMesh() of ForceTests.cpp:40-53: NACA 0012, closed edge, coarse preset."""
    return O.cgrid(0, 0, 12, options=O.cgrid_options("coarse"))


def _wall(m):
    """Wall faces of the j = 0 row in i order, with their nodes (i, 0), (i + 1, 0)."""
    faces = np.array([f for f in range(m.ni) if m.boundary[f] == O.WALL])
    a = m.node_xy[faces]
    b = m.node_xy[faces + 1]
    return faces, a, b


def _surface(m, u, v, p, mu, alpha=0.0, pref=0.0):
    return O.surface(m, u, v, p, mu, alpha_deg=alpha, reference_pressure=pref, **Q_STREAM)


def _constant(m, p, vel, mu=1.0):
    n = m.cells
    return np.full(n, vel[0]), np.full(n, vel[1]), np.full(n, float(p)), np.full(n, mu)


def test_wall_faces_form_a_closed_polygon(mesh):
    faces, a, b = _wall(mesh)
    assert np.array_equal(a[1:], b[:-1])
    assert np.array_equal(a[0], b[-1])  # closed trailing edge
    t = (b - a) / np.hypot(*(b - a).T)[:, None]
    n = mesh.fnormal[faces] / np.hypot(*mesh.fnormal[faces].T)[:, None]
    assert np.max(np.abs(np.sum(t * n, axis=1))) < 1e-12


def test_uniform_pressure_exerts_no_net_force(mesh):
    f, _ = _surface(mesh, *_constant(mesh, 7.5, (0.0, 0.0)))
    for name in ("fx_pressure", "fy_pressure", "pitching_moment", "cl", "cd"):
        assert abs(getattr(f, name)) < 1e-12, name
    f, _ = _surface(mesh, *_constant(mesh, 1.0e5, (0.0, 0.0)), pref=1.0e5)
    assert abs(f.fx_pressure) < 1e-9 and abs(f.fy_pressure) < 1e-9


def test_still_fluid_exerts_no_friction(mesh):
    f, rows = _surface(mesh, *_constant(mesh, 3.0, (0.0, 0.0)))
    assert abs(f.fx_friction) < 1e-14 and abs(f.fy_friction) < 1e-14 and abs(f.cd_friction) < 1e-14
    assert np.all(rows[:, 8] == 0.0)
    assert f.upper_separation == -1.0 and f.lower_separation == -1.0


def test_parts_add_up_and_wind_axes_keep_the_magnitude(mesh):
    field = _constant(mesh, 2.0, (0.6, 0.2))
    for alpha in (-12.0, -3.0, 0.0, 7.5, 15.0):
        f, _ = _surface(mesh, *field, alpha=alpha)
        assert abs(f.cd - (f.cd_pressure + f.cd_friction)) < 1e-12
        assert abs(f.cd - f.drag) < 1e-12 and abs(f.cl - f.lift) < 1e-12  # q c = 1
        body = np.hypot(f.fx_pressure + f.fx_friction, f.fy_pressure + f.fy_friction)
        assert abs(body - np.hypot(f.lift, f.drag)) < 1e-12, alpha
    # the rotation's direction: at 90 deg the stream runs along +y, so drag is
    # the body y force and lift the body -x force
    f, _ = _surface(mesh, *field, alpha=90.0)
    assert abs(f.drag - (f.fy_pressure + f.fy_friction)) < 1e-12
    assert abs(f.lift + (f.fx_pressure + f.fx_friction)) < 1e-12


def _polygon(a, b):
    """Signed shoelace area and centroid of the closed polygon through the segments."""
    cross = a[:, 0] * b[:, 1] - b[:, 0] * a[:, 1]
    area = 0.5 * np.sum(cross)
    cx = np.sum((a[:, 0] + b[:, 0]) * cross) / (6.0 * area)
    cy = np.sum((a[:, 1] + b[:, 1]) * cross) / (6.0 * area)
    return abs(area), cx, cy


@pytest.mark.parametrize("ga,gb,c", [(0.7, -1.3, 4.0), (-0.4, 2.0, -1.0)])
def test_linear_pressure_force_is_minus_its_gradient_times_the_area(mesh, ga, gb, c):
    faces, a, b = _wall(mesh)
    area, xc, yc = _polygon(a, b)
    assert abs(area - 0.0821) < 1e-3  # 0.685 t c^2 for the closed-edge NACA 0012
    u, v, p, mu = _constant(mesh, 0.0, (0.0, 0.0))
    centre = mesh.fcentre[faces]
    p[mesh.owner[faces]] = ga * centre[:, 0] + gb * centre[:, 1] + c
    f, _ = _surface(mesh, u, v, p, mu)
    assert abs(f.fx_pressure - (-ga * area)) < 1e-12
    assert abs(f.fy_pressure - (-gb * area)) < 1e-12
    # pitching moment (nose up positive) about (c/4, 0): the clockwise moment of
    # -p n over the body, A (b (xc - 1/4) - a yc) by the divergence theorem;
    # the midpoint rule errs by at most len^3/24 |f''| per segment, and
    # |f''| <= 2 |grad p| for f = (r - r_ref) x (p n)
    expected = area * (gb * (xc - 0.25) - ga * yc)
    seg = np.hypot(*(b - a).T)
    bound = np.sum(seg ** 3) / 24.0 * 2.0 * np.hypot(ga, gb)
    assert abs(f.pitching_moment - expected) <= bound
    assert bound < 1e-2 * abs(expected)


def test_pressure_coefficient_is_pressure_over_dynamic_pressure(mesh):
    _, rows = _surface(mesh, *_constant(mesh, 0.0, (0.0, 0.0)))
    assert np.max(np.abs(rows[:, 7])) < 1e-12
    _, rows = _surface(mesh, *_constant(mesh, 1.0, (0.0, 0.0)))
    assert np.max(np.abs(rows[:, 7] - 1.0)) < 1e-12
    _, rows = _surface(mesh, *_constant(mesh, 501.0, (0.0, 0.0)), pref=500.0)
    assert np.max(np.abs(rows[:, 7] - 1.0)) < 1e-12


def test_stations_cover_both_surfaces_from_the_leading_edge(mesh):
    f, rows = _surface(mesh, *_constant(mesh, 0.0, (1.0, 0.0)))
    assert f.n_upper > 10 and f.n_lower > 10 and f.stations == f.n_upper + f.n_lower
    for side in (rows[:f.n_upper], rows[f.n_upper:]):
        assert abs(side[0, 4]) < 0.01 and abs(side[-1, 4] - 1.0) < 0.01
        assert side[0, 5] == 0.0 and np.all(np.diff(side[:, 5]) > 0.0)
        assert 0.9 < side[-1, 5] < 1.2  # half the perimeter, give or take
    assert np.all(rows[:f.n_upper, 0] == 1.0) and np.all(rows[f.n_upper:, 0] == 0.0)


def _profile_field(m, slope, curvature, mu):
    """wallProfileField of SurfaceDataTests.cpp:186-235: speed slope d +
    curvature d^2 along the nearest wall face's tangent, d the distance from
    that face along its normal into the fluid."""
    faces, a, b = _wall(m)
    t = (b - a) / np.hypot(*(b - a).T)[:, None]
    n = -m.fnormal[faces] / np.hypot(*m.fnormal[faces].T)[:, None]
    centre = m.fcentre[faces]
    gap = np.hypot(m.centroid[:, None, 0] - centre[None, :, 0], m.centroid[:, None, 1] - centre[None, :, 1])
    near = np.argmin(gap, axis=1)
    d = np.abs(np.sum((m.centroid - centre[near]) * n[near], axis=1))
    speed = slope * d + curvature * d * d
    return speed * t[near, 0], speed * t[near, 1], np.zeros(m.cells), np.full(m.cells, mu)


def test_wall_shear_is_exact_for_a_linear_profile(mesh):
    f, rows = _surface(mesh, *_profile_field(mesh, 7.0, 0.0, 3.0))
    mid = rows[f.n_upper // 4:3 * f.n_upper // 4]
    assert np.max(np.abs(np.abs(mid[:, 8]) - 21.0)) <= 1e-6 * 21.0
    assert np.max(np.abs(mid[:, 9] - mid[:, 8])) < 1e-12  # Cf = tau / q, q = 1


def test_wall_shear_is_second_order_accurate(mesh):
    f, rows = _surface(mesh, *_profile_field(mesh, 1.0, 1000.0, 1.0))
    mid = rows[f.n_upper // 4:3 * f.n_upper // 4]
    assert np.max(np.abs(np.abs(mid[:, 8]) - 1.0)) <= 0.01


def test_separation_and_stagnation_points(mesh):
    """A profile linear in wall distance, built cell by cell over the two rows
    above each wall face, with slope +s along the flow (away from the leading
    edge) everywhere except the upper surface aft of mid-chord, where it is -s:
    the upper shear changes sign between the two stations either side of
    x = 0.5 with equal magnitudes, so separation sits at their mid x/c; the
    lower surface stays attached. The pressure -x at the wall cells puts the
    stagnation point at the face whose cell is furthest forward."""
    faces, a, b = _wall(mesh)
    ni = mesh.ni
    t = (b - a) / np.hypot(*(b - a).T)[:, None]
    n = -mesh.fnormal[faces] / np.hypot(*mesh.fnormal[faces].T)[:, None]
    centre = mesh.fcentre[faces]
    owner = mesh.owner[faces]
    u, v, p, mu = _constant(mesh, 0.0, (0.0, 0.0), mu=2.0)
    p[owner] = -mesh.centroid[owner, 0]
    stag = int(np.argmax(p[owner]))
    le = int(np.argmin(a[:, 0]))  # the wall face starting at the foremost node
    upper = np.arange(len(faces)) >= le
    sign = np.where(upper, 1.0, -1.0)  # +t runs LE -> TE on the upper side only
    sign[upper & (centre[:, 0] > 0.5)] *= -1.0
    for row in (0, 1):
        cells = owner + row * ni
        d = np.abs(np.sum((mesh.centroid[cells] - centre) * n, axis=1))
        u[cells] = sign * 3.0 * d * t[:, 0]
        v[cells] = sign * 3.0 * d * t[:, 1]
    f, rows = _surface(mesh, u, v, p, mu)
    assert f.stagnation_xc == centre[stag, 0]
    assert np.count_nonzero(rows[:, 10]) == 1 and rows[rows[:, 10] == 1.0, 1][0] == faces[stag]
    up = rows[:f.n_upper]
    k = int(np.argmax(up[:, 2] > 0.5))
    assert abs(up[k - 1, 8] - 6.0) < 1e-9 and abs(up[k, 8] + 6.0) < 1e-9
    assert abs(f.upper_separation - 0.5 * (up[k - 1, 4] + up[k, 4])) < 1e-9
    assert f.lower_separation == -1.0


@pytest.fixture(scope="module")
def solved():
    """solvedForces of ForceTests.cpp:219-250: the small NACA 0012 C-grid,
    U = 1, rho = 2, Re 500, default settings, 900 iterations from the
    uniform stream."""
    m = O.cgrid(0, 0, 12, options=O.cgrid_options("coarse", **SMALL))
    out = {}
    for alpha in (0.0, 2.0, 8.0, -8.0):
        kind, vel, pres = O.face_conditions(m, 1.0, alpha)
        s = O.Solver(m, kind, vel, pres, None)
        s.initialise(*O.uniform_field(m.cells, speed=1.0, alpha_deg=alpha, density=2.0, reynolds=500.0))
        for _ in range(9):
            mon = s.iterate(100)
            assert np.isfinite(mon.continuity)
        out[alpha], _ = _surface(m, *s.field(), alpha=alpha)
    return out


def test_symmetric_section_at_zero_incidence_produces_no_lift(solved):
    f = solved[0.0]
    assert abs(f.cl) < 5e-3 and abs(f.cm) < 5e-3
    assert f.cd > 0.05 and f.cd_friction > 0.0


def test_lift_follows_the_sign_of_incidence_and_grows_with_it(solved):
    up, down, low = solved[8.0], solved[-8.0], solved[2.0]
    assert up.cl > 0.0 > down.cl
    assert abs(up.cl + down.cl) < 1e-2 and abs(up.cd - down.cd) < 1e-2
    assert 0.0 < low.cl < up.cl
