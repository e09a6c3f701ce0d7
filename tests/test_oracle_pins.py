"""Pins of the FP64 oracle's mesh, gradient and linear-algebra routines (CPU).

Each check compares the oracle with something it does not share code with: a
closed form, a dense brute-force computation, a textbook algorithm run on the
assembled matrix, or an invariant of the discretisation. The cases follow the
reference's own tests (tests/SolverTests.cpp, MeshTests.cpp, CGridTests.cpp,
Naca4Tests.cpp of sidhanth-sridhar/CFD-SIMULATOR); file:line on each.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import SMALL

O = pytest.importorskip("oracle_py")

WALL, FARFIELD, OUTLET, WAKECUT = O.WALL, O.FARFIELD, O.OUTLET, O.WAKECUT


def face_values(mesh, fn):
    return np.ascontiguousarray(fn(mesh.fcentre[:, 0], mesh.fcentre[:, 1]), dtype=np.float64)


def green_gauss(mesh, phi_f):
    g = np.zeros(2 * mesh.cells)
    assert O.lib().orc_green_gauss(mesh.h, phi_f, g) == mesh.cells
    return g.reshape(-1, 2)


def laplacian(mesh, rhs):
    """SolverTests.cpp:222-240: unit coupling on interior faces, diagonal + 1."""
    diag = np.ones(mesh.cells)
    upper = np.zeros(mesh.faces)
    lower = np.zeros(mesh.faces)
    interior = mesh.neighbour >= 0
    np.add.at(diag, mesh.owner[interior], 1.0)
    np.add.at(diag, mesh.neighbour[interior], 1.0)
    upper[interior] = -1.0
    lower[interior] = -1.0
    return diag, upper, lower, np.ascontiguousarray(rhs, dtype=np.float64)


def dense(mesh, diag, upper, lower):
    """The LDU convention (LinearSystem.hpp:29-72): row owner gets upper * x[neighbour],
    row neighbour gets lower * x[owner]."""
    A = np.diag(diag.copy())
    for f in np.nonzero(mesh.neighbour >= 0)[0]:
        o, n = mesh.owner[f], mesh.neighbour[f]
        A[o, n] += upper[f]
        A[n, o] += lower[f]
    return A


# --- gradient (Gradient.cpp:20-50) -----------------------------------------

def test_green_gauss_is_exact_for_a_linear_field_on_a_graded_box():
    """SolverTests.cpp:159-175: phi = 3x - 2y + 1 on a 10 x 12 graded box."""
    mesh = O.box(length=2.0, height=1.0, cells_x=10, cells_y=12, grading=1, first_cell_height=0.01)
    g = green_gauss(mesh, face_values(mesh, lambda x, y: 3.0 * x - 2.0 * y + 1.0))
    assert np.max(np.abs(g[:, 0] - 3.0)) < 1e-10
    assert np.max(np.abs(g[:, 1] + 2.0)) < 1e-10


def test_green_gauss_is_exact_for_a_linear_field_on_a_cgrid():
    """The same closure on the skewed, stretched cells of a C-grid (interior cells:
    every face value is the exact field value at the face centre)."""
    mesh = O.cgrid(2, 4, 12, options=O.cgrid_options("coarse", **SMALL))
    g = green_gauss(mesh, face_values(mesh, lambda x, y: 0.7 * x + 1.9 * y - 0.3))
    assert np.max(np.abs(g[:, 0] - 0.7)) < 1e-8
    assert np.max(np.abs(g[:, 1] - 1.9)) < 1e-8


def test_green_gauss_converges_at_second_order():
    """SolverTests.cpp:182-211: sin(2x) cos(3y) on 10, 20, 40 cells."""
    def err(n):
        mesh = O.box(length=1.0, height=1.0, cells_x=n, cells_y=n, grading=0)
        g = green_gauss(mesh, face_values(mesh, lambda x, y: np.sin(2 * x) * np.cos(3 * y)))
        x, y = mesh.centroid[:, 0], mesh.centroid[:, 1]
        dx = 2 * np.cos(2 * x) * np.cos(3 * y)
        dy = -3 * np.sin(2 * x) * np.sin(3 * y)
        return max(np.max(np.abs(g[:, 0] - dx)), np.max(np.abs(g[:, 1] - dy)))
    coarse, fine, finer = err(10), err(20), err(40)
    assert coarse > 0 and fine < coarse and finer < fine
    assert finer < coarse / 8.0


# --- LDU algebra (LinearSystem.cpp:18-210) ----------------------------------

def test_ldu_multiply_matches_the_dense_product():
    """SolverTests.cpp:242-270, strengthened: every row, asymmetric coefficients."""
    mesh = O.box(cells_x=5, cells_y=4, grading=0)
    rng = np.random.default_rng(7)
    diag = rng.uniform(2, 3, mesh.cells)
    upper = np.where(mesh.neighbour >= 0, rng.uniform(-1, 0, mesh.faces), 0.0)
    lower = np.where(mesh.neighbour >= 0, rng.uniform(-1, 0, mesh.faces), 0.0)
    x = np.arange(1, mesh.cells + 1, dtype=np.float64)
    y = np.zeros(mesh.cells)
    O.lib().orc_ldu_multiply(mesh.h, diag, upper, lower, x, y)
    np.testing.assert_allclose(y, dense(mesh, diag, upper, lower) @ x, rtol=0, atol=1e-12)


def test_gauss_seidel_is_the_lexicographic_in_place_sweep():
    """LinearSystem.cpp:77-116 against the textbook forward sweep on the dense matrix:
    x_c <- (b_c - sum_{k<c} a_ck x_k(new) - sum_{k>c} a_ck x_k(old)) / a_cc."""
    mesh = O.box(cells_x=6, cells_y=5, grading=0)
    rng = np.random.default_rng(3)
    diag = rng.uniform(4, 5, mesh.cells)
    upper = np.where(mesh.neighbour >= 0, rng.uniform(-1, 0, mesh.faces), 0.0)
    lower = np.where(mesh.neighbour >= 0, rng.uniform(-1, 0, mesh.faces), 0.0)
    b = rng.uniform(-1, 1, mesh.cells)
    x0 = rng.uniform(-1, 1, mesh.cells)
    x = x0.copy()
    O.lib().orc_ldu_gauss_seidel(mesh.h, diag, upper, lower, b, x, 2)
    A = dense(mesh, diag, upper, lower)
    ref = x0.copy()
    for _ in range(2):
        for c in range(mesh.cells):
            ref[c] = (b[c] - (A[c] @ ref - A[c, c] * ref[c])) / A[c, c]
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-13)


def test_gauss_seidel_drives_the_residual_down():
    """SolverTests.cpp:352-363: 200 sweeps of the 8 x 8 Laplacian, residual / 1e8."""
    mesh = O.box(cells_x=8, cells_y=8, grading=0)
    diag, upper, lower, b = laplacian(mesh, np.ones(mesh.cells))
    A = dense(mesh, diag, upper, lower)
    x = np.zeros(mesh.cells)
    before = np.sum(np.abs(b - A @ x))
    O.lib().orc_ldu_gauss_seidel(mesh.h, diag, upper, lower, b, x, 200)
    assert np.sum(np.abs(b - A @ x)) < 1e-8 * before


def test_cg_recovers_a_known_solution():
    """SolverTests.cpp:365-385: 12 x 12 Laplacian, rhs from sin(3x) cos(2y)."""
    mesh = O.box(cells_x=12, cells_y=12, grading=0)
    expected = np.sin(3 * mesh.centroid[:, 0]) * np.cos(2 * mesh.centroid[:, 1])
    diag, upper, lower, _ = laplacian(mesh, np.zeros(mesh.cells))
    b = dense(mesh, diag, upper, lower) @ expected
    x = np.zeros(mesh.cells)
    r0, r1 = C.c_double(), C.c_double()
    its = O.lib().orc_ldu_cg(mesh.h, diag, upper, lower, b, x, 1e-12, 500, C.byref(r0), C.byref(r1))
    assert 0 < its < 500
    assert r1.value <= 1e-12 * r0.value
    np.testing.assert_allclose(x, expected, rtol=0, atol=1e-8)


def test_cg_reports_an_already_solved_system():
    """SolverTests.cpp:387-395: zero right-hand side, zero iterations."""
    mesh = O.box(cells_x=4, cells_y=4, grading=0)
    diag, upper, lower, b = laplacian(mesh, np.zeros(mesh.cells))
    x = np.zeros(mesh.cells)
    r0, r1 = C.c_double(), C.c_double()
    assert O.lib().orc_ldu_cg(mesh.h, diag, upper, lower, b, x, 1e-10, 100, C.byref(r0), C.byref(r1)) == 0


@pytest.mark.parametrize("block", [0, 7, 64, -1])
def test_cg_norms_are_plain_sums_in_every_probe_order(block):
    """The round-off probes (orc_set_sum_block) change only the association of
    the global sums: the initial residual norm |b|_1 of an integer right-hand
    side is exact in any order, the CG still recovers the solution, and the
    default order is restored afterwards."""
    mesh = O.box(cells_x=12, cells_y=12, grading=0)
    diag, upper, lower, _ = laplacian(mesh, np.zeros(mesh.cells))
    rng = np.random.default_rng(3)
    b = rng.integers(-50, 50, mesh.cells).astype(np.float64)
    x = np.zeros(mesh.cells)
    r0, r1 = C.c_double(), C.c_double()
    O.set_sum_block(block)
    try:
        its = O.lib().orc_ldu_cg(mesh.h, diag, upper, lower, b, x, 1e-12, 500, C.byref(r0), C.byref(r1))
    finally:
        O.set_sum_block(0)
    assert r0.value == np.sum(np.abs(b))
    assert 0 < its < 500
    np.testing.assert_allclose(dense(mesh, diag, upper, lower) @ x, b, rtol=0, atol=1e-8 * np.abs(b).max())


def test_probe_orders_move_a_solve_by_rounding_only():
    """Three laminar outer iterations with pairwise sums and the alternative
    tanh stay within 1e-10 of the default run but not bit-identical (the
    probes are live), and a run after resetting them repeats the default run
    exactly."""
    def run():
        mesh = O.cgrid(0, 0, 12, options=O.cgrid_options("coarse", **SMALL))
        kind, vel, pres = O.face_conditions(mesh, 1.0, 3.0)
        s = O.Solver(mesh, kind, vel, pres)
        s.initialise(*O.uniform_field(mesh.cells, speed=1.0, alpha_deg=3.0, density=1.0, reynolds=500.0))
        s.iterate(3)
        return s.field()[:3]

    base = run()
    O.set_sum_block(-1)
    O.set_tanh_variant(1)
    try:
        probed = run()
    finally:
        O.set_sum_block(0)
        O.set_tanh_variant(0)
    again = run()
    moved = False
    for a, b, c in zip(base, probed, again):
        assert np.array_equal(a, c)
        assert np.abs(b - a).max() <= 1e-10 * np.abs(a).max()
        moved |= not np.array_equal(a, b)
    assert moved


# --- distributions and geometry (Distribution.cpp, Naca4.cpp) ---------------

def test_geometric_distribution():
    """SolverTests.cpp:45-59: first spacing, end points, growing intervals."""
    t = np.zeros(41)
    assert O.lib().orc_geometric_distribution(0.001, 1.0, 40, t) == 41
    assert t[0] == 0.0 and t[-1] == 1.0
    assert abs((t[1] - t[0]) - 0.001) < 1e-6
    d = np.diff(t)
    assert np.all(d > 0) and np.all(d[1:] >= d[:-1] - 1e-15)


def test_naca0012_ordinate_at_thirty_percent():
    """Naca4Tests.cpp:106-116: the tabulated 0.06002 c half-thickness at x = 0.3 c with
    the standard (open trailing edge) polynomial, zero at and ahead of the leading edge.
    The closed-edge a4 (Naca4.cpp:18-23) moves it by 1e-5 c."""
    assert abs(O.lib().orc_thickness(0.30, 0.12, 0) - 0.06002) < 1e-5
    closed = O.lib().orc_thickness(0.30, 0.12, 1)
    assert abs(closed - 5 * 0.12 * (0.2969 * 0.3 ** 0.5 - 0.1260 * 0.3 - 0.3516 * 0.09 + 0.2843 * 0.027
                                     - 0.1036 * 0.0081)) < 1e-15
    assert O.lib().orc_thickness(0.0, 0.12, 1) == 0.0
    assert O.lib().orc_thickness(-0.5, 0.12, 1) == 0.0


# --- box meshes (BoxGrid.cpp:10-66) ------------------------------------------

def test_box_size_area_and_orthogonality():
    """SolverTests.cpp:88-98."""
    mesh = O.box(length=2.0, height=0.5, cells_x=8, cells_y=4, grading=0)
    assert (mesh.nodes_i, mesh.nodes_j, mesh.cells) == (9, 5, 32)
    assert mesh.inverted == 0
    assert abs(mesh.total_area - 1.0) < 1e-12
    assert abs(mesh.max_nonortho_deg) < 1e-9


def test_box_sides_carry_the_requested_boundary_types():
    """SolverTests.cpp:100-118."""
    mesh = O.box(cells_x=6, cells_y=3, grading=0, left=FARFIELD, right=OUTLET, lower=WALL, upper=WALL)
    assert np.sum(mesh.boundary == FARFIELD) == 3
    assert np.sum(mesh.boundary == OUTLET) == 3
    assert np.sum(mesh.boundary == WALL) == 12
    assert np.all(np.abs(mesh.fcentre[mesh.boundary == FARFIELD, 0]) < 1e-12)


# --- structured mesh invariants (MeshTests.cpp, CGridTests.cpp) -------------

@pytest.fixture(scope="module")
def small_cgrid():
    return O.cgrid(2, 4, 12, options=O.cgrid_options("coarse", **SMALL))


def _cell_closure(mesh):
    """Sum over each cell's four faces of the outward area vector, and of
    (x_f . S_f) / 2 (the divergence theorem gives the cell area)."""
    S = mesh.fnormal * mesh.farea[:, None]
    close = np.zeros((mesh.cells, 2))
    div = np.zeros(mesh.cells)
    xs = np.einsum("fi,fi->f", mesh.fcentre, S)
    np.add.at(close, mesh.owner, S)
    np.add.at(div, mesh.owner, 0.5 * xs)
    inner = mesh.neighbour >= 0
    np.add.at(close, mesh.neighbour[inner], -S[inner])
    np.add.at(div, mesh.neighbour[inner], -0.5 * xs[inner])
    return close, div


def test_area_vectors_close_and_areas_obey_the_divergence_theorem(small_cgrid):
    """MeshTests.cpp:195-230, CGridTests.cpp:116-130."""
    close, div = _cell_closure(small_cgrid)
    scale = np.max(small_cgrid.farea)
    assert np.max(np.abs(close)) < 1e-12 * max(scale, 1.0)
    np.testing.assert_allclose(div, small_cgrid.area, rtol=1e-9, atol=1e-15)
    assert small_cgrid.inverted == 0 and np.all(small_cgrid.area > 0)


def test_euler_formula_and_counts(small_cgrid):
    """MeshTests.cpp:46-67: faces = nI (nJ + 1) + (nI + 1) nJ for the structured block."""
    m = small_cgrid
    assert m.faces == m.ni * (m.nj + 1) + (m.ni + 1) * m.nj
    assert m.cells == m.ni * m.nj
    # each cell owns or neighbours exactly four faces
    refs = np.bincount(m.owner, minlength=m.cells) + np.bincount(m.neighbour[m.neighbour >= 0], minlength=m.cells)
    assert np.all(refs == 4)


def test_wake_cut_pairs_are_coincident_and_opposite():
    """CGridTests.cpp:237-263 on the medium preset."""
    m = O.cgrid(2, 4, 12, options=O.cgrid_options("medium"))
    wake = np.nonzero(m.boundary == WAKECUT)[0]
    assert len(wake) > 0
    assert np.all(m.partner[m.boundary != WAKECUT] == -1)
    p = m.partner[wake]
    assert np.all(p >= 0) and np.all(m.partner[p] == wake) and np.all(m.boundary[p] == WAKECUT)
    assert np.max(np.linalg.norm(m.fcentre[wake] - m.fcentre[p], axis=1)) < 1e-12
    assert np.max(np.abs(m.farea[wake] - m.farea[p])) < 1e-12
    assert np.max(np.abs(np.einsum("fi,fi->f", m.fnormal[wake], m.fnormal[p]) + 1.0)) < 1e-12
    assert np.all(m.owner[wake] != m.owner[p])
    assert np.max(np.abs(m.fcentre[wake, 1])) < 1e-12  # on the chord line (CGridTests.cpp:265-280)


@pytest.mark.parametrize("level,dims,wall", [("coarse", (222, 39), 158), ("medium", (430, 71), 318),
                                             ("fine", (830, 127), 638)])
def test_preset_dimensions(level, dims, wall):
    """CGridTests.cpp:369-378 and the preset sizes (CGrid.cpp:73-96): cells 8,658 /
    30,530 / 105,410, wall faces 2 surfacePoints - 2."""
    m = O.cgrid(2, 4, 12, options=O.cgrid_options(level))
    assert (m.ni, m.nj) == dims
    assert np.sum(m.boundary == WALL) == wall


def test_total_area_matches_the_analytic_domain():
    """CGridTests.cpp:282-292: rectangle behind the section + half ellipse in
    front, less the section, to 1e-4 on the fine preset."""
    o = O.cgrid_options("fine")
    m = O.cgrid(2, 4, 12, options=o)
    wall = m.boundary == WALL
    # section area from the wall faces (closed polygon, shoelace via x_f . S_f)
    S = m.fnormal[wall] * m.farea[wall, None]
    foil = -0.5 * np.sum(np.einsum("fi,fi->f", m.fcentre[wall], S))
    xt = 1.0  # closed trailing edge at the chord
    half = o.vertical
    front = xt + o.upstream
    expected = 2 * half * o.downstream + 0.5 * np.pi * front * half - foil
    assert 0.05 < foil < 0.1
    assert abs(m.total_area - expected) < 1e-4 * expected


@pytest.mark.parametrize("h", [1e-3, 3e-4, 1e-4])
def test_first_layer_height_is_honoured(h):
    """CGridTests.cpp:333-344."""
    m = O.cgrid(2, 4, 12, options=O.cgrid_options("medium", first_layer=h))
    assert abs(m.min_wall_spacing - h) < 0.05 * h
