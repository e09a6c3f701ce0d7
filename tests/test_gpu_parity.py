"""HIP path vs the FP64 oracle, stage by stage and iteration by iteration.

The plain FP64 path sums every cell's face contributions in the reference's
order with FP contraction off, and sweeps Gauss-Seidel in the reference's
lexicographic order, so everything upstream of the first global reduction is
bit-identical; the CG's dot products and L1 norms are tree sums, which moves
the pressure correction by round-off (bounded here at 1e-10 relative per
iteration on the small C-grid).
"""
import numpy as np
import pytest

from helpers import SMALL, cgrid_pair, ldu_to_rows, rel_err, solver_pair

pytestmark = pytest.mark.gpu


def _small_laminar(G, O, alpha=4.0, **mesh_kw):
    gm, om = cgrid_pair(G, O, **(mesh_kw or SMALL))
    kind, vel, pres = G.face_conditions(gm, 1.0, alpha)
    field = G.uniform_field(gm.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=500.0)
    gs, os_ = solver_pair(G, O, gm, om, kind, vel, pres, field)
    return gm, om, gs, os_


def test_first_iteration_stages_bit_identical(G, O, gpu_available):
    gm, om, gs, os_ = _small_laminar(G, O)
    print("case: NACA0012 118x27 laminar Re=500 alpha=4, one iteration", flush=True)
    mg = gs.iterate()
    mo = os_.iterate()
    # gradients of the start of the iteration (Green-Gauss)
    for name in ("gradU", "gradV", "gradP"):
        assert np.array_equal(gs.debug_array(name), os_.array(name)), name
    # momentum system, after relaxation (the oracle keeps the last component, y)
    aS, aN, aW, aE = ldu_to_rows(om, os_.array("momUpper"), os_.array("momLower"))
    assert np.array_equal(gs.debug_array("momDiag"), os_.array("momDiag"))
    assert np.array_equal(gs.debug_array("momS"), aS)
    assert np.array_equal(gs.debug_array("momN"), aN)
    assert np.array_equal(gs.debug_array("momW"), aW)
    assert np.array_equal(gs.debug_array("momE"), aE)
    assert np.array_equal(gs.debug_array("momBy"), os_.array("momSource"))
    # pressure system from the Rhie-Chow fluxes of the swept velocities
    pS, pN, pW, pE = ldu_to_rows(om, os_.array("pUpper"), os_.array("pLower"))
    assert np.array_equal(gs.debug_array("pDiag"), os_.array("pDiag"))
    assert np.array_equal(gs.debug_array("pN"), pN)
    assert np.array_equal(gs.debug_array("pE"), pE)
    assert np.array_equal(gs.debug_array("pSource"), os_.array("pSource"))
    # monitor: continuity is an L1 sum (re-associated), CG count must match
    assert mg.pressure_iterations == mo.pressure_iterations
    assert abs(mg.continuity - mo.continuity) <= 1e-12 * abs(mo.continuity)
    assert abs(mg.momentum_x - mo.momentum_x) <= 1e-12 * abs(mo.momentum_x)
    assert abs(mg.momentum_y - mo.momentum_y) <= 1e-12 * abs(mo.momentum_y)
    # pressure correction from CG: round-off only
    assert rel_err(gs.debug_array("correction"), os_.array("correction")) < 1e-10


def test_ten_iterations_track_the_oracle(G, O, gpu_available):
    gm, om, gs, os_ = _small_laminar(G, O, alpha=2.0)
    print("case: NACA0012 118x27 laminar Re=500 alpha=2, ten iterations", flush=True)
    for it in range(10):
        mg = gs.iterate()
        mo = os_.iterate()
        assert mg.pressure_iterations == mo.pressure_iterations, it
        assert abs(mg.continuity - mo.continuity) <= 1e-8 * abs(mo.continuity), it
    ug, vg, pg, _ = gs.field()
    uo, vo, po, _ = os_.field()
    assert rel_err(ug, uo) < 1e-9
    assert rel_err(vg, vo) < 1e-9
    assert rel_err(pg, po) < 1e-9
    assert rel_err(gs.mass_flux(), os_.flux()) < 1e-9
    assert rel_err(gs.divergence(), os_.divergence()) < 1e-7


def test_sst_iterations_track_the_oracle(G, O, gpu_available):
    gm, om = cgrid_pair(G, O, **SMALL)
    kind, vel, pres = G.face_conditions(gm, 1.0, 4.0)
    field = G.uniform_field(gm.cells, speed=1.0, alpha_deg=4.0, density=1.0, reynolds=1e5)
    gs, os_ = solver_pair(G, O, gm, om, kind, vel, pres, field, model=G.MODEL_SST)
    print("case: NACA0012 118x27 SST Re=1e5 alpha=4, five iterations", flush=True)
    tg = gs.turbulence()
    to = os_.turbulence()
    assert np.array_equal(tg["wall_distance"], to["wall_distance"])
    assert np.array_equal(tg["k"], to["k"])
    assert np.array_equal(tg["omega"], to["omega"])
    assert np.array_equal(tg["mut"], to["mut"])
    for it in range(5):
        mg = gs.iterate()
        mo = os_.iterate()
        assert abs(mg.continuity - mo.continuity) <= 1e-8 * abs(mo.continuity), it
        assert abs(mg.turb_first - mo.turb_first) <= 1e-8 * abs(mo.turb_first), it
        assert abs(mg.turb_second - mo.turb_second) <= 1e-8 * abs(mo.turb_second), it
    tg = gs.turbulence()
    to = os_.turbulence()
    for name in ("k", "omega", "mut"):
        assert rel_err(tg[name], to[name]) < 1e-8, name
    ug, vg, pg, mug = gs.field()
    uo, vo, po, muo = os_.field()
    assert rel_err(ug, uo) < 1e-9
    assert rel_err(pg, po) < 1e-9
    assert rel_err(mug, muo) < 1e-8


def test_non_orthogonal_correctors_track_the_oracle(G, O, gpu_available):
    """SimpleSolver.cpp:716-748 with nonOrthogonalCorrectors = 2: two pressure
    assemblies, solves and corrections per outer iteration."""
    gm, om = cgrid_pair(G, O, **SMALL)
    kind, vel, pres = G.face_conditions(gm, 1.0, 3.0)
    field = G.uniform_field(gm.cells, speed=1.0, alpha_deg=3.0, density=1.0, reynolds=500.0)
    gset = G.default_settings(non_orthogonal_correctors=2)
    oset = O.default_settings()
    oset.non_orthogonal_correctors = 2
    gs, os_ = solver_pair(G, O, gm, om, kind, vel, pres, field, gsettings=gset, osettings=oset)
    print("case: NACA0012 118x27 laminar Re=500 alpha=3, two correctors, five iterations", flush=True)
    for it in range(5):
        mg = gs.iterate()
        mo = os_.iterate()
        assert mg.pressure_iterations == mo.pressure_iterations, it
        assert abs(mg.continuity - mo.continuity) <= 1e-8 * abs(mo.continuity), it
    for a, b in zip(gs.field()[:3], os_.field()[:3]):
        assert rel_err(a, b) < 1e-9
    assert rel_err(gs.mass_flux(), os_.flux()) < 1e-9
