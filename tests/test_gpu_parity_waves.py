"""HIP path vs the FP64 oracle on meshes whose Gauss-Seidel sweep spans several
wavefronts: 99 rows (two waves of the 256-thread sweep), 300 rows (five waves
of the 1024-thread sweep, read-ahead 3) and the 1024 x 256 NACA 2412 workload
with k-omega SST at Re 1e6 (four waves).

The sweep hands each wave's last row to the next wave through LDS, so the
momentum sweep's result (seen through the pressure system built from the
swept velocities) must stay bit-identical to the oracle's lexicographic
sweep on every mesh. The pressure correction carries the CG's tree-sum
round-off, and the turbulence update after it too.
"""
import numpy as np
import pytest

from helpers import cgrid_pair, ldu_to_rows, rel_err, solver_pair

pytestmark = pytest.mark.gpu

# (id, NACA m p t, mesh options, model, Reynolds, alpha)
CASES = [
    ("naca0012_118x99_laminar", (0, 0, 12), dict(surface_points=40, wake_points=20, normal_points=100), 0, 500.0, 4.0),
    ("naca0012_118x300_laminar", (0, 0, 12), dict(surface_points=40, wake_points=20, normal_points=301), 0, 500.0, 4.0),
    ("naca2412_1024x256_sst", (2, 4, 12),
     dict(surface_points=385, wake_points=128, normal_points=257, first_layer=1e-6, upstream=15.0, vertical=15.0,
          downstream=20.0), 2, 1e6, 4.0),
]


@pytest.mark.parametrize("name,section,mesh_kw,model,reynolds,alpha", CASES, ids=[c[0] for c in CASES])
def test_first_iteration_on_multi_wave_sweeps(G, O, gpu_available, name, section, mesh_kw, model, reynolds, alpha):
    gm, om = cgrid_pair(G, O, *section, **mesh_kw)
    print(f"case: {name} ({gm.ni} x {gm.nj} cells), one iteration", flush=True)
    kind, vel, pres = G.face_conditions(gm, 1.0, alpha)
    field = G.uniform_field(gm.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=reynolds)
    gset = G.default_settings(intensity=0.001, viscosity_ratio=10.0)
    oset = O.default_settings()
    oset.intensity, oset.viscosity_ratio = 0.001, 10.0
    gs, os_ = solver_pair(G, O, gm, om, kind, vel, pres, field, gset, oset, model=model)
    mg = gs.iterate()
    mo = os_.iterate()
    for g in ("gradU", "gradV", "gradP"):
        assert np.array_equal(gs.debug_array(g), os_.array(g)), g
    aS, aN, aW, aE = ldu_to_rows(om, os_.array("momUpper"), os_.array("momLower"))
    assert np.array_equal(gs.debug_array("momDiag"), os_.array("momDiag"))
    for g, ref in (("momS", aS), ("momN", aN), ("momW", aW), ("momE", aE)):
        assert np.array_equal(gs.debug_array(g), ref), g
    assert np.array_equal(gs.debug_array("momBy"), os_.array("momSource"))
    # built from the swept velocities: bit-identical only if every wave swept in the reference's order
    pS, pN, pW, pE = ldu_to_rows(om, os_.array("pUpper"), os_.array("pLower"))
    assert np.array_equal(gs.debug_array("pDiag"), os_.array("pDiag"))
    assert np.array_equal(gs.debug_array("pN"), pN)
    assert np.array_equal(gs.debug_array("pE"), pE)
    assert np.array_equal(gs.debug_array("pSource"), os_.array("pSource"))
    assert mg.pressure_iterations == mo.pressure_iterations
    assert abs(mg.momentum_x - mo.momentum_x) <= 1e-12 * abs(mo.momentum_x)
    assert abs(mg.momentum_y - mo.momentum_y) <= 1e-12 * abs(mo.momentum_y)
    ug, vg, pg, _ = gs.field()
    uo, vo, po, _ = os_.field()
    errs = [rel_err(gs.debug_array("correction"), os_.array("correction")), rel_err(ug, uo), rel_err(vg, vo),
            rel_err(pg, po)]
    print("correction, u, v, p relative errors:", errs, flush=True)
    assert max(errs) < 1e-10  # measured: 2e-14 on the laminar meshes, 4.5e-12 on the workload (200 CG iterations)
    if model:
        assert abs(mg.turb_first - mo.turb_first) <= 1e-8 * abs(mo.turb_first)
        assert abs(mg.turb_second - mo.turb_second) <= 1e-8 * abs(mo.turb_second)
        tg, to = gs.turbulence(), os_.turbulence()
        terrs = [rel_err(tg[g], to[g]) for g in ("k", "omega", "mut")]
        print("k, omega, mut relative errors:", terrs, flush=True)
        assert max(terrs) < 1e-9  # measured: 3.4e-11 (k) on the workload
