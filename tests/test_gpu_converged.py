"""Converged aerofoil solves against the FP64 oracle. Each case runs the
reference's driver loop to its stop rule (the worst of the three residuals
below 1e-6, SolverWorker.cpp:166-190) on the device (cfdgpu_solve, the
device-side stop test at every iteration) and on the oracle (one outer
iteration at a time), then extracts the surface and the forces of both
(SurfaceData.cpp:208-348, Forces.cpp:60-130). The two differ only in the
summation order of the pressure solver's reductions, so the iteration counts
agree within two and the converged coefficients within 1e-6 relative; the
surface Cp and Cf within 1e-5 of their largest magnitude; the separation
points within 1e-4 of the chord (a separation found on one side is found on
the other)."""
import numpy as np
import pytest

from helpers import SMALL, cgrid_pair, solver_pair

pytestmark = pytest.mark.gpu

CASES = [
    # (NACA m, p, t), alpha, Reynolds number
    ((0, 0, 12), 4.0, 1e5),
    ((2, 4, 12), 2.0, 1e5),
]


def _converge_oracle(s, tol, cap):
    for it in range(cap):
        mon = s.iterate(1)
        if mon.worst() < tol:
            return it + 1
    return cap


@pytest.mark.parametrize("section,alpha,reynolds", CASES)
def test_converged_sst_forces_match_the_oracle(G, O, gpu_available, section, alpha, reynolds):
    print(f"case: NACA {section} 118x27 SST Re {reynolds:g} alpha {alpha}, solve to 1e-6", flush=True)
    gm, om = cgrid_pair(G, O, *section, **SMALL)
    kind, vel, pres = G.face_conditions(gm, 1.0, alpha)
    field = G.uniform_field(gm.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=reynolds)
    gs, os_ = solver_pair(G, O, gm, om, kind, vel, pres, field, model=G.MODEL_SST)
    tol, cap = 1e-6, 3000
    rep = gs.solve(tol, cap, check_every=1)
    taken = _converge_oracle(os_, tol, cap)
    print(f"iterations: device {rep.iterations} oracle {taken}", flush=True)
    assert rep.status == 0 and taken < cap
    assert abs(rep.iterations - taken) <= 2

    fg, rg = gs.surface(speed=1.0, alpha_deg=alpha, density=1.0, reference_pressure=0.0)
    fo, ro = O.surface(om, *os_.field(), speed=1.0, alpha_deg=alpha, density=1.0, reference_pressure=0.0)
    print(f"cl {fg.cl:.10f} ({fo.cl:.10f}) cd {fg.cd:.10f} ({fo.cd:.10f}) cm {fg.cm:.10f} ({fo.cm:.10f})",
          flush=True)
    for name in ("cl", "cd", "cm", "cd_pressure", "cd_friction"):
        a, b = getattr(fg, name), getattr(fo, name)
        assert abs(a - b) <= 1e-6 * abs(b), name
    assert rg.shape == ro.shape
    # side, face, x, y, x/c, arc and the station lengths are mesh-only
    for col in (0, 1, 10):
        assert np.array_equal(rg[:, col], ro[:, col]), col
    for col in (2, 3, 4, 5, 11):
        assert np.allclose(rg[:, col], ro[:, col], rtol=0, atol=1e-14), col
    for col in (7, 9):  # Cp, Cf
        scale = np.max(np.abs(ro[:, col]))
        assert np.max(np.abs(rg[:, col] - ro[:, col])) <= 1e-5 * scale, col
    assert (fg.n_upper, fg.n_lower, fg.stations) == (fo.n_upper, fo.n_lower, fo.stations)
    for name in ("upper_separation", "lower_separation", "stagnation_xc"):
        a, b = getattr(fg, name), getattr(fo, name)
        assert (a < 0.0) == (b < 0.0) and abs(a - b) <= 1e-4, name
    # the same separation wall face (the first shear sign change, + to -, away from
    # stagnation, SurfaceData.cpp:179-204, found in the oracle's station rows)
    for side, name in ((1, "upper_separation_face"), (0, "lower_separation_face")):
        st = ro[ro[:, 0] == side]
        face = -1
        for q in range(1, len(st)):
            if st[q, 10] or st[q - 1, 10]:
                continue
            if st[q - 1, 8] > 0.0 and st[q, 8] < 0.0:
                face = int(st[q, 1])
                break
        assert getattr(fg, name) == face, name
