"""The reference's continued polar sweep on the device (cfdgpu_polar_sweep),
against the polar the reference wrote itself
(tests/golden/naca0012_re1e6_sst_polar.csv: NACA 0012, medium C-grid with a
2.5e-5 chord first layer, k-omega SST, Re 1e6, U 50 m/s, Tu 0.1 %,
mu_t/mu 1, tolerance 1e-6, 4,000-iteration cap, alpha 0..18 in 2 degree
steps, continued between points). A continued sweep is path-dependent: a
difference of a few iterations at one point (a round-off difference in a
reduction is enough) moves the next point's start, and near stall that is
amplified. Measured here: the device stops within 2 iterations of the
reference up to 8 deg (0 deg: 531 against 533, so Cd, still moving, differs
by 1.6e-5), within 12 beyond; |dCl| 2e-7 at 0 deg, 8e-5 at 6 deg, 2e-4 at
10 deg, 1.3e-3 at 14 deg, 2.9e-3 at 18 deg; |dCd| and |dCm| at most 1.6e-5
up to 8 deg. The test anchors Cl, Cd, Cm at 2e-4 up to 8 deg, 1e-3 at 10 and
12 deg and 5e-3 beyond, the statuses exactly, the counts within 2 %, and the
separation columns: attached where the reference's is empty, elsewhere
within 1e-3 chord (measured: at most 2.2e-4, at 10 deg)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _reference_rows():
    rows = []
    for line in open(os.path.join(HERE, "golden", "naca0012_re1e6_sst_polar.csv")):
        if line[0].isdigit() or line[0] == "-":
            r = line.strip().split(",")
            sep = tuple(float(v) if v else None for v in (r[7], r[8]))
            rows.append((float(r[0]), float(r[1]), float(r[2]), float(r[5]), r[9], int(r[10]), sep))
    return rows


def test_continued_sst_polar_matches_the_reference_file(G, gpu_available):
    print("case: NACA0012 medium preset (430x71), y1 2.5e-5, SST Re 1e6, alpha 0..18 step 2, continued", flush=True)
    rows = _reference_rows()
    mesh = G.cgrid(0, 0, 12, options=G.cgrid_options("medium", first_layer_height=2.5e-5))
    assert (mesh.ni, mesh.nj) == (430, 71)
    st = G.default_settings(intensity=0.001, viscosity_ratio=1.0)
    pts = G.polar_sweep(mesh, [r[0] for r in rows], settings=st, model=G.MODEL_SST, speed=50.0, density=1.225,
                        reynolds=1e6, tolerance=1e-6, max_iterations=4000, check_every=1)
    for p, (alpha, cl, cd, cm, status, its, sep) in zip(pts, rows):
        print(f"alpha {alpha:5.1f}: cl {p.cl:.8f} ({cl:.8f}) cd {p.cd:.8f} ({cd:.8f}) cm {p.cm:.8f} ({cm:.8f}) "
              f"its {p.iterations} ({its}) status {p.status} separation {p.upper_separation:.6f},"
              f"{p.lower_separation:.6f} ({sep[0]},{sep[1]})", flush=True)
    status_code = {"converged": 0, "iteration-limit": 1, "diverged": 2}
    assert len(pts) == len(rows) == 10
    for p, (alpha, cl, cd, cm, status, its, sep) in zip(pts, rows):
        tol = 2e-4 if alpha <= 8.0 else (1e-3 if alpha <= 12.0 else 5e-3)
        assert p.alpha_deg == alpha
        assert abs(p.cl - cl) < tol and abs(p.cd - cd) < tol and abs(p.cm - cm) < tol, alpha
        assert p.status == status_code[status], alpha
        assert abs(p.iterations - its) <= 0.02 * its, alpha
        for got, ref in zip((p.upper_separation, p.lower_separation), sep):
            if ref is None:
                assert got < 0.0, (alpha, got)  # attached
            else:
                assert abs(got - ref) < 1e-3, (alpha, got, ref)
    assert np.all(np.isfinite([p.cl for p in pts]))
