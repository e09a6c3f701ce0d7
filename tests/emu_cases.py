"""Cases run on the CPU emulation of the HIP library (tests/emu, built with
AddressSanitizer) by test_emu_asan.py, in a child process that preloads the
sanitizer runtime. Each prints one line 'ok <case>' or raises."""
import sys

import numpy as np

import cfdgpu as G
import oracle_py as O
from helpers import SMALL, cgrid_pair, solver_pair


def parity(normal_points):
    """One laminar outer iteration (CG capped at 3) against the oracle: the
    lexicographic sweep spans normal_points - 1 rows (two waves of the
    256-lane sweep at 99 rows, five of the 1024-lane sweep at 299)."""
    kw = dict(SMALL, normal_points=normal_points)
    gm, om = cgrid_pair(G, O, 0, 0, 12, **kw)
    kind, vel, pres = G.face_conditions(gm, 1.0, 2.0)
    field = G.uniform_field(gm.cells, speed=1.0, alpha_deg=2.0, density=1.0, reynolds=500.0)
    gs_set = G.default_settings(pressure_iterations=3)
    os_set = O.default_settings()
    os_set.pressure_iterations = 3
    gs, os_ = solver_pair(G, O, gm, om, kind, vel, pres, field, gsettings=gs_set, osettings=os_set)
    gs.iterate(1)
    os_.iterate(1)
    for a, b in zip(gs.field()[:3], os_.field()[:3]):
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
        assert err < 1e-12, err
    print(f"ok parity {gm.ni}x{gm.nj}", flush=True)


def batch(fast):
    """A batch of two meshes of different widths (27 and 99 rows) against the
    narrower one alone: bit-identical."""
    st = dict(precision=1, gs_ordering=1, pressure_solver=1, pressure_iterations=4, refinement_passes=2) if fast \
        else dict(pressure_iterations=4)

    def make(npts, m, p, t, a):
        g = G.cgrid(m, p, t, options=G.cgrid_options("coarse", **dict(SMALL, normal_points=npts)))
        kind, vel, pres = G.face_conditions(g, 1.0, a)
        s = G.SimpleSolver(g, kind, vel, pres, G.default_settings(**st))
        s.set_turbulence_model(G.MODEL_SST)
        s.initialise(*G.uniform_field(g.cells, speed=1.0, alpha_deg=a, density=1.0, reynolds=1e5))
        return g, s

    alone = make(28, 0, 0, 12, 4.0)
    pair = [make(28, 0, 0, 12, 4.0), make(100, 2, 4, 12, 1.0)]
    alone[1].iterate(2)
    G.Batch([s for _, s in pair]).iterate(2)
    for a, b in zip(alone[1].field(), pair[0][1].field()):
        assert np.array_equal(a, b)
    assert np.isfinite(pair[1][1].field()[0]).all()
    print(f"ok batch {'fast' if fast else 'plain'}", flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:]:
        if name.startswith("parity"):
            parity(int(name[len("parity"):]))
        elif name == "batch_plain":
            batch(False)
        elif name == "batch_fast":
            batch(True)
        else:
            raise SystemExit(f"unknown case {name}")
