"""Graph replay of the outer iteration: `iterate`, `solve`, batches and the
serving loop queue one captured HIP graph per iteration; `iterate_timed`
queues the same launches one by one (with events between the stages). The
replays must reproduce the direct launches bit for bit, on every replay, with
host reads of the fields and a settings change (which captures again)
between them."""
import numpy as np
import pytest

from helpers import SMALL

pytestmark = pytest.mark.gpu


def _solver(G, precision=0):
    grid = G.cgrid(2, 4, 12, options=G.cgrid_options("coarse", **SMALL))
    kind, vel, pres = G.face_conditions(grid, 1.0, 3.0)
    s = G.SimpleSolver(grid, kind, vel, pres, G.default_settings(precision=precision))
    s.set_turbulence_model(G.MODEL_SST)
    s.initialise(*G.uniform_field(grid.cells, speed=1.0, alpha_deg=3.0, density=1.0, reynolds=1e5))
    return grid, s


@pytest.mark.parametrize("precision", [0, 1])
def test_replayed_iterations_match_direct_launches(G, gpu_available, precision):
    print(f"case: NACA 2412 118x27 SST, precision {precision}, graph replays against direct launches", flush=True)
    (_, replayed), (_, direct) = _solver(G, precision), _solver(G, precision)
    for rnd in range(4):
        mr = replayed.iterate(5)
        md, _, _ = direct.iterate_timed(5)
        assert mr.continuity == md.continuity and mr.momentum_x == md.momentum_x, rnd
        for a, b in zip(replayed.field(), direct.field()):  # eager reads between the replays
            assert np.array_equal(a, b), rnd
    # a settings change alters the launch schedule (fewer CG iterations): captured again
    for s in (replayed, direct):
        st = s.settings
        st.pressure_iterations = 50
        s.set_settings(st)
    for rnd in range(3):
        mr = replayed.iterate(4)
        md, _, _ = direct.iterate_timed(4)
        assert mr.pressure_iterations == md.pressure_iterations <= 50
        for a, b in zip(replayed.field(), direct.field()):
            assert np.array_equal(a, b), rnd
    assert np.array_equal(replayed.turbulence()["omega"], direct.turbulence()["omega"])

