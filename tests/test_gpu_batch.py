"""Batched solves (cfdgpu_batch_*): several solvers advanced by one launch per
stage. Each member's result must be bit-identical to the same solve run on
its own: the per-instance arithmetic, reduction partials and sweep order do
not depend on the batch (only the launch width does)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# three C-grids of different size (rows 27, 35, 71: the widest needs 128 lanes)
MESHES = [(0, 0, 12, dict(surface_points=40, wake_points=20, normal_points=28), 2.0),
          (2, 4, 12, dict(surface_points=56, wake_points=24, normal_points=36), 4.0),
          (4, 4, 15, dict(surface_points=48, wake_points=16, normal_points=72, first_layer_height=1e-4), -2.0)]


def _make(G, spec, model):
    m, p, t, kw, alpha = spec
    mesh = G.cgrid(m, p, t, options=G.cgrid_options("coarse", **kw))
    kind, vel, pres = G.face_conditions(mesh, 1.0, alpha)
    field = G.uniform_field(mesh.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=1e5)
    s = G.SimpleSolver(mesh, kind, vel, pres, G.default_settings(viscosity_ratio=10.0))
    if model:
        s.set_turbulence_model(model)
    s.initialise(*field)
    return mesh, s


@pytest.mark.parametrize("model", [0, 2], ids=["laminar", "sst"])
def test_batch_members_match_solo_runs(G, gpu_available, model):
    print(f"case: batch of 3 C-grids (118x27, 158x35, 126x71), model={model}, 6 iterations", flush=True)
    solo = [_make(G, spec, model) for spec in MESHES]
    for _, s in solo:
        s.iterate(6)
    members = [_make(G, spec, model) for spec in MESHES]
    batch = G.Batch([s for _, s in members])
    mons = batch.iterate(6)
    for (mesh, a), (_, b), mon in zip(solo, members, mons):
        for fa, fb in zip(a.field(), b.field()):
            assert np.array_equal(fa, fb)
        assert np.array_equal(a.mass_flux(), b.mass_flux())
        if model == 2:
            ta, tb = a.turbulence(), b.turbulence()
            for key in ("k", "omega", "mut"):
                assert np.array_equal(ta[key], tb[key]), key
        assert np.isfinite(mon.continuity)


def test_batch_solve_stops_each_member_by_its_own_rule(G, gpu_available):
    print("case: batch solve of 3 laminar C-grids, tol 1e-4, cap 400", flush=True)
    solo = [_make(G, spec, 0) for spec in MESHES]
    solo_reports = [s.solve(1e-4, 400, check_every=1) for _, s in solo]
    members = [_make(G, spec, 0) for spec in MESHES]
    batch = G.Batch([s for _, s in members])
    reports = batch.solve(1e-4, 400, check_every=10)
    for (_, a), (_, b), ra, rb in zip(solo, members, solo_reports, reports):
        assert ra.iterations == rb.iterations and ra.status == rb.status
        for fa, fb in zip(a.field(), b.field()):
            assert np.array_equal(fa, fb)
    # members that stopped early were left untouched while the others ran on
    assert len({r.iterations for r in reports}) >= 1
