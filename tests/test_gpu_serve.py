"""The serving loop (cfdgpu_server_*): requests admitted as they arrive, by
slot count and memory budget, advanced together on the device and retired
between chunks. Each result must be the same whether the request was served
in a mix or alone: here bit-identical to the drop-in solver's own device solve
from the same starting field (SimpleSolver.solve, the worker's stop rule of
SolverWorker.cpp:161-192). Warm starts take the nearest converged alpha
(Panels.cpp:1137-1157) and stay within |dCl| < 1e-3 of the cold solve; a
divergence is retried once from cold and reported (Polar.cpp:79-84, 112-117)
without disturbing the other resident solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# three small C-grids (rows 27, 35, 71) that converge to 1e-5 in 150-250 iterations
MESHES = [(0, 0, 12, dict(surface_points=40, wake_points=20, normal_points=28)),
          (2, 4, 12, dict(surface_points=56, wake_points=24, normal_points=36)),
          (4, 4, 15, dict(surface_points=48, wake_points=16, normal_points=72, first_layer_height=1e-4))]
TOL, CAP, RE = 1e-5, 2000, 1e5


def _settings(G, **kw):
    return G.default_settings(viscosity_ratio=10.0, **kw)


def _request(G, spec, alpha, arrival=0.0, settings=None):
    m, p, t, kw = spec
    return G.serve_request(m, p, t, alpha, RE, G.cgrid_options("coarse", **kw), arrival_s=arrival,
                           settings=settings)


def _server(G, **kw):
    base = dict(tolerance=TOL, max_iterations=CAP, settings=_settings(G), warm_start=0, keep_fields=1,
                chunk_iterations=7)
    base.update(kw)
    return G.Server(G.server_options(**base))


def _solo(G, spec, alpha, start=None, settings=None):
    """The same request through the drop-in solver: returns (report, fields, forces)."""
    m, p, t, kw = spec
    mesh = G.cgrid(m, p, t, options=G.cgrid_options("coarse", **kw))
    kind, vel, pres = G.face_conditions(mesh, 1.0, alpha)
    u, v, pr, rho, mu = G.uniform_field(mesh.cells, speed=1.0, alpha_deg=alpha, density=1.0, reynolds=RE)
    if start is not None:
        u, v, pr = (np.array(a) for a in start)
    s = G.SimpleSolver(mesh, kind, vel, pres, settings or _settings(G))
    s.set_turbulence_model(G.MODEL_SST)
    s.initialise(u, v, pr, rho, mu)
    rep = s.solve(TOL, CAP, check_every=1)
    fields = s.field()
    forces, _ = G.surface(mesh, *fields, speed=1.0, alpha_deg=alpha, density=1.0)
    return rep, fields, forces


def _run_all(srv):
    st = srv.run(0.0)
    return {r.ticket: r for r in srv.collect()}, st


def test_each_result_is_the_same_alone_and_in_the_mix(G, gpu_available):
    print("case: 5 SST requests on 3 C-grids, 3 slots, 3 joining after 3 chunks, chunk 7", flush=True)
    jobs = [(MESHES[0], 2.0), (MESHES[1], 4.0), (MESHES[2], -2.0), (MESHES[0], 5.0), (MESHES[1], 0.0)]
    srv = _server(G, max_slots=3)
    tickets = [srv.submit(_request(G, spec, a)) for spec, a in jobs[:2]]
    for _ in range(3):
        st = srv.step()
    assert st.active == 2 and st.finished == 0 and st.member_iterations > 0
    # arrivals while the first two are resident: one joins at once, two wait for a slot
    tickets += [srv.submit(_request(G, spec, a, arrival=st.clock_s)) for spec, a in jobs[2:]]
    res, st = _run_all(srv)
    assert st.finished == len(jobs) and st.queued == 0 and st.active == 0
    assert st.peak_active == 3
    for ticket, (spec, alpha) in zip(tickets, jobs):
        r = res[ticket]
        assert r.status == 0 and r.warm_from == -1 and r.retried_cold == 0
        assert r.admitted_s >= r.arrival_s and r.finished_s > r.admitted_s
        rep, fields, forces = _solo(G, spec, alpha)
        assert (rep.status, rep.iterations) == (r.status, r.iterations)
        for a, b in zip(srv.fields(ticket, r.cells), fields):
            assert np.array_equal(a, b), f"ticket {ticket}"
        assert (r.cl, r.cd, r.cm) == (forces.cl, forces.cd, forces.cm)


def test_warm_start_takes_the_nearest_converged_alpha(G, gpu_available):
    print("case: warm starts on the 118x27 C-grid, alpha 2 -> 3 -> 3.4", flush=True)
    spec = MESHES[0]
    srv = _server(G, warm_start=1)
    first = srv.submit(_request(G, spec, 2.0))
    done, _ = _run_all(srv)
    second = srv.submit(_request(G, spec, 3.0))
    done.update(_run_all(srv)[0])
    third = srv.submit(_request(G, spec, 3.4))
    done.update(_run_all(srv)[0])
    assert done[first].warm_from == -1
    assert done[second].warm_from == first
    assert done[third].warm_from == second  # 3.0 is nearer to 3.4 than 2.0
    for ticket, donor, alpha in ((second, first, 3.0), (third, second, 3.4)):
        r = done[ticket]
        assert r.status == 0
        start = srv.fields(donor, done[donor].cells)[:3]
        rep, fields, _ = _solo(G, spec, alpha, start=start)
        assert rep.iterations == r.iterations
        for a, b in zip(srv.fields(ticket, r.cells), fields):
            assert np.array_equal(a, b)
        _, _, cold = _solo(G, spec, alpha)
        assert abs(r.cl - cold.cl) < 1e-3, (r.cl, cold.cl)
        assert abs(r.cd - cold.cd) < 1e-4, (r.cd, cold.cd)


def test_divergence_is_retried_once_and_leaves_the_others_alone(G, gpu_available):
    print("case: 3 requests, the middle one with relaxation 1.0 (diverges)", flush=True)
    bad = _settings(G, velocity_relaxation=1.0, pressure_relaxation=1.0)
    jobs = [(MESHES[0], 2.0, None), (MESHES[1], 4.0, bad), (MESHES[2], -2.0, None)]
    srv = _server(G, max_slots=3)
    tickets = [srv.submit(_request(G, spec, a, settings=s)) for spec, a, s in jobs]
    res, st = _run_all(srv)
    r = res[tickets[1]]
    assert r.status == 2 and r.retried_cold == 1 and not np.isfinite(r.continuity)
    for ticket, (spec, alpha, _) in zip((tickets[0], tickets[2]), (jobs[0], jobs[2])):
        rep, fields, _ = _solo(G, spec, alpha)
        assert res[ticket].status == 0 and res[ticket].iterations == rep.iterations
        for a, b in zip(srv.fields(ticket, res[ticket].cells), fields):
            assert np.array_equal(a, b)


def test_admission_waits_for_memory_and_rejects_what_never_fits(G, gpu_available):
    print("case: budget of 1.5 solves of the 118x27 grid, 3 requests + one 126x71", flush=True)
    probe = _server(G)
    probe.submit(_request(G, MESHES[0], 2.0))
    one = _run_all(probe)[0][0].device_bytes
    srv = _server(G, max_slots=4, hbm_budget_bytes=1.5 * one)
    small = [srv.submit(_request(G, MESHES[0], a)) for a in (0.0, 1.0, 2.0)]
    big = srv.submit(_request(G, MESHES[2], 0.0))
    res, st = _run_all(srv)
    assert st.peak_active == 1 and st.peak_resident_bytes <= 1.5 * one
    assert res[big].status == 3
    order = sorted(small, key=lambda t: res[t].admitted_s)
    assert order == small  # arrival order
    for a, b in zip(order, order[1:]):
        assert res[b].admitted_s >= res[a].finished_s
        assert res[a].status == 0
