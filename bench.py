"""Benchmark of the MI355X SIMPLE outer iteration.

Default run (the benchmark line): the metric, converged solves/s of the
batched fast path against the plain FP64 path at |dCl| < 1e-3 and
|dCd| < 1e-4, with p50/p99 time to convergence, on the committed
64-instance mixed-size batch (workloads/mixed_batch_64_seed0.json: random
NACA four-digit sections, alpha -5..12 deg, Re 2e5..5e6, C-grids of 256x64,
512x128 and 1024x256 cells, k-omega SST, Tu 0.1 %, mu_t/mu 10). All 64
solves arrive at once and are served with continuous batching (64 slots,
chunks of 20 outer iterations), first on the fast path (mixed precision,
two-colour FP32 defect sweeps, line-preconditioned CG), then on the plain
FP64 path (the reference's algorithm). Both run the worker's stop rule
(residuals.worst() < 1e-6, 10,000-iteration limit, SolverWorker.cpp:161-192)
with coefficient settling (Cl, Cd sampled every 500 iterations; estimated
remaining change at most half the acceptance bar). `value` counts the fast
path's converged solves whose Cl and Cd are within the bar of the plain
path's answer for the same instance, per second of the fast run's wall time.

Beside it the line carries the plain path batched and one solve at a time
(each level's measured ms per outer iteration alone times every solve's
iterations), the speedups, p50/p99 latency, the roofline of the dominant unit
(one line-CG iteration over the whole batch, from HIP events on the batch's
stream inside the timed region: K outer iterations after W untimed), the
per-stage times, and the CPU baseline (the FP64 oracle, a restatement of the
reference solver, timed on this host's cores). The line CG runs with the
launch shapes of cfd-simulator_amd/workloads/launch_tuning.json (written by
cfd-simulator_amd/tune.py) when that file exists.

Other modes: --roofline-only (the timed region and its roofline alone, the
command kernel traces are taken of), --single (one 1024x256 SST solve:
per-step times, roofline, mixed comparison), --solve, --converge,
--batch-polar, --mixed-batch, --serve, --calibrate-copy.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(_HERE, "cfd-simulator_amd"))

METRIC = "converged solves/s batched-mixed vs plain FP64 at |ΔCl|<1e-3; p50/p99 latency"
UNIT = "converged solves/s"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published

# HBM bytes per launch of the dominant kernels from rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE, scaled by the same counters on a copy of known size (FETCH_SIZE
# x 2.00, WRITE_SIZE x 1.00): the plain path's single 1024x256 solve
# (profiles/naca2412_1024x256_sst_traffic.md) and the fast path's line-CG
# iteration over the 64-instance batch with every member iterating (launch A
# 299.3 MB + launch B 194.8 MB; profiles/mixed_batch_64_fast_kernels.md).
TRAFFIC_BYTES = {"k_gs_wave_mom": 142.32e6, "k_gs_wave_turb": 135.59e6, "k_cg_a + k_cg_b": 32.08e6,
                 "line_cg_iteration_mixed": 494.1e6}

# HBM bytes per launch of the fast path's FP64 stencil kernels over the whole
# 64-instance batch, counted the same way (profiles/mixed_batch_64_fast_kernels.md),
# keyed by the stage that launches each once per outer iteration. These kernels
# read about their compulsory bytes, so the counted bytes stand for the
# algorithmic ones.
BATCH_KERNEL_BYTES = {"gradients": ("k_grad", 1169.6e6), "momentum_assembly": ("k_mom", 1678.0e6),
                      "rhie_chow": ("k_rhiechow", 1355.8e6), "pressure_assembly": ("k_pressure", 1104.7e6),
                      "sst_blending": ("k_sst_blend", 1737.7e6), "sst_assembly": ("k_sst_asm", 1483.2e6)}

WORKLOAD = dict(section=(2, 4, 12), surface_points=385, wake_points=128, normal_points=257,
                first_layer_height=1e-6, upstream=15.0, vertical=15.0, downstream=20.0,
                reynolds=1e6, alpha=4.0, speed=50.0, density=1.225, intensity=0.001, viscosity_ratio=10.0)


def traffic_model():
    """Compulsory HBM words per cell for each unit of work (each array read or
    written once per kernel, perfect neighbour reuse, geometry per cell)."""
    return {
        "cg_iteration": 16,
        "gradients": 15,
        "momentum_assembly": 36,
        "momentum_sweeps": 33,      # 3 sweeps, u and v: 5 coefficients + 2 sources + 2 reads + 2 writes
        "sst_sweeps": 48,           # 3 sweeps each of k and omega: 5 coefficients + 1 source + 1 read + 1 write
        "rhie_chow_pressure": 34,
        "correction": 23,
        "sst_update": 130,
        # bytes per cell of one line-CG iteration on the mixed path (FP32 storage throughout,
        # the correction x included): launch A reads z, p_old, aP, aN, aE and x (4 B each) and
        # writes x, p, Ap = 36 (x += alpha p of the previous iteration rides in launch A);
        # launch B (rows in registers) reads Ap, r, 1/pivot and aN (the back substitution's
        # factor aN / pivot formed from the two) and writes r, z = 24
        "line_cg_iteration_mixed": 60,
    }


# The kernels that can dominate an outer iteration: (stage, kernel, unit of
# the traffic model, launches per stage call and CG iteration).
DOMINANT_CANDIDATES = [("momentum_sweeps", "k_gs_wave_mom", "momentum_sweeps"),
                       ("sst_sweeps", "k_gs_wave_turb", "sst_sweeps"),
                       ("cg_iterations", "k_cg_a + k_cg_b", "cg_iteration")]


def build_case(G):
    w = WORKLOAD
    m, p, t = w["section"]
    opts = G.cgrid_options("coarse", upstream_chords=w["upstream"], vertical_chords=w["vertical"],
                           downstream_chords=w["downstream"], surface_points=w["surface_points"],
                           wake_points=w["wake_points"], normal_points=w["normal_points"],
                           first_layer_height=w["first_layer_height"])
    mesh = G.cgrid(m, p, t, options=opts)
    kind, vel, pres = G.face_conditions(mesh, w["speed"], w["alpha"])
    field = G.uniform_field(mesh.cells, speed=w["speed"], alpha_deg=w["alpha"], density=w["density"],
                            reynolds=w["reynolds"])
    return mesh, kind, vel, pres, field


def precision_of(G, args):
    return G.MIXED if getattr(args, "precision", "fp64") == "mixed" else G.FP64


def settings_summary(st):
    return {"precision": ["fp64", "mixed"][st.precision], "gs_ordering": ["lexicographic", "two-colour"][st.gs_ordering],
            "pressure_solver": ["jacobi-cg", "line-cg"][st.pressure_solver], "cg_cap": st.pressure_iterations,
            "refinement_passes": st.refinement_passes, "momentum_sweeps": st.momentum_sweeps}


def path_settings(G, args, **kw):
    """SimpleSettings of the run: the reference defaults with the workload's
    inflow turbulence, and the path options of the command line."""
    st = G.default_settings(intensity=0.001, viscosity_ratio=10.0, precision=precision_of(G, args), **kw)
    if getattr(args, "fast", False):
        st.precision, st.gs_ordering, st.pressure_solver, st.pressure_iterations = G.MIXED, 1, 1, 15
    if getattr(args, "gs", None) is not None:
        st.gs_ordering = {"lex": 0, "rb": 1}[args.gs]
    if getattr(args, "psolver", None) is not None:
        st.pressure_solver = {"jacobi": 0, "line": 1}[args.psolver]
    if getattr(args, "cg_max", None) is not None:
        st.pressure_iterations = args.cg_max
    if getattr(args, "refine", None) is not None:
        st.refinement_passes = args.refine
    return st


def gpu_solver(G, mesh, kind, vel, pres, field, precision=0):
    w = WORKLOAD
    st = G.default_settings(intensity=w["intensity"], viscosity_ratio=w["viscosity_ratio"], precision=precision)
    s = G.SimpleSolver(mesh, kind, vel, pres, st)
    s.set_turbulence_model(G.MODEL_SST)
    s.initialise(*field)
    return s


def coefficients(G, mesh, solver):
    w = WORKLOAD
    u, v, p, mu = solver.field()
    f, _ = G.surface(mesh, u, v, p, mu, speed=w["speed"], alpha_deg=w["alpha"], density=w["density"])
    return f


def cpu_baseline(iterations: int, iters_per_solve):
    """The oracle on the same solve, one thread, a bounded number of outer iterations."""
    sys.path.insert(0, os.path.join(_HERE, "oracle"))
    import oracle_py as O  # the checker: timed here as the CPU baseline only

    w = WORKLOAD
    m, p, t = w["section"]
    opts = O.cgrid_options("coarse", upstream=w["upstream"], vertical=w["vertical"], downstream=w["downstream"],
                           surface_points=w["surface_points"], wake_points=w["wake_points"],
                           normal_points=w["normal_points"], first_layer=w["first_layer_height"])
    mesh = O.cgrid(m, p, t, options=opts)
    kind, vel, pres = O.face_conditions(mesh, w["speed"], w["alpha"])
    st = O.default_settings()
    st.intensity = w["intensity"]
    st.viscosity_ratio = w["viscosity_ratio"]
    s = O.Solver(mesh, kind, vel, pres, st)
    s.set_model(2)
    s.initialise(*O.uniform_field(mesh.cells, speed=w["speed"], alpha_deg=w["alpha"], density=w["density"],
                                  reynolds=w["reynolds"]))
    s.iterate(1)  # first iteration outside the sample (page faults, first touch)
    t0 = time.perf_counter()
    s.iterate(iterations)
    sec = (time.perf_counter() - t0) / iterations
    value = (1.0 / (sec * iters_per_solve)) if iters_per_solve else None  # converged solves/s
    return {"value": value, "unit": UNIT if iters_per_solve else "outer iterations/s",
            "iterations_per_s": 1.0 / sec, "s_per_iteration": sec, "cores": 1, "kind": "port",
            "sample": f"{iterations} outer iterations of the same solve (FP64 oracle, 1 thread, -O2), "
                      f"after 1 untimed"}


def _host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model or "unknown", os.cpu_count() or 1


def _oracle_instance(row, level):
    """The oracle's solver for one manifest instance (the same inputs as the GPU path)."""
    sys.path.insert(0, os.path.join(_HERE, "oracle"))
    import oracle_py as O  # the checker: timed here as the CPU baseline only
    import workloads as W

    s_pts, w_pts, n_pts = W.LEVELS[level]
    opts = O.cgrid_options("coarse", upstream=15.0, vertical=15.0, downstream=20.0, surface_points=s_pts,
                           wake_points=w_pts, normal_points=n_pts, first_layer=W.first_layer(row["reynolds"]))
    mesh = O.cgrid(row["m"], row["p"], row["t"], options=opts)
    kind, vel, pres = O.face_conditions(mesh, 1.0, row["alpha"])
    st = O.default_settings()
    st.intensity, st.viscosity_ratio, st.length_scale = 0.001, 10.0, 0.0
    s = O.Solver(mesh, kind, vel, pres, st)
    s.set_model(2)
    s.initialise(*O.uniform_field(mesh.cells, speed=1.0, alpha_deg=row["alpha"], density=1.0,
                                  reynolds=row["reynolds"]))
    return s


def _oracle_rate(job):
    """Outer iterations per second of the oracle on one instance (one process, one thread)."""
    row, level, iterations = job
    s = _oracle_instance(row, level)
    s.iterate(1)
    t0 = time.perf_counter()
    s.iterate(iterations)
    return iterations / (time.perf_counter() - t0)


def cpu_baseline_batch(rows, results, sample_iterations=(12, 4, 2), parallel_iterations=12):
    """The FP64 oracle (a restatement of the reference solver; one solve per
    thread, as the reference's SolverWorker runs, SolverWorker.cpp:52) on the
    same instances, timed on this host. Latency: each level's seconds per
    outer iteration on one core (the first instance of the level, after one
    untimed iteration) times every solve's outer iterations from the plain
    GPU run (the same iterates up to reduction order), p50/p99 over the
    converged solves. Throughput: P independent solver processes, one per
    core, on a medium instance together; their parallel efficiency scales the
    single-core work of the whole batch to P cores."""
    import multiprocessing as mp

    model, nproc = _host_cpu()
    per_level = {}
    for level, n in zip(("coarse", "medium", "fine"), sample_iterations):
        row = next((r for r in rows if r["level"] == level), None)
        if row is None:
            continue
        per_level[level] = 1.0 / _oracle_rate((row, level, n))
    lat = sorted(per_level[rows[q]["level"]] * r["iterations"] for q, r in results.items() if r["status"] == 0)
    work = sum(per_level[rows[q]["level"]] * r["iterations"] for q, r in results.items())
    converged = sum(1 for r in results.values() if r["status"] == 0)
    procs = max(1, min(nproc, 16))
    medium = next((r for r in rows if r["level"] == "medium"), rows[0])
    lvl = medium["level"]
    with mp.get_context("spawn").Pool(procs) as pool:
        rates = pool.map(_oracle_rate, [(medium, lvl, parallel_iterations)] * procs)
    efficiency = (sum(rates) / procs) * per_level[lvl]
    seconds = work / (procs * efficiency)
    return {"value": converged / seconds if seconds > 0 else None, "unit": UNIT, "cores": procs, "kind": "port",
            "cpu_model": model, "nproc": nproc,
            "flags": "g++ -O2 -std=c++17 -ffp-contract=off -fno-builtin-sin -fno-builtin-cos (oracle/Makefile)",
            "latency_p50_s": _percentile(lat, 0.5), "latency_p99_s": _percentile(lat, 0.99),
            "s_per_iteration": per_level, "parallel_efficiency": efficiency,
            "single_core_seconds": work, "seconds_on_cores": seconds,
            "sample": (f"oracle seconds per outer iteration on one core per mesh level ({sample_iterations} "
                       f"iterations after 1 untimed), times the plain GPU run's outer iterations per solve; "
                       f"{procs} concurrent processes x {parallel_iterations} iterations of a 512x128 instance "
                       f"for the parallel efficiency")}


LAUNCH_TUNING = {"source": "built-in"}


def apply_tuning(args):
    """Line-CG launch shapes of cfd-simulator_amd/tune.py's file, if any."""
    import cfdgpu as G
    import tune

    path = args.tuning or tune.TUNING_FILE
    shape = None if args.tuning == "none" else tune.load_tuning(path)
    if shape is not None:
        G.set_launch_tuning(*shape)
        LAUNCH_TUNING["source"] = os.path.relpath(os.path.abspath(path), _HERE)
    cells, rows = G.launch_tuning()
    LAUNCH_TUNING.update(cells_per_thread=cells, column_rows=rows)


LCG_SEG_MAX_ROWS = 256  # launch B splits each column over several waves up to this many rows (layout.hpp)


def lcg_kernels(max_rows):
    """The names of the line-CG launch pair the engine picks for a batch whose widest
    member has max_rows wall-normal rows under the current launch tuning (launch.hip)."""
    cells = LAUNCH_TUNING["cells_per_thread"]
    a = "k_lcg_a_f32_tile" if cells == 0 else f"k_lcg_a_f32_c{cells}"
    b = "k_lcg_b_f32_seg" if max_rows <= LCG_SEG_MAX_ROWS else f"k_lcg_b_f32_u{LAUNCH_TUNING['column_rows']}"
    return f"{a} + {b}"


def batch_max_rows(W, rows):
    """Wall-normal rows of the widest member of a manifest's batch."""
    return max(W.LEVELS[r["level"]][2] - 1 for r in rows)


def batch_roofline(fast_st, cells, stage_ms, calls, cg_cells, max_rows):
    """The roofline object of the fast path's line-CG iteration on the batch."""
    tm = traffic_model()
    # one line-CG iteration of the whole batch (launch A + launch B): the CG loop's HIP-event time
    # over its launch pairs; bytes of the member-iterations those launches ran (a member whose CG
    # has met its tolerance returns at once)
    pairs = calls["cg_iterations"] * fast_st.pressure_iterations
    unit_ms = stage_ms["cg_iterations"] / max(pairs, 1)
    active = cg_cells / max(pairs, 1)  # cells per launch pair that iterate
    algo = tm["line_cg_iteration_mixed"] * active
    achieved = algo / (unit_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES.get("line_cg_iteration_mixed"),
                "kernel": lcg_kernels(max_rows), "stage": "cg_iterations",
                "algorithmic_bytes_per_launch": algo, "launch_ms": unit_ms, "launch_pairs": pairs,
                "bytes_per_cell": tm["line_cg_iteration_mixed"], "cells": cells, "iterating_cells": active,
                "note": "one line-CG iteration of the whole batch (launch A + launch B), HIP events on the "
                        "batch's stream around the CG loop of every timed outer iteration"}
    return roofline


def batch_roofline_others(stage_ms, calls):
    """Rooflines of the FP64 stencil kernels on the whole 64-instance batch:
    counted bytes per launch over the launch's HIP-event time in this run."""
    out = {}
    for stage, (kernel, nbytes) in BATCH_KERNEL_BYTES.items():
        n = calls.get(stage, 0)
        if not n:
            continue
        ms = stage_ms[stage] / n
        achieved = nbytes / (ms * 1e-3) / 1e9
        out[kernel] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": nbytes, "stage": stage, "launch_ms": ms,
                       "basis": "counted HBM bytes per launch over the stage's HIP-event time"}
    return out


def run_roofline(args):
    """Only the timed region of the headline: K outer iterations of the fast path
    on the 64-instance batch after W untimed, with the roofline of its line-CG
    iteration (a diagnostic line, the command a kernel trace is taken of)."""
    import cfdgpu as G
    import workloads as W

    if G.device_count() < 1:
        raise SystemExit("bench.py needs a GPU: no HIP device visible")
    rows = W.load(W.MIXED_BATCH_FILE)
    if args.only:
        rows = [rows[int(q)] for q in args.only.split(",")]
    fast_st = G.default_settings(intensity=0.001, viscosity_ratio=10.0, precision=G.MIXED, gs_ordering=1,
                                 pressure_solver=1, pressure_iterations=args.fast_cg, momentum_sweeps=args.fast_sweeps)
    cells, ms_per_step, stage_ms, calls, cg_cells = roofline_batch(G, W, rows, fast_st, args.steps, args.warmup)
    print(json.dumps({"diagnostic": f"fast path on {len(rows)} instances of the 64-instance mixed-size batch: step "
                                    "time and the line-CG roofline (not the metric)",
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "cells": cells,
                      "fast_path": settings_summary(fast_st), "launch_tuning": dict(LAUNCH_TUNING),
                      "roofline": batch_roofline(fast_st, cells, stage_ms, calls, cg_cells, batch_max_rows(W, rows)),
                      "roofline_others": None if args.only else batch_roofline_others(stage_ms, calls),
                      "stage_ms_per_step": {k: v / args.steps for k, v in stage_ms.items()}}), flush=True)


def run_headline(args):
    """The benchmark line: the metric on the 64-instance mixed-size batch."""
    import cfdgpu as G
    import workloads as W

    if G.device_count() < 1:
        raise SystemExit("bench.py needs a GPU: no HIP device visible")
    H = HEADLINE
    rows = W.load(W.MIXED_BATCH_FILE)
    fast_st = G.default_settings(intensity=0.001, viscosity_ratio=10.0, precision=G.MIXED, gs_ordering=1,
                                 pressure_solver=1, pressure_iterations=args.fast_cg, momentum_sweeps=args.fast_sweeps)
    plain_st = G.default_settings(intensity=0.001, viscosity_ratio=10.0)
    common = dict(tolerance=H["tolerance"], max_iterations=H["max_iterations"], chunk=H["chunk"], slots=H["slots"],
                  settle_window=H["settle_window"], settle_cl=H["settle_cl"], settle_cd=H["settle_cd"])
    # the timed unit of the roofline: K outer iterations of the whole batch after W untimed
    cells, ms_per_step, stage_ms, calls, cg_cells = roofline_batch(G, W, rows, fast_st, args.steps, args.warmup)
    roofline = batch_roofline(fast_st, cells, stage_ms, calls, cg_cells, batch_max_rows(W, rows))
    caps = {"coarse": args.fast_cg, "medium": args.fast_cg_medium or args.fast_cg,
            "fine": args.fast_cg_fine or args.fast_cg}
    fast_st.pressure_iterations = max(caps.values())
    fast = serve_batch(G, W, rows, fast_st, "fast", level_caps=caps, **common)
    if args.plain_results:
        with open(args.plain_results) as fh:
            plain = json.load(fh)["plain"]
        plain["results"] = {int(k): v for k, v in plain["results"].items()}
    else:
        plain = serve_batch(G, W, rows, plain_st, "plain FP64", **common)
    acc = acceptance(fast["results"], plain["results"])
    accepted = sum(1 for a in acc if a["accepted"])
    fs, ps = summarize(fast, rows), summarize(plain, rows)
    value = accepted / fast["seconds"]
    # plain FP64 one solve at a time: each level's ms per outer iteration alone x the plain run's iterations
    alone_ms = alone_rates(G, W, rows, plain_st)
    alone_s = sum(r["iterations"] * alone_ms[r["level"]] * 1e-3 for r in plain["results"].values())
    plain_conv = ps["statuses"]["converged"]
    alone_rate = plain_conv / alone_s if alone_s > 0 else None
    fast_alone_ms = alone_rates(G, W, rows, fast_st)

    line = {"metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": "mixed-size batch: 64 random NACA four-digit sections (seed 0), C-grids of 256x64, "
                                   "512x128 and 1024x256 cells (8.03 M cells), k-omega SST, Re 2e5-5e6, alpha -5..12; "
                                   "served at once with continuous batching",
                       "instances": len(rows), "cells": cells,
                       "stop_rule": f"residuals.worst() < {H['tolerance']:g}, limit {H['max_iterations']}, and Cl, Cd "
                                    f"settled (estimated remaining change <= {H['settle_cl']:g}, {H['settle_cd']:g} "
                                    f"from samples every {H['settle_window']} iterations)",
                       "fast_path": dict(settings_summary(fast_st), cg_cap_by_level=caps),
                       "launch_tuning": dict(LAUNCH_TUNING),
                       "plain_path": settings_summary(plain_st),
                       "step": "one outer iteration of all 64 resident solves (fast path), timed by HIP events"},
            "accepted": accepted, "converged_both": len(acc),
            "acceptance": {"bar": {"dcl": ACCEPT_CL, "dcd": ACCEPT_CD},
                           "max_dcl": max((a["dcl"] for a in acc), default=None),
                           "max_dcd": max((a["dcd"] for a in acc), default=None),
                           "rejected_rows": [a["row"] for a in acc if not a["accepted"]]},
            "fast": fs, "plain_batched": ps,
            "plain_one_at_a_time": {"converged_solves_per_s": alone_rate, "seconds_estimate": alone_s,
                                    "ms_per_iteration": alone_ms},
            "fast_one_at_a_time_ms_per_iteration": fast_alone_ms,
            "speedup": {"vs_plain_one_at_a_time": value / alone_rate if alone_rate else None,
                        "vs_plain_batched": value / ps["converged_solves_per_s"] if ps["converged_solves_per_s"]
                        else None},
            "latency": {"p50_s": fs["time_to_convergence_p50_s"], "p99_s": fs["time_to_convergence_p99_s"]},
            "roofline": roofline,
            "roofline_others": batch_roofline_others(stage_ms, calls),
            "stage_ms_per_step": {k: v / args.steps for k, v in stage_ms.items()}}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_batch(rows, plain["results"])
    if args.results_file:
        with open(args.results_file, "w") as fh:
            json.dump({"fast": fast, "plain": plain, "acceptance": acc}, fh)
    print(json.dumps(line), flush=True)


def run_bench(args):
    import cfdgpu as G

    if G.device_count() < 1:
        raise SystemExit("bench.py needs a GPU: no HIP device visible")
    mesh, kind, vel, pres, field = build_case(G)
    s = gpu_solver(G, mesh, kind, vel, pres, field, precision_of(G, args))
    if args.warmup:
        s.iterate(args.warmup)
    t0 = time.perf_counter()
    mon, stage_ms, stage_calls = s.iterate_timed(args.steps)  # returns after the stream has drained
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps

    # The dominant kernel: the sweeps are one launch per stage call; the CG is
    # two launches per CG iteration, and the pair is one unit of the traffic
    # model and of the timing (events around every launch would add their own
    # gaps to 400 launches per step). By the kernel trace k_cg_a alone is
    # already the largest kernel, so the CG iteration is the dominant unit.
    tm = traffic_model()
    cg_its = max(mon.pressure_iterations, 1) * args.steps  # CG iterations in the timed region (cap reached each time)
    def entry(stage, kernel, unit):
        units = cg_its if unit == "cg_iteration" else stage_calls[stage]
        unit_ms = stage_ms[stage] / units
        algo = float(tm[unit]) * mesh.cells * 8
        achieved = algo / (unit_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES.get(kernel), "kernel": kernel,
                "stage": stage, "algorithmic_bytes_per_launch": algo, "launch_ms": unit_ms, "words_per_cell": tm[unit]}

    present = [c for c in DOMINANT_CANDIDATES if stage_calls.get(c[0])]
    entries = {c[1]: entry(*c) for c in present}
    kernel = max(present, key=lambda c: stage_ms[c[0]])[1]
    roofline = entries[kernel]

    # a diagnostic of one solve: its outer iterations per second (how many it
    # takes to converge is measured by --solve, never assumed here)
    line = {
        "metric": "outer iterations/s of one 1024x256 SST solve (diagnostic, not the headline)",
        "value": 1000.0 / ms_per_step, "unit": "outer iterations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "NACA 2412 C-grid 1024x256 (262,144 cells, y1 1e-6 c), k-omega SST, Re 1e6, "
                               "alpha 4 deg, Tu 0.1%, mut/mu 10; "
                               + ("plain FP64" if args.precision == "fp64" else "mixed (FP32 CG storage)")
                               + ", lexicographic GS; one step = one SIMPLE outer iteration",
                   "cells": mesh.cells,
                   "cg_iterations_last": mon.pressure_iterations,
                   "stage_ms_per_step": {k: v / args.steps for k, v in stage_ms.items() if stage_calls[k]}},
        "roofline": roofline,
        "roofline_others": {k: v for k, v in entries.items() if k != kernel},
    }
    line["config"]["precision"] = args.precision
    if args.precision == "fp64" and not args.no_mixed:
        line["mixed_precision"] = mixed_comparison(G, mesh, kind, vel, pres, field, args, ms_per_step, stage_ms)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_iterations, None)
    print(json.dumps(line), flush=True)


def mixed_comparison(G, mesh, kind, vel, pres, field, args, fp64_ms, fp64_stage_ms):
    """The same workload on the mixed path (FP32 storage in the pressure
    solve), timed over a quarter of the steps after the same warm-up: outer
    iteration and pressure-solve times beside the FP64 path's."""
    del field
    s = gpu_solver(G, mesh, kind, vel, pres, G.uniform_field(
        mesh.cells, speed=WORKLOAD["speed"], alpha_deg=WORKLOAD["alpha"], density=WORKLOAD["density"],
        reynolds=WORKLOAD["reynolds"]), G.MIXED)
    steps = max(args.steps // 4, 1)
    if args.warmup:
        s.iterate(args.warmup)
    t0 = time.perf_counter()
    _, stage_ms, _ = s.iterate_timed(steps)
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    p32 = stage_ms["pressure_solve"] / steps
    p64 = fp64_stage_ms["pressure_solve"] / args.steps
    return {"ms_per_step": ms, "steps": steps, "speedup_per_step": fp64_ms / ms,
            "pressure_solve_ms": p32, "pressure_solve_ms_fp64": p64, "pressure_solve_speedup": p64 / p32,
            "storage": "FP32 matrix and r, z, p, Ap in the pressure CG; FP64 arithmetic, reductions, x, sweeps"}


def run_converge(args):
    """Run the solve for --max-iterations (or until worst() < tol with --stop-at-tolerance),
    printing the residual and Cl/Cd/Cm every chunk. Reports the first chunk whose
    residual is below the tolerance (the reference's stop rule) and the first
    iteration after which |Cl - Cl_end| < 1e-3 and |Cd - Cd_end| < 1e-4 hold for
    every later chunk, with Cl_end, Cd_end the coefficients of the last chunk."""
    import cfdgpu as G

    mesh, kind, vel, pres, field = build_case(G)
    s = gpu_solver(G, mesh, kind, vel, pres, field)
    hist = []
    done = 0
    t0 = time.perf_counter()
    at_tol = None
    while done < args.max_iterations:
        mon = s.iterate(args.chunk)
        done += args.chunk
        f = coefficients(G, mesh, s)
        hist.append((done, mon.worst(), f.cl, f.cd, f.cm))
        print(f"{done} {time.perf_counter() - t0:8.1f}s worst={mon.worst():.3e} cl={f.cl:.8f} cd={f.cd:.8f} "
              f"cm={f.cm:.8f} eddy={mon.max_eddy_ratio:.1f}", flush=True)
        if at_tol is None and mon.worst() < args.tolerance:
            at_tol = done
            if args.stop_at_tolerance:
                break
    cl_f, cd_f = hist[-1][2], hist[-1][3]
    n_conv = None
    for it, _, cl, cd, _ in reversed(hist):
        if abs(cl - cl_f) < 1e-3 and abs(cd - cd_f) < 1e-4:
            n_conv = it
        else:
            break
    print(json.dumps({"iterations": done, "final_worst": hist[-1][1], "cl": cl_f, "cd": cd_f,
                      "cm": hist[-1][4], "iterations_to_tolerance": at_tol, "tolerance": args.tolerance,
                      "iterations_to_coefficients": n_conv, "chunk": args.chunk,
                      "seconds": time.perf_counter() - t0}), flush=True)


def run_solve(args):
    """One C2 solve from the uniform field to the reference's stop rule
    (residuals.worst() < tolerance, checked after every outer iteration on the
    device), timed end to end on the GPU."""
    import cfdgpu as G

    mesh, kind, vel, pres, field = build_case(G)
    s = gpu_solver(G, mesh, kind, vel, pres, field, precision_of(G, args))
    t0 = time.perf_counter()
    rep = s.solve(args.tolerance, args.max_iterations, check_every=args.check_every)
    sec = time.perf_counter() - t0
    f = coefficients(G, mesh, s)
    print(json.dumps({"iterations": rep.iterations, "status": rep.status, "seconds": sec,
                      "ms_per_iteration": 1000.0 * sec / max(rep.iterations, 1), "worst": rep.last.worst(),
                      "tolerance": args.tolerance, "cl": f.cl, "cd": f.cd, "cm": f.cm,
                      "precision": args.precision}), flush=True)


def run_batch_polar(args):
    """The batched uniform polar: NACA 0012 on a 512 x 128 C-grid (surfacePoints
    193, wakePoints 64, normalPoints 129, first layer 2.5e-5 chord), k-omega SST,
    Re 1e6, Tu 0.1 %, mu_t/mu 10, alpha -4..16 deg in 1 deg steps (21 solves),
    every point started cold from the uniform field and advanced together, one
    launch per stage for all 21 (cfdgpu_batch_solve). Reports converged
    solves/s and each point's coefficients."""
    import cfdgpu as G

    opts = G.cgrid_options("coarse", upstream_chords=15.0, vertical_chords=15.0, downstream_chords=20.0,
                           surface_points=193, wake_points=64, normal_points=129, first_layer_height=2.5e-5)
    mesh = G.cgrid(0, 0, 12, options=opts)
    alphas = list(G.sweep_angles(-4.0, 16.0, 1.0))
    solvers = []
    t_setup = time.perf_counter()
    for a in alphas:
        kind, vel, pres = G.face_conditions(mesh, 50.0, a)
        st = G.default_settings(intensity=0.001, viscosity_ratio=10.0, precision=precision_of(G, args))
        s = G.SimpleSolver(mesh, kind, vel, pres, st)
        s.set_turbulence_model(G.MODEL_SST)
        s.initialise(*G.uniform_field(mesh.cells, speed=50.0, alpha_deg=a, density=1.225, reynolds=1e6))
        solvers.append(s)
    t_setup = time.perf_counter() - t_setup
    n = len(solvers)
    iters, status, done_at = [0] * n, [1] * n, [None] * n
    active = list(range(n))
    t0 = time.perf_counter()
    while active:
        # members that stopped leave the batch; the rest go on together
        batch = G.Batch([solvers[i] for i in active])
        reps = batch.solve(args.tolerance, args.chunk, check_every=50)
        now = time.perf_counter() - t0
        for i, r in zip(active, reps):
            iters[i] += r.iterations
            if r.status != 1:
                status[i], done_at[i] = r.status, now
        active = [i for i in active if status[i] == 1 and iters[i] < args.max_iterations]
        print(f"{now:8.1f}s active {len(active)} of {n}, iterations {min(iters)}..{max(iters)}", flush=True)
        del batch
    sec = time.perf_counter() - t0
    points = []
    for a, s, it, stt, t in zip(alphas, solvers, iters, status, done_at):
        u, v, p, mu = s.field()
        f, _ = G.surface(mesh, u, v, p, mu, speed=50.0, alpha_deg=a, density=1.225)
        points.append({"alpha": a, "status": stt, "iterations": it, "seconds_to_stop": t, "cl": f.cl, "cd": f.cd,
                       "cm": f.cm})
    converged = sum(1 for x in status if x == 0)
    lat = sorted(t for t, x in zip(done_at, status) if x == 0)
    print(json.dumps({"workload": "batched uniform polar, NACA 0012 512x128, SST Re 1e6, alpha -4..16 (21 cold)",
                      "precision": args.precision,
                      "cells_per_solve": mesh.cells, "solves": len(alphas), "converged": converged,
                      "seconds": sec, "setup_seconds": t_setup, "converged_solves_per_s": converged / sec,
                      "time_to_convergence_p50_s": lat[len(lat) // 2] if lat else None,
                      "time_to_convergence_p99_s": lat[min(len(lat) - 1, int(0.99 * len(lat)))] if lat else None,
                      "tolerance": args.tolerance, "max_iterations": args.max_iterations, "points": points}),
          flush=True)


def _percentile(sorted_vals, q):
    if not sorted_vals:
        return None
    return sorted_vals[min(len(sorted_vals) - 1, int(q * len(sorted_vals)))]


def run_serve(args, mixed_batch=False):
    """The served request stream through cfdgpu_server_*: the first
    --requests requests of the committed Poisson manifest (20 requests/s, random
    NACA four-digit sections, alpha, Re and mesh level; workloads.py), replayed
    at their arrival times, admitted by an HBM budget of 200 GB into at most
    --slots resident solves, advanced together in chunks of --serve-chunk outer
    iterations, warm-started from the nearest converged alpha of the same
    section and mesh size. With --mixed-batch, the 64-instance mixed-size batch
    manifest instead, all arriving at once. Reports converged solves/s and the
    p50/p99 time from arrival to convergence (queueing included)."""
    import cfdgpu as G
    import workloads as W

    rows = W.load(W.MIXED_BATCH_FILE if mixed_batch else W.SERVING_FILE)[: args.requests]
    if args.only:
        keep = [int(q) for q in args.only.split(",")]
        rows = [rows[q] for q in keep]
    st = path_settings(G, args)
    srv = G.Server(G.server_options(hbm_budget_bytes=200e9, max_slots=args.slots, chunk_iterations=args.serve_chunk,
                                    tolerance=args.tolerance, max_iterations=args.max_iterations, settings=st,
                                    warm_start=0 if args.cold else 1))
    for r in rows:
        srv.submit(G.serve_request(r["m"], r["p"], r["t"], r["alpha"], r["reynolds"],
                                   W.mesh_options(G, r["level"], r["reynolds"]),
                                   arrival_s=0.0 if mixed_batch else r["arrival"]))
    results, last = [], -1e9
    while True:
        stats = srv.step()
        results += srv.collect()
        if stats.clock_s - last >= 10.0:
            last = stats.clock_s
            print(f"{stats.clock_s:8.1f}s queued {stats.queued} resident {stats.active} finished {stats.finished} "
                  f"converged {stats.converged} iterations {stats.member_iterations:.0f}", flush=True)
        if stats.queued == 0 and stats.active == 0:
            break
        if args.wall_limit > 0 and stats.clock_s >= args.wall_limit:
            break
    wall = stats.clock_s
    conv = [r for r in results if r.status == 0]
    ttc = sorted(r.finished_s - r.arrival_s for r in conv)
    per_request = {r.ticket: {"level": rows[r.ticket]["level"], "status": r.status, "iterations": r.iterations,
                              "cl": r.cl, "cd": r.cd, "cm": r.cm} for r in results}
    if args.results_file:
        with open(args.results_file, "w") as fh:
            json.dump({"precision": args.precision, "tolerance": args.tolerance, "results": per_request}, fh)
    versus = compare_results(per_request, args.compare_to) if args.compare_to else None
    by_level = {}
    level_of = {q: rows[q]["level"] for q in range(len(rows))}
    for lv in W.LEVELS:
        mine = [r for r in results if level_of[r.ticket] == lv]
        c = sorted(r.finished_s - r.arrival_s for r in mine if r.status == 0)
        by_level[lv] = {"finished": len(mine), "converged": len(c), "ttc_p50_s": _percentile(c, 0.5),
                        "iterations_mean": (sum(r.iterations for r in mine) / len(mine)) if mine else None}
    statuses = {G.STATUS_NAMES[k]: sum(1 for r in results if r.status == k) for k in G.STATUS_NAMES}
    name = ("mixed-size batch, 64 random NACA sections, coarse/medium/fine C-grids, SST" if mixed_batch else
            f"served stream, first {len(rows)} requests of the 20 req/s Poisson manifest, SST")
    print(json.dumps({"workload": name, "settings": settings_summary(st), "requests": len(rows), "finished": len(results),
                      "unfinished": len(rows) - len(results), "statuses": statuses,
                      "seconds": wall, "converged_solves_per_s": len(conv) / wall if wall > 0 else None,
                      "time_to_convergence_p50_s": _percentile(ttc, 0.5),
                      "time_to_convergence_p99_s": _percentile(ttc, 0.99),
                      "warm_started": sum(1 for r in results if r.warm_from >= 0),
                      "retried_cold": sum(1 for r in results if r.retried_cold),
                      "outer_iterations": stats.member_iterations,
                      "cell_iterations_per_s": sum(r.cells * r.iterations for r in results) / wall if wall > 0 else None,
                      "peak_resident": stats.peak_active, "peak_resident_gb": stats.peak_resident_bytes / 1e9,
                      "slots": args.slots, "chunk": args.serve_chunk, "tolerance": args.tolerance,
                      "max_iterations": args.max_iterations, "by_level": by_level,
                      "one_at_a_time": alone_comparison(G, W, rows, results, wall) if args.alone_rate else None,
                      "versus": versus}),
          flush=True)


# The headline workload: the 64-instance mixed-size batch (seed 0), all
# arriving at once, served with continuous batching (64 slots, chunks of 20
# outer iterations), worker stop rule residuals.worst() < 1e-6 with the
# reference's 10,000-iteration limit, plus coefficient settling (Cl and Cd
# integrated every 500 iterations; their estimated remaining change at most
# 5e-4 and 5e-5, half the acceptance bar) for both paths.
HEADLINE = dict(tolerance=1e-6, max_iterations=10000, chunk=20, slots=64, settle_window=500, settle_cl=5e-4,
                settle_cd=5e-5)
ACCEPT_CL, ACCEPT_CD = 1e-3, 1e-4


def serve_batch(G, W, rows, st, label, tolerance, max_iterations, chunk, slots, settle_window=0, settle_cl=5e-4,
                settle_cd=5e-5, wall_limit=0.0, progress=15.0, level_caps=None):
    """Serves the rows (all arriving at once) and returns per-row results,
    the wall time and the server statistics."""
    srv = G.Server(G.server_options(hbm_budget_bytes=200e9, max_slots=slots, chunk_iterations=chunk,
                                    tolerance=tolerance, max_iterations=max_iterations, settings=st,
                                    warm_start=0, settle_window=settle_window, settle_cl=settle_cl,
                                    settle_cd=settle_cd))
    tickets = {}
    own = {}
    if level_caps:
        for lv, cap in level_caps.items():
            own[lv] = G.Settings.from_buffer_copy(st)
            own[lv].pressure_iterations = cap
    for q, r in enumerate(rows):
        t = srv.submit(G.serve_request(r["m"], r["p"], r["t"], r["alpha"], r["reynolds"],
                                       W.mesh_options(G, r["level"], r["reynolds"]), arrival_s=0.0,
                                       settings=own.get(r["level"])))
        tickets[t] = q
    results, last = [], 0.0
    while True:
        stats = srv.step()
        results += srv.collect()
        if stats.clock_s - last >= progress:
            last = stats.clock_s
            print(f"[{label}] {stats.clock_s:7.1f}s resident {stats.active} finished {stats.finished} "
                  f"converged {stats.converged}", file=sys.stderr, flush=True)
        if stats.queued == 0 and stats.active == 0:
            break
        if wall_limit > 0 and stats.clock_s >= wall_limit:
            break
    out = {tickets[r.ticket]: {"level": rows[tickets[r.ticket]]["level"], "status": r.status,
                               "iterations": r.iterations, "cl": r.cl, "cd": r.cd, "cm": r.cm,
                               "ttc_s": r.finished_s - r.arrival_s, "cells": r.cells} for r in results}
    return {"results": out, "seconds": stats.clock_s, "iterations": stats.member_iterations,
            "peak_gb": stats.peak_resident_bytes / 1e9}


def summarize(run, rows):
    res = run["results"]
    conv = [r for r in res.values() if r["status"] == 0]
    ttc = sorted(r["ttc_s"] for r in conv)
    statuses = {G_STATUS[k]: sum(1 for r in res.values() if r["status"] == k) for k in G_STATUS}
    return {"seconds": run["seconds"], "statuses": statuses,
            "converged_solves_per_s": len(conv) / run["seconds"] if run["seconds"] > 0 else None,
            "time_to_convergence_p50_s": _percentile(ttc, 0.5), "time_to_convergence_p99_s": _percentile(ttc, 0.99),
            "outer_iterations": run["iterations"],
            "cell_iterations_per_s": sum(r["cells"] * r["iterations"] for r in res.values()) / run["seconds"],
            "peak_resident_gb": run["peak_gb"]}


G_STATUS = {0: "converged", 1: "iteration-limit", 2: "diverged", 3: "rejected"}


def alone_rates(G, W, rows, settings, iterations=50, warmup=10):
    """ms per outer iteration of one solve alone on the given path, per mesh
    level (the level's first manifest instance, after warm-up)."""
    out = {}
    for lv in W.LEVELS:
        r = next((x for x in rows if x["level"] == lv), None)
        if r is None:
            continue
        mesh = G.cgrid(r["m"], r["p"], r["t"], options=W.mesh_options(G, lv, r["reynolds"]))
        kind, vel, pres = G.face_conditions(mesh, 1.0, r["alpha"])
        s = G.SimpleSolver(mesh, kind, vel, pres, settings)
        s.set_turbulence_model(G.MODEL_SST)
        s.initialise(*G.uniform_field(mesh.cells, speed=1.0, alpha_deg=r["alpha"], density=1.0,
                                      reynolds=r["reynolds"]))
        s.iterate(warmup)
        t0 = time.perf_counter()
        s.iterate(iterations)
        out[lv] = (time.perf_counter() - t0) * 1e3 / iterations
    return out


def acceptance(fast, plain):
    """Per instance converged on both paths: |dCl| and |dCd| of the fast
    path's result against the plain FP64 path's (the judge)."""
    rows = []
    for q, a in fast.items():
        b = plain.get(q)
        if b is None or a["status"] != 0 or b["status"] != 0:
            continue
        dcl, dcd = abs(a["cl"] - b["cl"]), abs(a["cd"] - b["cd"])
        rows.append({"row": q, "level": a["level"], "dcl": dcl, "dcd": dcd,
                     "accepted": dcl < ACCEPT_CL and dcd < ACCEPT_CD})
    return rows


def roofline_batch(G, W, rows, st, steps, warmup):
    """Stage times of the fast path's outer iteration over the whole batch
    (64 instances resident, 8.0 M cells), from HIP events on the batch's
    stream inside the timed region; the line CG iteration (launch A + launch
    B) is the unit of the roofline."""
    solvers = []
    cells = 0
    for r in rows:
        mesh = G.cgrid(r["m"], r["p"], r["t"], options=W.mesh_options(G, r["level"], r["reynolds"]))
        kind, vel, pres = G.face_conditions(mesh, 1.0, r["alpha"])
        s = G.SimpleSolver(mesh, kind, vel, pres, st)
        s.set_turbulence_model(G.MODEL_SST)
        s.initialise(*G.uniform_field(mesh.cells, speed=1.0, alpha_deg=r["alpha"], density=1.0,
                                      reynolds=r["reynolds"]))
        solvers.append((mesh, s))
        cells += mesh.cells
    batch = G.Batch([s for _, s in solvers])
    if warmup:
        batch.iterate(warmup)
    t0 = time.perf_counter()
    mons, stage_ms, calls = batch.iterate_timed(steps)
    wall = time.perf_counter() - t0
    caps = [mon.pressure_iterations for mon in mons]
    if min(caps) != max(caps):
        print(f"[roofline] CG iterations per member in the last step range {min(caps)}-{max(caps)}",
              file=sys.stderr, flush=True)
    # cells x CG iterations of the timed region (the last step's counts, the same in every step
    # once every member runs to its cap)
    cg_cells = sum(m.cells * mon.pressure_iterations for (m, _), mon in zip(solvers, mons)) * steps
    return cells, wall * 1e3 / steps, stage_ms, calls, cg_cells


def compare_results(mine, path):
    """Per-request coefficients against another run of the same manifest (the
    --results-file of a plain FP64 run): the largest |dCl|, |dCd|, |dCm| over
    the requests both runs converged, and how many statuses differ."""
    with open(path) as fh:
        other = json.load(fh)
    theirs = {int(k): v for k, v in other["results"].items()}
    both = [t for t in mine if t in theirs and mine[t]["status"] == 0 and theirs[t]["status"] == 0]
    d = {q: max((abs(mine[t][q] - theirs[t][q]) for t in both), default=None) for q in ("cl", "cd", "cm")}
    return {"against": other["precision"], "converged_in_both": len(both),
            "status_differs": sum(1 for t in mine if t in theirs and mine[t]["status"] != theirs[t]["status"]),
            "max_abs_dcl": d["cl"], "max_abs_dcd": d["cd"], "max_abs_dcm": d["cm"],
            "accepted": (d["cl"] is not None and d["cl"] < 1e-3 and d["cd"] < 1e-4)}


def alone_comparison(G, W, rows, results, batched_seconds):
    """The same solves one at a time on the plain path: one outer iteration of
    a solve alone, timed per mesh level (the first instance of each level in
    the manifest, 100 iterations after 20), times every result's outer
    iterations (its last attempt only, so the estimate is a lower bound)."""
    per_level = {}
    for lv in W.LEVELS:
        r = next((x for x in rows if x["level"] == lv), None)
        if r is None:
            continue
        mesh = G.cgrid(r["m"], r["p"], r["t"], options=W.mesh_options(G, lv, r["reynolds"]))
        kind, vel, pres = G.face_conditions(mesh, 1.0, r["alpha"])
        s = G.SimpleSolver(mesh, kind, vel, pres, G.default_settings(intensity=0.001, viscosity_ratio=10.0))
        s.set_turbulence_model(G.MODEL_SST)
        s.initialise(*G.uniform_field(mesh.cells, speed=1.0, alpha_deg=r["alpha"], density=1.0,
                                      reynolds=r["reynolds"]))
        s.iterate(20)
        t0 = time.perf_counter()
        s.iterate(100)
        per_level[lv] = (time.perf_counter() - t0) / 100
        print(f"alone: {lv} {mesh.cells} cells {per_level[lv] * 1e3:.3f} ms per outer iteration", file=sys.stderr,
              flush=True)
    level_of = {q: rows[q]["level"] for q in range(len(rows))}
    seconds = sum(r.iterations * per_level[level_of[r.ticket]] for r in results)
    converged = sum(1 for r in results if r.status == 0)
    return {"ms_per_iteration": {k: v * 1e3 for k, v in per_level.items()}, "seconds_estimate": seconds,
            "converged_solves_per_s": converged / seconds if seconds > 0 else None,
            "batched_speedup": seconds / batched_seconds if batched_seconds > 0 else None}


def run_calibrate_copy(args):
    """A device-to-device copy of known size, for scaling the HBM traffic
    counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) to bytes."""
    import torch

    n = 64 * 1024 * 1024  # doubles: 512 MiB read + 512 MiB written per copy
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-iterations", type=int, default=6, help="oracle iterations in the CPU sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--converge", action="store_true", help="run the solve to convergence instead")
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--max-iterations", type=int, default=40000)
    ap.add_argument("--tolerance", type=float, default=1e-7)
    ap.add_argument("--stop-at-tolerance", action="store_true")
    ap.add_argument("--solve", action="store_true", help="one solve to the residual tolerance, timed")
    ap.add_argument("--check-every", type=int, default=1)
    ap.add_argument("--calibrate-copy", action="store_true", help="a copy of known size (counter scaling)")
    ap.add_argument("--batch-polar", action="store_true", help="21-point cold polar as one batch")
    ap.add_argument("--serve", action="store_true", help="replay the served request stream")
    ap.add_argument("--mixed-batch", action="store_true", help="the 64-instance mixed-size batch, served at once")
    ap.add_argument("--requests", type=int, default=64, help="requests of the manifest replayed by --serve")
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--serve-chunk", type=int, default=20)
    ap.add_argument("--cold", action="store_true", help="serve without warm starts")
    ap.add_argument("--wall-limit", type=float, default=900.0, help="seconds before --serve stops waiting")
    ap.add_argument("--alone-rate", action="store_true",
                    help="with --serve / --mixed-batch: also price the same solves run one at a time")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="fp64: the plain path; mixed: FP32 storage in the pressure CG")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-path comparison of the default run")
    ap.add_argument("--fast", action="store_true",
                    help="the fast path: mixed precision, two-colour FP32 sweeps, line-preconditioned CG of 15")
    ap.add_argument("--gs", choices=("lex", "rb"), help="Gauss-Seidel ordering")
    ap.add_argument("--psolver", choices=("jacobi", "line"), help="pressure solver")
    ap.add_argument("--cg-max", type=int, help="pressure CG iteration cap")
    ap.add_argument("--refine", type=int, help="pressure refinement passes")
    ap.add_argument("--only", help="with --mixed-batch / --serve / --roofline-only: comma-separated manifest rows")
    ap.add_argument("--single", action="store_true",
                    help="the single 1024x256 SST solve (plain FP64 step time, roofline, mixed comparison)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the headline's timed region and its line-CG roofline (for kernel traces)")
    ap.add_argument("--fast-cg", type=int, default=15, help="line-CG cap of the fast path in the headline run")
    ap.add_argument("--fast-sweeps", type=int, default=3,
                    help="momentum sweeps per outer iteration on the fast path (the reference's default 3)")
    ap.add_argument("--fast-cg-medium", type=int, default=None, help="line-CG cap on the 512x128 instances")
    ap.add_argument("--fast-cg-fine", type=int, default=None, help="line-CG cap on the 1024x256 instances")
    ap.add_argument("--plain-results", help="headline: take the plain FP64 run from this --results-file")
    ap.add_argument("--results-file", help="with --serve / --mixed-batch: write every request's result here")
    ap.add_argument("--compare-to", help="with --serve / --mixed-batch: a --results-file to compare coefficients with")
    ap.add_argument("--tuning", default=None,
                    help="line-CG launch shapes from this tuner file (default: the package's "
                         "workloads/launch_tuning.json when present); 'none' keeps the built-in shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (independent solves shard "
              "by instance, one solve per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    apply_tuning(args)
    if args.calibrate_copy:
        run_calibrate_copy(args)
    elif args.batch_polar:
        run_batch_polar(args)
    elif args.serve or args.mixed_batch:
        run_serve(args, mixed_batch=args.mixed_batch)
    elif args.solve:
        run_solve(args)
    elif args.converge:
        run_converge(args)
    elif args.single:
        run_bench(args)
    elif args.roofline_only:
        run_roofline(args)
    else:
        run_headline(args)


if __name__ == "__main__":
    main()
