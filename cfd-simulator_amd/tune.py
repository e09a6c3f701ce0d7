"""Launch tuner of the line-preconditioned CG.

The CG iteration of the fast path is two launches: a cell-parallel one
(p = z + beta p, A p and its p.Ap partial sums; 1, 2 or 4 cells per thread,
or tiles of 64 x 16 cells staged in LDS) and a column-parallel one (alpha,
the residual update and the tridiagonal solve of each wall-normal column; 16
or 32 rows loaded ahead, or, for meshes of at most 256 rows, each column's
rows split over the waves of a block, which has no depth to tune). Every shape
computes the same iterates (the block partial sums split differently, so
results agree to rounding), and which is fastest depends on the batch's mesh
sizes and the device. The tuner builds the 64-instance mixed-size batch of
workloads/mixed_batch_64_seed0.json on the fast-path settings, times the
CG loop of a few outer iterations of each shape with HIP events on the
batch's stream, and writes the fastest shape with every measurement to
workloads/launch_tuning.json (or --out), which bench.py applies.

    python cfd-simulator_amd/tune.py [--steps 6] [--repeats 3] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import cfdgpu as G  # noqa: E402
import workloads as W  # noqa: E402

TUNING_FILE = os.path.join(HERE, "workloads", "launch_tuning.json")
CELLS = (0, 1, 2, 4)  # launch A: 0 = LDS tiles of 64 x 16 cells, else cells per thread
SEG_MAX_ROWS = 256    # meshes of at most this many rows take the segmented launch B (no depth to tune)


def shapes_for(max_rows: int):
    rows = (32,) if max_rows <= SEG_MAX_ROWS else (16, 32)
    return [(c, r) for c in CELLS for r in rows]


def fast_settings(cg_cap: int = 15):
    st = G.default_settings(intensity=0.001, viscosity_ratio=10.0, precision=G.MIXED)
    st.gs_ordering, st.pressure_solver, st.pressure_iterations = 1, 1, cg_cap
    return st


def build_batch(rows, st):
    solvers = []
    for r in rows:
        mesh = G.cgrid(r["m"], r["p"], r["t"], options=W.mesh_options(G, r["level"], r["reynolds"]))
        kind, vel, pres = G.face_conditions(mesh, 1.0, r["alpha"])
        s = G.SimpleSolver(mesh, kind, vel, pres, st)
        s.set_turbulence_model(G.MODEL_SST)
        s.initialise(*G.uniform_field(mesh.cells, speed=1.0, alpha_deg=r["alpha"], density=1.0,
                                      reynolds=r["reynolds"]))
        solvers.append((mesh, s))
    return solvers, G.Batch([s for _, s in solvers])


def tune(steps: int = 6, repeats: int = 3, warmup: int = 20, cg_cap: int = 15, log=print):
    rows = W.load(W.MIXED_BATCH_FILE)
    st = fast_settings(cg_cap)
    solvers, batch = build_batch(rows, st)
    cells = sum(m.cells for m, _ in solvers)
    shapes = shapes_for(max(m.nj for m, _ in solvers))
    batch.iterate(warmup)  # past the start-up transient, every member's CG at its cap
    before = G.launch_tuning()
    results = []
    try:
        for rep in range(repeats):
            for cells_per_thread, column_rows in shapes:
                G.set_launch_tuning(cells_per_thread, column_rows)
                batch.iterate(1)  # the first launches of a shape are not timed
                mons, stage_ms, calls = batch.iterate_timed(steps)
                cg = max(m.pressure_iterations for m in mons)
                results.append({"cells_per_thread": cells_per_thread, "column_rows": column_rows, "repeat": rep,
                                "cg_loop_ms": stage_ms["cg_iterations"] / steps, "cg_iterations": cg})
                log(f"[tune] repeat {rep} shape ({cells_per_thread}, {column_rows}): "
                    f"CG loop {stage_ms['cg_iterations'] / steps:.3f} ms per outer iteration")
    finally:
        G.set_launch_tuning(*before)
    best = {}
    for r in results:
        key = (r["cells_per_thread"], r["column_rows"])
        best[key] = min(best.get(key, float("inf")), r["cg_loop_ms"])
    winner = min(best, key=best.get)
    return {
        "cells_per_thread": winner[0], "column_rows": winner[1],
        "objective": "the CG loop (launches A and B of every iteration) per outer iteration, best of repeats",
        "workload": {"manifest": os.path.basename(W.MIXED_BATCH_FILE), "instances": len(rows), "cells": cells,
                     "cg_cap": cg_cap, "steps": steps, "repeats": repeats, "warmup": warmup},
        "best_ms": {f"{c}x{r}": round(v, 4) for (c, r), v in sorted(best.items())},
        "measurements": results,
        "date": time.strftime("%Y-%m-%d %H:%M:%S"),
    }


def load_tuning(path: str = TUNING_FILE):
    """(cells_per_thread, column_rows) of a tuning file, or None if there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        doc = json.load(f)
    return int(doc["cells_per_thread"]), int(doc["column_rows"])


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cg-cap", type=int, default=15)
    ap.add_argument("--out", default=TUNING_FILE)
    args = ap.parse_args()
    if G.device_count() < 1:
        sys.exit("tune.py needs a GPU")
    doc = tune(args.steps, args.repeats, args.warmup, args.cg_cap)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps({k: doc[k] for k in ("cells_per_thread", "column_rows", "best_ms")}))


if __name__ == "__main__":
    main()
