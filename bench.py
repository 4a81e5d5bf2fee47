"""Throughput of the D3Q27 cumulant step on one MI355X, against the HBM roofline.

Workload (default): NACA0012, chord 128 cells at 8 deg, Re 1e5, Ma 0.1, uniform
lattice 1536x1024x64 (100.7M cells), z periodic, characteristic farfield on x/y,
Bouzidi walls, Cl/Cd/Cm and the low-Mach guards recorded on every step.  One
step = one full lattice-Boltzmann time step of that case (boundary kernels +
fused AA stream-collide sweep + load/guard reductions).  Populations are
resident in HBM before timing starts.  The headline value is the FP16
shifted-storage run (FP32 collision); the FP64 and FP32 storage runs of the same
case are reported beside it, with the loads of each compared to FP64.

roofline: algorithmic bytes of one sweep launch = fluid cells x 2 x 27 x bytes per
population (each population read once and written once, AA pattern), divided
by the sweep kernel's mean duration from HIP events on the lattice's stream.

cpu_baseline: the FP64 C/OpenMP oracle (same algorithm, reference data flow) on
the host, on the same workload cut to a 1536x1024x8 span slice.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
BYTES = {"fp64": 8, "fp32": 4, "fp16": 2}
METRIC = "MLUPS & %HBM roofline, FP64 vs FP16-store; cases/hour + p50/p99 latency"


def gpu_run(case, storage, steps, warmup, profile_steps):
    import numpy as np
    import torch
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = case.build()
    s = Solver(section, flow, grid, storage=storage, plan=plan)
    cells = int(np.prod(plan.shape))
    fluid_cells = int(s._fluid.sum())
    s.advance(warmup, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = s.advance(steps, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sweep_ms, bnd_ms = s.lattice.run_profiled(profile_steps, s.steps, 1)
    s.steps += profile_steps
    s.lattice.read_history()
    sweep_avg = sweep_ms / profile_steps / 1e3
    traffic = fluid_cells * 2 * 27 * BYTES[storage]
    achieved = traffic / sweep_avg / 1e9
    out = {"storage": storage, "mlups": cells * steps / dt / 1e6, "ms_per_step": dt / steps * 1e3,
           "sweep_ms": sweep_avg * 1e3, "boundary_ms": bnd_ms / profile_steps,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                        "algorithmic_bytes_per_launch": traffic},
           "cells": cells, "fluid_cells": fluid_cells,
           "loads": np.array([[r["cl"], r["cd"], r["cm"]] for r in recs]),
           "diag": s.diagnostics()}
    s.lattice.close()
    del s
    torch.cuda.empty_cache()
    return out


def cpu_baseline(case, seconds):
    """FP64 oracle on the host: same case, span cut to 8 cells, timed for ~`seconds`."""
    from oracle import oracle as O
    sub = case.with_shape((case.shape[0], case.shape[1], 8))
    section, flow, grid, plan = sub.build()
    o = O.from_case(section, flow, grid, plan)
    cells = plan.shape[0] * plan.shape[1] * plan.shape[2]
    for _ in range(2):
        o.step(sample=True)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(sample=True)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": cells * n / dt / 1e6, "unit": "MLUPS", "cores": O.Oracle.threads(), "kind": "port",
            "sample": f"FP64 C/OpenMP oracle, {plan.shape[0]}x{plan.shape[1]}x{plan.shape[2]} span slice of the "
                      f"same case, 2 warm-up + {n} timed steps with loads every step ({dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=("c128", "c32"), default="c128")
    ap.add_argument("--storages", default="fp16,fp64,fp32")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs a case on a single MI355X; "
              "multi-GPU runs are not supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    from aerokinetic_amd import cases
    case = cases.NACA0012_C128_1536x1024x64 if args.workload == "c128" else cases.NACA0012_C32_384x256x8
    storages = [s.strip() for s in args.storages.split(",") if s.strip()]
    results = {}
    for st in storages:
        print(f"[bench] {case.name} storage={st}", file=sys.stderr, flush=True)
        results[st] = gpu_run(case, st, args.steps, args.warmup, args.profile_steps)
        r = results[st]
        print(f"[bench]   {r['mlups']:.0f} MLUPS, sweep {r['sweep_ms']:.3f} ms, "
              f"{r['roofline']['achieved']:.0f} GB/s ({100 * r['roofline']['frac']:.1f}% of HBM peak)",
              file=sys.stderr, flush=True)
    head = results[storages[0]]
    import numpy as np
    extra = {}
    for st, r in results.items():
        e = {k: r[k] for k in ("mlups", "ms_per_step", "sweep_ms", "boundary_ms", "roofline")}
        if "fp64" in results and st != "fp64":
            a, b = r["loads"], results["fp64"]["loads"]
            e["max_abs_dcoef_vs_fp64"] = dict(zip(("cl", "cd", "cm"), map(float, np.abs(a - b).max(axis=0))))
            e["mean_coef_rel_err_vs_fp64"] = dict(zip(("cl", "cd", "cm"), map(
                float, np.abs(a.mean(axis=0) - b.mean(axis=0)) / np.abs(b.mean(axis=0)))))
        e["local_mach_max"] = r["diag"]["local_mach_max"]
        extra[st] = e
    cpu = cpu_baseline(case, args.cpu_seconds) if args.cpu_seconds > 0 else None
    line = {"metric": METRIC, "value": head["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if storages[0] == "fp64" else "f32", "data": "synthetic",
            "config": {"workload": case.name + ", characteristic farfield, Bouzidi walls, loads+guards every step",
                       "storage": storages[0], "cells": head["cells"], "fluid_cells": head["fluid_cells"],
                       "parallelism": "single GPU"},
            "roofline": head["roofline"], "cpu_baseline": cpu, "by_storage": extra}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
