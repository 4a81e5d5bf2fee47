"""CPU baseline timing of the FP64 oracle (test and measurement infrastructure only).

    python -m oracle.cpu_baseline --case NACA0012_C128_1536x1024x64 --shape 1536,1024,8 \
        --seconds 15 --threads 16

Runs the oracle's uniform step (the reference's two-array pull data flow, boundary objects
in installation order, loads every step) on a named workload of aerokinetic_amd.cases,
optionally on another lattice shape with the same leading-edge cell, for `warmup` untimed
steps and then until `seconds` have passed (or for exactly `--steps` timed steps, as the
reference's own profiling script does: 16 warm-up and 256 timed steps), and prints one JSON object:
{"value": MLUPS, "cores": threads, "steps": n, "seconds": s, "case": name}.  Started in a
process of its own by the benchmark, with OMP_PROC_BIND / OMP_PLACES set there, so that its
OpenMP threads are pinned without touching the benchmark's own process.
"""
import argparse
import json
import sys
import time


def main(argv=None):
    from aerokinetic_amd import cases
    from oracle import oracle as O
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", required=True)
    ap.add_argument("--shape", default=None)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--min-steps", type=int, default=4)
    ap.add_argument("--steps", type=int, default=0, help="exactly this many timed steps (0: bounded by --seconds)")
    a = ap.parse_args(argv)
    case = getattr(cases, a.case)
    if a.shape:
        case = case.with_shape(tuple(int(v) for v in a.shape.split(",")))
    section, flow, grid, plan = case.build()
    O.Oracle.set_threads(a.threads)
    o = O.from_case(section, flow, grid, plan)
    cells = plan.shape[0] * plan.shape[1] * plan.shape[2]
    for _ in range(a.warmup):
        o.step(sample=True)
    n, t0 = 0, time.perf_counter()
    if a.steps > 0:
        for _ in range(a.steps):
            o.step(sample=True)
        n = a.steps
    else:
        while n < a.min_steps or time.perf_counter() - t0 < a.seconds:
            o.step(sample=True)
            n += 1
    dt = time.perf_counter() - t0
    print(json.dumps({"value": cells * n / dt / 1e6, "cores": a.threads, "steps": n, "warmup": a.warmup,
                      "seconds": dt, "case": case.name, "cells": cells}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
