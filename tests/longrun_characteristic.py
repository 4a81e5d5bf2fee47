"""Long-run parity of the characteristic farfield: the GPU FP64 path against the oracle for
thousands of steps of an unsteady case, until the run fails its guards.

Case: the 1536x1024x64 precision-study case scaled by 1/4 (384x256x16, NACA0012 chord 32 at
8 deg, leading edge at (128, 128), Re 2.5e4 so that the lattice viscosity is the same, Ma 0.1,
characteristic farfield, seeded perturbation), Cl/Cd/Cm every step.  Not collected by
pytest (the oracle needs ~20 minutes of host time for it).

    python tests/longrun_characteristic.py gpu --steps 7000 --out longrun_gpu.npy
    python tests/longrun_characteristic.py oracle --steps 7000 --threads 8 --out longrun_oracle.npy
    python tests/longrun_characteristic.py compare longrun_oracle.npy longrun_gpu.npy

Each run stops at the first step whose guards fail (local Mach > 0.3, max |rho - 1| > 0.1 or a
non-finite state; the oracle checks every 100 steps) and saves the loads of every step.
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def case():
    from aerokinetic_amd import cases
    return cases.UniformCase("naca0012 c32 384x256x16 a8 Re2.5e4 Ma0.1", (384, 256, 16), 32, 8.0, 2.5e4, 0.1,
                             (128, 128))


def guards_fail(mach, derr, finite=True):
    return (not finite) or mach > 0.3 or derr > 0.1


def run_gpu(steps):
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = case().build()
    s = Solver(section, flow, grid, storage="fp64", plan=plan)
    loads, failed = [], None
    while s.steps < steps and failed is None:
        for r in s.advance(min(500, steps - s.steps), 1):
            loads.append((r["cl"], r["cd"], r["cm"]))
            finite = bool(np.isfinite([r["cl"], r["cd"], r["cm"]]).all()) and not r["nonfinite"]
            if failed is None and guards_fail(r["local_mach_max"], r["density_error_max"], finite):
                failed = r["step"]
        print(f"gpu step {s.steps} cl {loads[-1][0]:.6f}", flush=True)
    return np.array(loads), failed


def run_oracle(steps, threads):
    from oracle.oracle import Oracle, from_case
    Oracle.set_threads(threads)
    section, flow, grid, plan = case().build()
    o = from_case(section, flow, grid, plan)
    loads, failed = [], None
    for step in range(1, steps + 1):
        r = o.step(sample=True)
        loads.append((r["cl"], r["cd"], r["cm"]))
        if step % 100 == 0:
            d = o.diagnostics()
            print(f"oracle step {step} cl {r['cl']:.6f} mach {d['local_mach_max']:.4f}", flush=True)
            if guards_fail(d["local_mach_max"], d["density_error_max"], d["finite"]):
                failed = step
                break
    return np.array(loads), failed


def compare(a_path, b_path):
    a, b = np.load(a_path), np.load(b_path)
    n = min(len(a), len(b))
    d = np.abs(a[:n, 0] - b[:n, 0])
    for s in (100, 1000, 2000, 3000, 4000, 5000, 6000):
        if s <= n:
            print(f"step {s}: cl {a[s - 1, 0]:.15f} vs {b[s - 1, 0]:.15f}, |d| {d[s - 1]:.2e}")
    for tol in (1e-12, 1e-9, 1e-6, 1e-3):
        off = np.nonzero(d > tol)[0]
        print(f"first step with |d cl| > {tol:g}: {int(off[0]) + 1 if len(off) else None} (of {n})")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("mode", choices=("gpu", "oracle", "compare"))
    ap.add_argument("files", nargs="*")
    ap.add_argument("--steps", type=int, default=7000)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.mode == "compare":
        compare(*a.files)
        return 0
    loads, failed = run_gpu(a.steps) if a.mode == "gpu" else run_oracle(a.steps, a.threads)
    print(f"{a.mode}: {len(loads)} steps, guards failed at step {failed}", flush=True)
    if a.out:
        np.save(a.out, loads)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
