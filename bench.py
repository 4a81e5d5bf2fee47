"""Throughput of the D3Q27 cumulant lattice-Boltzmann step on one MI355X.

Headline workload (the two-level configuration): NACA0012, chord 64 fine cells
(32 coarse) at 6 deg, Re 3e4, Ma 0.1; coarse lattice 256x192x8 (legacy farfield,
z periodic) with a fine patch of 160x96x16 fine cells (80x48x8 coarse) sub-cycled
twice per coarse step; Bouzidi walls on the fine level; Cl/Cd/Cm recorded every
coarse step.  One step = one coarse step: donor moments, coarse sweep, coarse face
flux and farfield, two fine substeps (prolongation in time, fine sweep, fine face
flux, walls + loads), reflux and restriction.  Populations are resident in HBM
before timing.  MLUPS counts coarse cells + 2 x fine cells per coarse step (the
reference's own two-level MLUPS convention).  `value` is the FP64-storage run; the
FP32-storage run of the same case is reported beside it, with its loads compared
to FP64.

roofline: the dominant kernel (largest total time inside the step) is the fine
sweep; its algorithmic bytes per launch = updated fine cells x 2 x 27 x bytes per
population (one read + one write per population, AA pattern), divided by its mean
duration from HIP events recorded on the stream it is launched on.

exterior_farfield: the plain 384x256x8 NACA0012 case (chord 32, 4 deg, Re 1e4, Ma 0.1)
with the exterior farfield (a 386x258x8 padded lattice whose ring is rebuilt every step):
step rate and the split of a step into sweep, boundary and ring phases.

uniform_sweep: the same step on a uniform 1536x1024x64 lattice (100.7M cells,
NACA0012 chord 128 at 8 deg, Re 1e5, characteristic farfield) in FP64, FP32 and
FP16 shifted storage -- the HBM-roofline view of the fused AA sweep at full size.

copy: the library's streaming device copy of 4 GiB (attainable HBM read+write
bandwidth on this GPU), reported beside the vendor peak.

traffic: HBM bytes per launch of the dominant kernel from the memory counters
(FETCH_SIZE, WRITE_SIZE per dispatch), each scaled by what the counter reads for the
streaming copy of a known size; measured separately (the counters serialise the
GPU) and recorded below with the command that produced them.

cpu_baseline: the FP64 oracle (C/OpenMP lattices + NumPy transfer, the
reference's data flow) on the host, same two-level case, a bounded number of
coarse steps, all host cores; beside it the uniform 384x256x8 oracle step on all
cores and on one core.
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
BYTES = {"fp64": 8, "fp32": 4, "fp16": 2}

# HBM bytes per launch from the memory counters (rocprofv3 --pmc FETCH_SIZE, then
# WRITE_SIZE, KiB per dispatch; FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, the factors that
# map the counters onto the 8 GiB read + 8 GiB written by the streaming copy), for the
# kernels whose roofline is reported: (workload, storage, kernel) -> bytes.
MEASURED_TRAFFIC = {  # profiles/hbm_traffic.md (mean over both AA parities)
    ('two-level', 'fp64', 'fine k_sweep'): 117721287,
    ('two-level', 'fp64', 'coarse k_sweep'): 173270528,
    ('two-level', 'fp32', 'fine k_sweep'): 62851843,
    ('two-level', 'fp32', 'coarse k_sweep'): 87715949,
    ('uniform-100m', 'fp16', 'k_sweep'): 11123233305,
    ('uniform-100m', 'fp32', 'k_sweep'): 22186486482,
    ('uniform-100m', 'fp64', 'k_sweep'): 44633303245,
}
METRIC = "MLUPS & %HBM roofline, FP64 vs FP16-store; cases/hour + p50/p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def roofline(bytes_per_launch, seconds_per_launch, traffic_key=None, copy_gbs=None):
    achieved = bytes_per_launch / seconds_per_launch / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": achieved / PEAK_HBM_GBS, "traffic": MEASURED_TRAFFIC.get(traffic_key),
           "algorithmic_bytes_per_launch": bytes_per_launch, "launch_ms": seconds_per_launch * 1e3}
    if copy_gbs:
        out["frac_of_copy"] = achieved / copy_gbs
    return out


def refined_fields(s):
    """rho and u of both levels (coarse cells, then the padded fine lattice without its solid
    cells), flattened into one set for the field L2 against the FP64-storage run."""
    import numpy as np
    from aerokinetic_amd.d3q27 import C
    rhos, us, masks = [], [], []
    for level in (0, 1):
        g = s.populations(level)
        rho = 1.0 + g.sum(axis=-1)
        u = (g @ C.astype(float)) / rho[..., None]
        mask = np.ones(rho.shape, dtype=bool)
        if level == 1:  # the owned part of the padded fine lattice holds the solid cells
            mask[s.pad:-s.pad, s.pad:-s.pad] &= ~s._fine_solid_xy[:, :, None]
        rhos.append(rho.ravel()); us.append(u.reshape(-1, 3)); masks.append(mask.ravel())
    return np.concatenate(rhos), np.concatenate(us), np.concatenate(masks)


def refined_gpu(case, storage, steps, warmup, profile_steps, copy_gbs=None):
    import numpy as np
    import torch
    from aerokinetic_amd.refined import RefinedSolver
    project, settings, plan = case.build()
    s = RefinedSolver(project, settings, plan=plan, storage=storage)
    fine_owned = 8 * plan.box.shape[0] * plan.box.shape[1] * plan.box.shape[2]
    coarse = int(np.prod(plan.shape))
    updates = coarse + 2 * fine_owned
    s.advance(warmup, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = s.advance(steps, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fields = refined_fields(s)
    tot, cs, fs = s.run_profiled(profile_steps)
    fnx, fny = s.fine_shape[0] + 4, s.fine_shape[1] + 4
    fine_cells = (fnx * fny - int(s._fine_solid_xy.sum())) * s.fine_shape[2]
    b = BYTES[storage]
    fine_sweep = fs / (2 * profile_steps) / 1e3
    coarse_sweep = cs / profile_steps / 1e3
    out = {"storage": storage, "mlups": updates * steps / dt / 1e6, "ms_per_step": dt / steps * 1e3,
           "updates_per_step": updates, "coarse_cells": coarse, "fine_owned_cells": fine_owned,
           "event_ms_per_step": tot / profile_steps, "coarse_sweep_ms": coarse_sweep * 1e3,
           "fine_sweep_ms": fine_sweep * 1e3,
           "fine_sweep_cells": fine_cells,
           "roofline": roofline(fine_cells * 2 * 27 * b, fine_sweep, ("two-level", storage, "fine k_sweep"), copy_gbs),
           "coarse_sweep_roofline": roofline(coarse * 2 * 27 * b, coarse_sweep, ("two-level", storage, "coarse k_sweep"),
                                             copy_gbs),
           "loads": np.array([[r["cl"], r["cd"], r["cm"]] for r in recs]),
           "fields": fields,
           "diag": s.diagnostics()}
    s.close()
    torch.cuda.empty_cache()
    return out


def uniform_gpu(case, storage, steps, warmup, profile_steps, copy_gbs=None):
    import numpy as np
    import torch
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = case.build()
    s = Solver(section, flow, grid, storage=storage, plan=plan)
    cells = int(np.prod(plan.shape))
    fluid = int(s._fluid.sum())
    s.advance(warmup, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = s.advance(steps, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rho, u = s.lattice.macroscopic()
    sweep_ms, bnd_ms = s.lattice.run_profiled(profile_steps, s.steps, 1)
    s.steps += profile_steps
    s.lattice.read_history()
    sweep = sweep_ms / profile_steps / 1e3
    out = {"storage": storage, "mlups": cells * steps / dt / 1e6, "ms_per_step": dt / steps * 1e3,
           "sweep_ms": sweep * 1e3, "boundary_ms": bnd_ms / profile_steps,
           "roofline": roofline(fluid * 2 * 27 * BYTES[storage], sweep, ("uniform-100m", storage, "k_sweep"), copy_gbs),
           "cells": cells, "steps_compared": warmup + steps,
           "loads": np.array([[r["cl"], r["cd"], r["cm"]] for r in recs]),
           "fields": (rho, u, s._fluid)}
    s.lattice.close()
    del s
    torch.cuda.empty_cache()
    return out


def exterior_gpu(case, storage, steps, warmup, profile_steps, copy_gbs=None):
    """The plain uniform case with the exterior farfield: step rate over the physical cells,
    and the split of a step into the sweep, the boundary phase and the rest (the exterior
    ring's band, rebuild, predictor stages and accumulators, plus the step's bookkeeping)."""
    import numpy as np
    import torch
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = case.build()
    s = Solver(section, flow, grid, storage=storage, plan=plan)
    cells = int(np.prod(plan.shape))
    X, Y, Z = s.storage_shape
    ring = (2 * X + 2 * (Y - 2)) * Z
    swept = int((~s._solid_xy).sum()) * Z + (X * Y - plan.shape[0] * plan.shape[1]) * Z - ring
    s.advance(warmup, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = s.advance(steps, 1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rho, u = s.lattice.macroscopic()
    rho, u = rho[s._core], u[s._core]  # the physical box
    sweep_ms, bnd_ms, tot_ms = s.lattice.run_profiled(profile_steps, s.steps, 1, total=True)
    s.steps += profile_steps
    s.lattice.read_history()
    sweep = sweep_ms / profile_steps / 1e3
    out = {"storage": storage, "mlups": cells * steps / dt / 1e6, "ms_per_step": dt / steps * 1e3,
           "loads": np.array([[r["cl"], r["cd"], r["cm"]] for r in recs]), "fields": (rho, u, s._fluid),
           "storage_shape": list(s.storage_shape), "ring_cells": ring,
           "sweep_ms": sweep * 1e3, "boundary_ms": bnd_ms / profile_steps,
           "ring_and_bookkeeping_ms": (tot_ms - sweep_ms - bnd_ms) / profile_steps,
           "step_ms_direct": tot_ms / profile_steps,
           "sweep_roofline": roofline(swept * 2 * 27 * BYTES[storage], sweep, None, copy_gbs),
           "last_sample": {k: recs[-1][k] for k in ("step", "cl", "cd", "cm", "exterior_source_mass_lattice",
                                                     "wall_mass_exchange_lattice")}}
    s.lattice.close()
    del s
    torch.cuda.empty_cache()
    return out


def compare_loads(results, ref="fp64"):
    """Loads (and, where recorded, rho/u fields) of each storage against the FP64-storage run."""
    import numpy as np
    out = {}
    for st, r in results.items():
        if st == ref or ref not in results:
            continue
        a, b = r["loads"], results[ref]["loads"]
        ma, mb = a.mean(axis=0), b.mean(axis=0)
        out[st] = {"max_abs_dcoef_vs_fp64": dict(zip(("cl", "cd", "cm"), map(float, np.abs(a - b).max(axis=0)))),
                   "mean_coef_vs_fp64": {k: [float(ma[i]), float(mb[i])] for i, k in enumerate(("cl", "cd", "cm"))},
                   "mean_coef_rel_err": {k: float(abs(ma[i] - mb[i]) / abs(mb[i])) for i, k in enumerate(("cl", "cd"))}}
        if "fields" in r and "fields" in results[ref]:
            (ra, ua, fl), (rb, ub, _) = r["fields"], results[ref]["fields"]
            dr, du = (ra - rb)[fl], (ua - ub)[fl]
            out[st]["field_l2_vs_fp64"] = {
                "rho_rel": float(np.linalg.norm(dr) / np.linalg.norm(rb[fl])),
                "rho_minus_1_rel": float(np.linalg.norm(dr) / np.linalg.norm(rb[fl] - 1.0)),
                "u_rel": float(np.linalg.norm(du) / np.linalg.norm(ub[fl])),
                "rho_max_abs": float(np.abs(dr).max()), "u_max_abs": float(np.abs(du).max())}
    return out


def cpu_uniform_oracle(seconds, threads):
    """The uniform 384x256x8 oracle step (loads every step) on `threads` host cores."""
    from oracle import oracle as O
    from aerokinetic_amd import cases
    section, flow, grid, plan = cases.NACA0012_C32_384x256x8.build()
    O.Oracle.set_threads(threads)
    o = O.from_case(section, flow, grid, plan)
    cells = plan.shape[0] * plan.shape[1] * plan.shape[2]
    for _ in range(2):
        o.step(sample=True)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.step(sample=True)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": cells * n / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"uniform 384x256x8 FP64 oracle, {n} steps with loads ({dt:.1f} s)"}


def cpu_baseline_refined(case, seconds, threads):
    from oracle import oracle as O
    from oracle.refined_oracle import RefinedOracle
    O.Oracle.set_threads(threads)
    project, settings, plan = case.build()
    ro = RefinedOracle(project, settings, plan)
    updates = plan.shape[0] * plan.shape[1] * plan.shape[2] + 2 * 8 * plan.box.shape[0] * plan.box.shape[1] * plan.box.shape[2]
    ro.step()
    n, t0 = 0, time.perf_counter()
    while True:
        ro.step()
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": updates * n / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (C/OpenMP lattices, NumPy transfer) on the same two-level case, "
                      f"1 warm-up + {n} timed coarse steps ({dt:.1f} s)"}


def serving_gpu(args):
    """Continuous batching (FP16 storage) vs one-at-a-time FP64 solves of the same requests."""
    from aerokinetic_amd import serving as S
    reqs = S.polar_requests(args.serve_cases, seed=2026, rate=args.serve_rate, chord=(16, 24), span=(8, 8),
                            reynolds=(1e4, 3e4))
    rule = dict(chunk=args.serve_chunk, window=args.serve_window, tol=0.005, max_steps=args.serve_max_steps)
    out = {"workload": (f"{len(reqs)} two-level NACA 4-digit requests (seed 2026: camber 0-4 %, thickness 9-15 %, "
                        f"alpha -4..12 deg, Re 1e4-3e4 log-uniform, coarse chord 16-24 cells, 12c x 8c x 8 cells, "
                        f"fine patch 2:1), Poisson arrivals {args.serve_rate}/s; retire when the "
                        f"{args.serve_window}-step Cl window varies < 0.5 % or at {args.serve_max_steps} steps"),
           "rule": rule}
    for st in [s for s in args.serve_storages.split(",") if s]:
        log(f"serving batched {st}")
        rep = S.Server(storage=st, **rule).serve(reqs)
        out[f"batched_{st}"] = rep.summary()
        log(f"  {out[f'batched_{st}']['cases_per_hour']:.0f} cases/h, p50 {out[f'batched_{st}']['latency_p50_s']:.2f} s, "
            f"p99 {out[f'batched_{st}']['latency_p99_s']:.2f} s, peak members {rep.peak_members}")
    log("serving sequential fp64")
    seq = S.serve_sequential(reqs, storage="fp64", **rule)
    out["sequential_fp64"] = seq.summary()
    log(f"  {out['sequential_fp64']['cases_per_hour']:.0f} cases/h, p50 {out['sequential_fp64']['latency_p50_s']:.2f} s")
    for k in list(out):
        if k.startswith("batched_"):
            out[f"{k}_vs_sequential_fp64"] = out[k]["cases_per_hour"] / out["sequential_fp64"]["cases_per_hour"]
    return out


def main():
    faulthandler.enable()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--storages", default="fp64,fp32")
    ap.add_argument("--profile-steps", type=int, default=40)
    ap.add_argument("--uniform-storages", default="fp64,fp32,fp16", help="empty to skip the 100M-cell sweep")
    ap.add_argument("--uniform-steps", type=int, default=100)
    ap.add_argument("--exterior-storages", default="fp64,fp16", help="empty to skip the exterior-farfield case")
    ap.add_argument("--exterior-steps", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-cases", type=int, default=48, help="0 skips the serving measurement")
    ap.add_argument("--serve-rate", type=float, default=64.0)
    ap.add_argument("--serve-storages", default="fp16,fp32")
    ap.add_argument("--serve-chunk", type=int, default=200)
    ap.add_argument("--serve-window", type=int, default=400)
    ap.add_argument("--serve-max-steps", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: a case runs on a single MI355X here; "
              "multi-GPU runs are not supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    from aerokinetic_amd import cases
    from aerokinetic_amd.profile_step import copy_bandwidth
    copy_ms, copy_gbs = copy_bandwidth(4.0, 10)
    log(f"streaming copy 4 GiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s read+write")
    case = cases.NACA0012_REFINED_256x192x8
    storages = [s for s in args.storages.split(",") if s]
    results = {}
    for st in storages:
        log(f"two-level {st}")
        results[st] = refined_gpu(case, st, args.steps, args.warmup, args.profile_steps, copy_gbs)
        r = results[st]
        log(f"  {r['mlups']:.0f} MLUPS, {r['ms_per_step']:.3f} ms/step, fine sweep {r['fine_sweep_ms']:.4f} ms "
            f"({100 * r['roofline']['frac']:.1f}% HBM), coarse sweep {r['coarse_sweep_ms']:.4f} ms")
    uni = {}
    ucase = cases.NACA0012_C128_1536x1024x64
    for st in [s for s in args.uniform_storages.split(",") if s]:
        log(f"uniform 100M {st}")
        uni[st] = uniform_gpu(ucase, st, args.uniform_steps, 10, 20, copy_gbs)
        r = uni[st]
        log(f"  {r['mlups']:.0f} MLUPS, sweep {r['sweep_ms']:.3f} ms ({100 * r['roofline']['frac']:.1f}% HBM)")
    uni_cmp = compare_loads(uni)
    ext = {}
    for st in [s for s in args.exterior_storages.split(",") if s]:
        log(f"exterior farfield 384x256x8 {st}")
        ext[st] = exterior_gpu(cases.NACA0012_C32_384x256x8_EXTERIOR, st, args.exterior_steps, 100, 50, copy_gbs)
        r = ext[st]
        log(f"  {r['mlups']:.0f} MLUPS, {r['ms_per_step']:.4f} ms/step; sweep {r['sweep_ms']:.4f} ms, boundary "
            f"{r['boundary_ms']:.4f} ms, ring + bookkeeping {r['ring_and_bookkeeping_ms']:.4f} ms")
    serving = serving_gpu(args) if args.serve_cases > 0 else None
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cpu = None
    if args.cpu_seconds > 0:
        log(f"cpu baseline on {cores} cores")
        cpu = cpu_baseline_refined(case, args.cpu_seconds, cores)
        cpu["uniform_384x256x8"] = {"all_cores": cpu_uniform_oracle(args.cpu_seconds / 3, cores),
                                    "one_core": cpu_uniform_oracle(args.cpu_seconds / 3, 1)}
    head = results[storages[0]]
    strip = lambda r: {k: v for k, v in r.items() if k not in ("loads", "diag", "fields")}
    line = {"metric": METRIC, "value": head["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if storages[0] == "fp64" else "f32", "data": "synthetic",
            "config": {"workload": case.name + ", legacy coarse farfield, fine Bouzidi walls, loads every coarse step",
                       "storage": storages[0], "updates_per_step": head["updates_per_step"],
                       "parallelism": "single GPU", "step_graphs": True},
            "roofline": head["roofline"], "cpu_baseline": cpu,
            "copy": {"bytes": 4 * 2**30, "ms": copy_ms, "gbs_read_write": copy_gbs},
            "by_storage": {st: strip(r) for st, r in results.items()},
            "loads_vs_fp64": compare_loads(results),
            "uniform_sweep": {"workload": ucase.name + ", characteristic farfield, loads every step",
                              "by_storage": {st: strip(r) for st, r in uni.items()},
                              "vs_fp64": uni_cmp},
            "exterior_farfield": {"workload": cases.NACA0012_C32_384x256x8_EXTERIOR.name + " (exterior_regularized_fd_v1; "
                                              "MLUPS over the physical cells)",
                                  "by_storage": {st: strip(r) for st, r in ext.items()},
                                  "vs_fp64": compare_loads(ext)},
            "serving": serving}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
