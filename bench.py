"""Throughput of the D3Q27 cumulant lattice-Boltzmann step on one MI355X.

Headline workload (the precision-study configuration): a uniform 1536x1024x64 lattice
(100,663,296 cells) around a NACA0012 of chord 128 cells at 8 deg, Re 1e5, Ma 0.1, z
periodic, characteristic farfield on the x/y faces, seeded perturbation, Cl/Cd/Cm recorded
every step.  One step = boundary phase (Bouzidi walls + loads, farfield) + the fused AA
stream-collide sweep.  `value` is the FP64-storage MLUPS (cells x steps / s over exactly
--steps timed steps, populations resident in HBM before timing); the same case in FP32 and
FP16 shifted storage (g = f - w, FP32 collision arithmetic) is reported beside it
(`by_storage`), with its loads and rho/u fields compared to FP64.  With --tile auto (the
default) each AA form's sweep launch (tile and cap on resident waves per SIMD) is the
autotuner's pick on the lattice itself (aerokinetic_amd.autotune.tune, reported as
`sweep_launch`); the solver seeds its initial state again after tuning, untimed.

roofline: the dominant kernel is the sweep (~97 % of a step).  Algorithmic bytes per launch
= fluid cells x 2 x 27 x bytes per population (one read + one write per population, AA
pattern in place), divided by its mean duration from HIP events that the sweep launch
records itself on the stream it runs on (both AA forms, direct launches).  `traffic` is
measured in this run: two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over a
child process that runs the streaming copy of 4 GiB and one sweep of each AA form per
storage; each counter is scaled by what it reads for the copy of known size.

two_level: coarse 256x192x8 + fine 160x96x16 around a NACA0012 of chord 64 fine cells at
6 deg, Re 3e4 (one step = one coarse step with both fine substeps and all transfers; MLUPS
counts coarse cells + 2 x fine cells per coarse step), FP64 and FP32 storage.  With
--tile auto (the default) the sweep tile of each level is the autotuner's pick
(aerokinetic_amd.autotune.tune_refined, reported as `tiles`).

exterior_farfield: the plain 384x256x8 NACA0012 case (chord 32, 4 deg, Re 1e4) with the
exterior farfield (a 386x258x8 padded lattice whose ring is rebuilt every step).

serving: continuous batching of two-level requests in FP16 storage against one-at-a-time
FP64 solves of the same requests: seeded Poisson arrivals of a NACA 4-digit main element +
25 %-chord flap (0-30 deg), lattice sizes log-uniform in 2M-200M cells, Re 1e4-1e5,
retirement when the 2000-step Cl window varies < 0.5 % or at the step cap; cases/hour and
p50/p99 latency.  With --tile auto the batch picks its sweep tile on its first real steps
(RefinedBatch.tune; reported as `tile`).

polar_sweep: 64 two-level FP32 cases of uneven size (chord 32-96, span 8-32, 540M cells in
all) stepped as one batch, against the same cases one at a time; MLUPS of both.

copy: the library's streaming device copy of 4 GiB (attainable HBM read+write bandwidth).

cpu_baseline: the FP64 oracle (C/OpenMP, the reference's two-array pull data flow) on the
host cores: a 1536x1024x8 slice of the headline lattice (same airfoil, same farfield) on all
cores, with the plain 384x256x8 case on all cores and on one core beside it.
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
STORE_ID = {"fp64": 0, "fp32": 1, "fp16": 2}

METRIC = "MLUPS & %HBM roofline, FP64 vs FP16-store; cases/hour + p50/p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def roofline(bytes_per_launch, seconds_per_launch, copy_gbs=None, traffic=None):
    achieved = bytes_per_launch / seconds_per_launch / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
           "algorithmic_bytes_per_launch": bytes_per_launch, "launch_ms": seconds_per_launch * 1e3}
    if copy_gbs:
        out["frac_of_copy"] = achieved / copy_gbs
    return out


def _counter_rows(directory):
    """(kernel name, counter value) of every dispatch in a rocprofv3 counter-collection CSV."""
    import csv
    rows = []
    for path in Path(directory).rglob("*counter_collection.csv"):
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
                value = r.get("Counter_Value") or r.get("Counter-Value")
                if value is not None:
                    rows.append((name, float(value)))
    return rows


def measure_traffic(storages, timeout=300, launches=None):
    """HBM bytes per launch of the headline sweep, per storage, from two rocprofv3 counter
    passes (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch) over `profile_step --workload
    traffic` in a child process: the streaming copy of 4 GiB (which reads and writes 4 GiB:
    the calibration) and one sweep of each AA form of the 1536x1024x64 lattice, with the
    sweep launches of the timed run (`launches` {storage: "even/odd"}).  Returns
    {storage: bytes per launch (mean of the two forms), ...} plus the details, or a reason."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    exe = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if Path("/opt/rocm/bin/rocprofv3").exists() else None)
    if exe is None:
        return None, "rocprofv3 not found"
    copy_bytes = 4 * 2**30
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out_dir = tempfile.mkdtemp(prefix=f"ak_pmc_{counter.lower()}_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "-d", out_dir, "-o", "pmc", "--output-format", "csv", "--", sys.executable,
               "-m", "aerokinetic_amd.profile_step", "--workload", "traffic", "--storages", ",".join(storages)]
        if launches:
            cmd += ["--launch", ",".join(f"{st}:{spec}" for st, spec in launches.items())]
        env = dict(os.environ, TMPDIR="/tmp", PYTHONPATH=str(ROOT))
        log(f"memory counters: {counter}")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, timeout=timeout, capture_output=True, text=True)
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 --pmc {counter} timed out"
        rows = _counter_rows(out_dir)
        shutil.rmtree(out_dir, ignore_errors=True)
        if not rows:
            return None, f"rocprofv3 --pmc {counter} gave no counters (rc {r.returncode}): {r.stderr[-300:]}"
        copy = [v for n, v in rows if "k_stream_copy" in n]
        if not copy:
            return None, f"no streaming-copy dispatch in the {counter} pass"
        scale = copy_bytes / (float(np.mean(copy)) * 1024.0)  # bytes per counter KiB
        per[counter] = {"scale": scale, "copy_kib": float(np.mean(copy)),
                        "sweeps": {st: [v * 1024.0 * scale for n, v in rows if f"k_sweep<{STORE_ID[st]}," in n]
                                   for st in storages}}
    out = {}
    for st in storages:
        fetch, write = per["FETCH_SIZE"]["sweeps"][st], per["WRITE_SIZE"]["sweeps"][st]
        if len(fetch) != 2 or len(write) != 2:
            return None, f"expected two {st} sweep dispatches per pass, got {len(fetch)} and {len(write)}"
        out[st] = {"read_bytes": float(np.mean(fetch)), "written_bytes": float(np.mean(write)),
                   "bytes_per_launch": float(np.mean(fetch) + np.mean(write)),
                   "by_form": [float(f + w) for f, w in zip(fetch, write)]}
    out["calibration"] = {"copy_bytes_each_way": copy_bytes,
                          "fetch_size_kib_of_copy": per["FETCH_SIZE"]["copy_kib"],
                          "write_size_kib_of_copy": per["WRITE_SIZE"]["copy_kib"],
                          "fetch_factor": per["FETCH_SIZE"]["scale"], "write_factor": per["WRITE_SIZE"]["scale"]}
    return out, None


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


def refined_gpu(case, storage, steps, warmup, profile_steps, copy_gbs=None, tile=None):
    import numpy as np
    import torch
    from aerokinetic_amd.refined import RefinedSolver
    project, settings, plan = case.build()
    s = RefinedSolver(project, settings, plan=plan, storage=storage, tile=tile)
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
           "roofline": roofline(fine_cells * 2 * 27 * b, fine_sweep, copy_gbs),
           "coarse_sweep_roofline": roofline(coarse * 2 * 27 * b, coarse_sweep, copy_gbs),
           "tiles": {"coarse": list(s.tile[0]), "fine": list(s.tile[1])},
           "tile_ms": {lvl: {f"{t[0]}x{t[1]}": round(v, 5) for t, v in d.items()} for lvl, d in s.tile_timings.items()},
           "loads": np.array([[r["cl"], r["cd"], r["cm"]] for r in recs]),
           "fields": fields,
           "diag": s.diagnostics()}
    s.close()
    torch.cuda.empty_cache()
    return out


def uniform_gpu(case, storage, steps, warmup, profile_steps, copy_gbs=None, tile=None, launch=None):
    import numpy as np
    import torch
    from aerokinetic_amd.autotune import apply_launch
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = case.build()
    t_tune = time.perf_counter()
    s = Solver(section, flow, grid, storage=storage, plan=plan, tile=None if launch else tile)
    if launch:  # a fixed launch per AA form instead of tuning (profiler runs see only these)
        apply_launch(s.lattice, launch)
        s.tile = launch
    t_tune = time.perf_counter() - t_tune
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
           "roofline": roofline(fluid * 2 * 27 * BYTES[storage], sweep, copy_gbs), "fluid_cells": fluid,
           "cells": cells, "steps_compared": warmup + steps,
           "sweep_launch": _launch_report(s.tile), "setup_and_tuning_s": t_tune,
           "launch_spec": _launch_spec(s.tile),
           "placement_sweep_ms": [round(t, 4) for t in s.placement_ms],
           "loads": np.array([[r["cl"], r["cd"], r["cm"]] for r in recs]),
           "fields": (rho, u, s._fluid)}
    s.lattice.close()
    del s
    torch.cuda.empty_cache()
    return out


def _launch_spec(tile):
    """The autotuned launch as profile_step --launch takes it ('even/odd'), or None."""
    from aerokinetic_amd.autotune import format_launch
    if tile is not None and len(tile) == 2 and all(isinstance(t, (tuple, list)) for t in tile):
        return format_launch(tile)
    return None


def _launch_report(tile):
    """Sweep launch of each AA form: tile (bx x by; bx 0 = row-linear) and resident-wave cap."""
    if tile is None:
        return {"tile": "library default (64x4 on whole-wave rows)", "autotuned": False}
    if len(tile) == 2 and all(isinstance(t, (tuple, list)) for t in tile):
        return {form: {"tile": f"{t[0]}x{t[1]}", "waves_per_simd_cap": (t[2] if len(t) > 2 else 0) or None}
                for form, t in zip(("even", "odd"), tile)} | {"autotuned": True}
    return {"tile": f"{tile[0]}x{tile[1]}", "autotuned": False}


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
           "sweep_roofline": roofline(swept * 2 * 27 * BYTES[storage], sweep, copy_gbs),
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


def host_cpu(usable):
    """CPU model, usable cores and the OpenMP placement of the host the baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo", encoding="utf-8"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "usable_cores": usable, "omp_proc_bind": "close", "omp_places": "cores"}


def cpu_uniform_oracle(case_name, seconds, threads, shape=None, warmup=2, steps=0):
    """The FP64 oracle's uniform step (loads every step) of a named workload on `threads` host
    cores, timed in a child process whose OpenMP threads are pinned (OMP_PROC_BIND=close,
    OMP_PLACES=cores): `warmup` untimed steps, then exactly `steps` timed steps, or (steps=0)
    steps until `seconds` have passed (at least 4)."""
    import subprocess
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--case", case_name, "--seconds", str(seconds),
           "--threads", str(threads), "--warmup", str(warmup), "--steps", str(steps)] + \
          (["--shape", ",".join(map(str, shape))] if shape else [])
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads),
               PYTHONPATH=str(ROOT))
    r = subprocess.run(cmd, cwd=str(ROOT), env=env, capture_output=True, text=True, timeout=600)
    if r.returncode:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-500:]}")
    t = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": t["value"], "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{t['case']}: FP64 oracle, {t['warmup']} warm-up + {t['steps']} timed steps with loads every "
                      f"step ({t['seconds']:.1f} s), OpenMP threads pinned (close, cores)"}


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
    """Continuous batching (FP16 storage) vs one-at-a-time FP64 solves of the same requests:
    the serving workload (main element + 25 %-chord flap at 0-30 deg, lattices of 2M-200M
    cells log-uniform, Re 1e4-1e5, Poisson arrivals), a reduced number of requests."""
    from aerokinetic_amd import serving as S
    reqs = S.serving_requests(args.serve_cases, seed=2026, rate=args.serve_rate, cells=(2e6, 2e8), flap=(0.0, 30.0))
    rule = dict(chunk=args.serve_chunk, window=args.serve_window, tol=0.005, max_steps=args.serve_max_steps)
    cells = [S.case_cells(r) for r in reqs]
    out = {"workload": (f"{len(reqs)} two-level requests (seed 2026): NACA 4-digit main element (camber 0-4 %, "
                        f"thickness 9-18 %) + 25 %-chord flap at 0-30 deg, gap/overlap 1-3 % chord, alpha -4..12 deg, "
                        f"Re 1e4-1e5 log-uniform, 12c x 8c domains, lattice sizes log-uniform in 2M-200M cells "
                        f"(drawn: {min(cells) / 1e6:.1f}M-{max(cells) / 1e6:.1f}M, total {sum(cells) / 1e6:.0f}M), "
                        f"Poisson arrivals {args.serve_rate}/s; retire when the {args.serve_window}-step Cl window "
                        f"varies < 0.5 % or at {args.serve_max_steps} coarse steps"),
           "rule": rule, "cells": cells}
    for st in [s for s in args.serve_storages.split(",") if s]:
        log(f"serving batched {st}")
        rep = S.Server(storage=st, tile=args.tile, **rule).serve(reqs, log=log)
        out[f"batched_{st}"] = rep.summary()
        log(f"  {out[f'batched_{st}']['cases_per_hour']:.0f} cases/h, p50 {out[f'batched_{st}']['latency_p50_s']:.2f} s, "
            f"p99 {out[f'batched_{st}']['latency_p99_s']:.2f} s, peak members {rep.peak_members}")
    log("serving sequential fp64")
    seq = S.serve_sequential(reqs, storage="fp64", log=log, **rule)
    out["sequential_fp64"] = seq.summary()
    log(f"  {out['sequential_fp64']['cases_per_hour']:.0f} cases/h, p50 {out['sequential_fp64']['latency_p50_s']:.2f} s")
    for k in list(out):
        if k.startswith("batched_"):
            out[f"{k}_vs_sequential_fp64"] = out[k]["cases_per_hour"] / out["sequential_fp64"]["cases_per_hour"]
    return out


def polar_gpu(args):
    """The batched polar sweep: `--polar-cases` two-level FP32 cases of uneven size (NACA 4-digit,
    camber 0-4 %, thickness 9-15 %, alpha -4..12 deg, Re 1e4-1e5 log-uniform, chord 32-96 coarse
    cells, span 8-32 cells, 12c x 8c domains; seed 2026), each kernel of a batch step one launch
    over all members' concatenated work lists, against the same cases stepped one at a time.
    Both modes run `--polar-steps` coarse steps per case (loads every step) after 10 untimed
    warm-up steps, on the same solver objects (the flow is re-initialised in between)."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import torch
    from aerokinetic_amd import serving as S
    from aerokinetic_amd.batch import RefinedBatch
    from aerokinetic_amd.refined import RefinedSolver, host_setup
    reqs = S.polar_requests(args.polar_cases, seed=2026, chord=(32, 96), span=(8, 32))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(8) as pool:
        bundles = list(pool.map(lambda r: (lambda b: (b, host_setup(b[0], b[2])))(S.project_of(r)), reqs))
    solvers = [RefinedSolver(b[0], b[1], plan=b[2], storage="fp32", host=h) for b, h in bundles]
    build_s = time.perf_counter() - t0
    updates = [int(np.prod(m.plan.shape)) + 2 * 8 * int(np.prod(m.plan.box.shape)) for m in solvers]
    steps = args.polar_steps
    log(f"polar sweep: {len(solvers)} cases, {sum(updates) / 1e6:.0f}M updates per coarse step, built in {build_s:.1f} s")
    # one at a time
    seq = 0.0
    for m in solvers:
        m.advance(10, 1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        m.advance(steps, 1)
        torch.cuda.synchronize()
        seq += time.perf_counter() - t
        m.initialize_flow()
    # all together
    batch = RefinedBatch("fp32")
    for m in solvers:
        batch.add(m)
    batch.advance(10)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, failed = batch.advance(steps)
    torch.cuda.synchronize()
    bat = time.perf_counter() - t
    batch.close()
    for m in solvers:
        m.close()
    torch.cuda.empty_cache()
    total = sum(updates) * steps
    out = {"workload": (f"{len(reqs)} two-level FP32 cases (seed 2026): NACA 4-digit camber 0-4 %, thickness 9-15 %, "
                        f"alpha -4..12 deg, Re 1e4-1e5, chord 32-96 coarse cells, span 8-32 cells, 12c x 8c domains; "
                        f"{steps} coarse steps per case, loads every step"),
           "cases": len(reqs), "updates_per_coarse_step": int(sum(updates)),
           "smallest_largest_updates": [int(min(updates)), int(max(updates))], "host_build_s": build_s,
           "batched_s": bat, "batched_mlups": total / bat / 1e6, "transfer_failures": len(failed),
           "sequential_s": seq, "sequential_mlups": total / seq / 1e6, "batched_vs_sequential": seq / bat}
    log(f"  batched {out['batched_mlups']:.0f} MLUPS ({bat:.2f} s), one at a time {out['sequential_mlups']:.0f} MLUPS "
        f"({seq:.2f} s): {seq / bat:.2f}x")
    return out


def main():
    faulthandler.enable()
    # host cores this process may use, read before any runtime binds a thread
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps of the headline lattice (each storage)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storages", default="fp64,fp32,fp16", help="headline storages; the first is `value`")
    ap.add_argument("--profile-steps", type=int, default=20, help="event-timed steps for the roofline")
    ap.add_argument("--traffic", type=int, default=1, help="0 skips the memory-counter passes")
    ap.add_argument("--two-level-storages", default="fp64,fp32", help="empty to skip the two-level case")
    ap.add_argument("--two-level-steps", type=int, default=2000)
    ap.add_argument("--tile", default="auto", help="sweep launches of the headline lattice (tile and wave cap per AA "
                    "form), the two-level case and the serving batch: 'auto' (autotuned) or 'default'")
    ap.add_argument("--launch", default="", help="headline lattice: a fixed sweep launch per storage instead of "
                    "tuning, 'fp64:BXxBYcW/BXxBYcW,...' (even / odd form tile and wave cap, as `launch_spec` reports)")
    ap.add_argument("--exterior-storages", default="fp64,fp16", help="empty to skip the exterior-farfield case")
    ap.add_argument("--exterior-steps", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-cases", type=int, default=6, help="0 skips the serving measurement")
    ap.add_argument("--polar-cases", type=int, default=64, help="0 skips the batched polar sweep")
    ap.add_argument("--polar-steps", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=1.0)
    ap.add_argument("--serve-storages", default="fp16")
    ap.add_argument("--serve-chunk", type=int, default=250)
    ap.add_argument("--serve-window", type=int, default=2000)
    ap.add_argument("--serve-max-steps", type=int, default=3000)
    args = ap.parse_args()
    args.tile = None if args.tile == "default" else args.tile
    from aerokinetic_amd.profile_step import _launches
    fixed_launch = _launches(args.launch)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: a case runs on a single MI355X here; "
              "multi-GPU runs are not supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    from aerokinetic_amd import cases
    from aerokinetic_amd.profile_step import copy_bandwidth
    copy_ms, copy_gbs = copy_bandwidth(4.0, 10)
    log(f"streaming copy 4 GiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s read+write")
    # headline: the 1536x1024x64 lattice in every storage
    ucase = cases.NACA0012_C128_1536x1024x64
    storages = [s for s in args.storages.split(",") if s]
    uni = {}
    for st in storages:
        log(f"{ucase.name} {st}")
        uni[st] = uniform_gpu(ucase, st, args.steps, args.warmup, args.profile_steps, copy_gbs, tile=args.tile,
                              launch=fixed_launch.get(st))
        r = uni[st]
        log(f"  {r['mlups']:.0f} MLUPS, {r['ms_per_step']:.3f} ms/step, sweep {r['sweep_ms']:.3f} ms "
            f"({100 * r['roofline']['frac']:.1f}% HBM), boundary {r['boundary_ms']:.3f} ms")
    uni_cmp = compare_loads(uni)
    two = {}
    case2 = cases.NACA0012_REFINED_256x192x8
    for st in [s for s in args.two_level_storages.split(",") if s]:
        log(f"two-level {st}")
        two[st] = refined_gpu(case2, st, args.two_level_steps, 500, 40, copy_gbs, tile=args.tile)
        r = two[st]
        log(f"  {r['mlups']:.0f} MLUPS, {r['ms_per_step']:.3f} ms/step, fine sweep {r['fine_sweep_ms']:.4f} ms "
            f"({100 * r['roofline']['frac']:.1f}% HBM), coarse sweep {r['coarse_sweep_ms']:.4f} ms")
    ext = {}
    for st in [s for s in args.exterior_storages.split(",") if s]:
        log(f"exterior farfield 384x256x8 {st}")
        ext[st] = exterior_gpu(cases.NACA0012_C32_384x256x8_EXTERIOR, st, args.exterior_steps, 100, 50, copy_gbs)
        r = ext[st]
        log(f"  {r['mlups']:.0f} MLUPS, {r['ms_per_step']:.4f} ms/step; sweep {r['sweep_ms']:.4f} ms, boundary "
            f"{r['boundary_ms']:.4f} ms, ring + bookkeeping {r['ring_and_bookkeeping_ms']:.4f} ms")
    serving = serving_gpu(args) if args.serve_cases > 0 else None
    polar = polar_gpu(args) if args.polar_cases > 0 else None
    import torch
    torch.cuda.empty_cache()
    launches = {st: r["launch_spec"] for st, r in uni.items() if r.get("launch_spec")}
    traffic, why = (measure_traffic(storages, launches=launches) if args.traffic else (None, "skipped (--traffic 0)"))
    if traffic:
        for st in storages:
            uni[st]["roofline"]["traffic"] = traffic[st]["bytes_per_launch"]
            uni[st]["roofline"]["traffic_detail"] = {k: v for k, v in traffic[st].items() if k != "bytes_per_launch"}
            uni[st]["roofline"]["traffic_calibration"] = traffic["calibration"]
    else:
        log(f"memory counters not measured: {why}")
        for st in storages:
            uni[st]["roofline"]["traffic_note"] = why
    cpu = None
    if args.cpu_seconds > 0:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        cores = min(omp, usable) if omp > 0 else usable
        log(f"cpu baseline on {cores} cores")
        cpu = cpu_uniform_oracle("NACA0012_C128_1536x1024x64", args.cpu_seconds, cores, shape=(1536, 1024, 8))
        cpu["sample"] += " -- a 1536x1024x8 slice of the headline 1536x1024x64 lattice (per-cell work is the same)"
        cpu["host"] = host_cpu(usable)
        # the plain case as the reference's own profiling script times it (16 warm-up + 256 timed
        # steps, loads every step) on all cores; one core is bounded by time instead
        cpu["uniform_384x256x8"] = {"all_cores": cpu_uniform_oracle("NACA0012_C32_384x256x8", 0, cores,
                                                                    warmup=16, steps=256),
                                    "one_core": cpu_uniform_oracle("NACA0012_C32_384x256x8", args.cpu_seconds / 3, 1)}
        if two:
            cpu["two_level"] = cpu_baseline_refined(case2, args.cpu_seconds / 2, cores)
    head = uni[storages[0]]
    strip = lambda r: {k: v for k, v in r.items() if k not in ("loads", "diag", "fields")}
    line = {"metric": METRIC, "value": head["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16-store/f32"}[storages[0]], "data": "synthetic",
            "config": {"workload": ucase.name + " (1536x1024x64 = 100,663,296 cells), characteristic farfield, "
                                               "loads every step",
                       "storage": storages[0], "cells": head["cells"], "parallelism": "single GPU",
                       "step_graphs": True,
                       "sweep_launch": head.get("sweep_launch")},
            "roofline": head["roofline"], "cpu_baseline": cpu,
            "copy": {"bytes": 4 * 2**30, "ms": copy_ms, "gbs_read_write": copy_gbs},
            "by_storage": {st: strip(r) for st, r in uni.items()},
            "vs_fp64": uni_cmp,
            "two_level": {"workload": case2.name + ", legacy coarse farfield, fine Bouzidi walls, loads every coarse "
                                                   "step; MLUPS = (coarse + 2 x fine cells) per coarse step",
                          "by_storage": {st: strip(r) for st, r in two.items()},
                          "loads_vs_fp64": compare_loads(two)},
            "exterior_farfield": {"workload": cases.NACA0012_C32_384x256x8_EXTERIOR.name + " (exterior_regularized_fd_v1; "
                                              "MLUPS over the physical cells)",
                                  "by_storage": {st: strip(r) for st, r in ext.items()},
                                  "vs_fp64": compare_loads(ext)},
            "serving": serving,
            "polar_sweep": polar}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
