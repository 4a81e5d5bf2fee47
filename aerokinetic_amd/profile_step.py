"""Run a fixed number of steps of a named workload (a short driver for profilers).

    python -m aerokinetic_amd.profile_step --workload uniform-100m --storage fp16 --steps 10
    python -m aerokinetic_amd.profile_step --workload two-level --storage fp64 --steps 40
    python -m aerokinetic_amd.profile_step --workload copy --copy-gib 4 --steps 10

Workloads: `uniform-c32` (384x256x8), `exterior-c32` (the same box with the exterior
farfield), `uniform-100m` (1536x1024x64) and
`two-level` (coarse 256x192x8 + fine 160x96x16), all NACA0012 (see
aerokinetic_amd.cases); `copy`, a device-to-device copy of a buffer of known size
(the calibration kernel for memory counters); `serve`, the batched service over
seeded two-level requests (--steps = the per-case step cap); `batch`, a fixed batch of
--cases seeded two-level serving requests (coarse chord 16-24) stepped together; `slab-y` (1536x128x64) and
`slab-x` (192x1024x64), one GPU's share of the 1536x1024x64 lattice at 8 GPUs as a plain
fluid slab (sweep kernel only); `traffic`, the streaming copy of a known size and one
sweep of each AA form of the 1536x1024x64 lattice per storage, for memory-counter runs.  Each workload prints one
line per phase with HIP-event timings of the lattice's stream, so a profiler's
per-kernel averages can be set beside them.
"""
from __future__ import annotations

import argparse
import sys
import time


def _log(msg):
    print(f"[profile_step] {msg}", flush=True)


def run_uniform(case, storage, steps, warmup, sample_every, graphs=True, launch=None):
    import numpy as np
    import torch
    from .autotune import apply_launch
    from .solver import Solver
    section, flow, grid, plan = case.build()
    s = Solver(section, flow, grid, storage=storage, plan=plan, graphs=graphs)
    if launch:
        apply_launch(s.lattice, launch)
        _log(f"sweep launch (even, odd) {launch}")
    fluid = int(s._fluid.sum())
    _log(f"{case.name} storage {storage}: {int(np.prod(plan.shape))} cells, {fluid} fluid")
    s.advance(warmup, sample_every)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.advance(steps, sample_every)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sweep_ms, bnd_ms, tot_ms = s.lattice.run_profiled(steps, s.steps, sample_every, total=True)
    s.steps += steps
    s.lattice.read_history()
    _log(f"steps {steps}: wall {dt / steps * 1e3:.4f} ms/step (graphs {graphs}); direct launches with events: "
         f"sweep {sweep_ms / steps:.4f} ms, boundary {bnd_ms / steps:.4f} ms, step {tot_ms / steps:.4f} ms")
    s.lattice.close()


def run_two_level(case, storage, steps, warmup, sample_every, graphs=True):
    from .refined import RefinedSolver
    project, settings, plan = case.build()
    s = RefinedSolver(project, settings, plan=plan, storage=storage, graphs=graphs)
    _log(f"{case.name} storage {storage}")
    import torch
    s.advance(warmup, sample_every)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.advance(steps, sample_every)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tot, cs, fs = s.run_profiled(steps, sample_every)
    _log(f"steps {steps}: wall {dt / steps * 1e3:.4f} ms/step (graphs {graphs}); direct launches with events "
         f"{tot / steps:.4f} ms/step, coarse sweep {cs / steps:.4f} ms, fine sweep {fs / (2 * steps):.4f} ms "
         f"per substep")
    s.close()


def copy_bandwidth(gib, steps):
    """(ms per copy, GB/s read + write) of the library's streaming copy of `gib` GiB."""
    import torch
    from . import _lib
    L = _lib.load()
    n = int(gib * 2**30) // 8 * 8
    src = torch.ones(n // 8, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream()
    copy = lambda: _lib.check(L.ak_device_copy(_lib._vp(dst.data_ptr()), _lib._vp(src.data_ptr()), n,
                                               _lib._vp(st.cuda_stream)), None, "ak_device_copy")
    copy()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(steps):
        copy()
    b.record(st)
    torch.cuda.synchronize()
    if not torch.equal(dst[:1024], src[:1024]):
        raise RuntimeError("device copy mismatch")
    ms = a.elapsed_time(b) / steps
    return ms, 2 * n / ms / 1e6


def run_copy(gib, steps):
    ms, gbs = copy_bandwidth(gib, steps)
    _log(f"copy {int(gib * 2**30)} bytes: {ms:.4f} ms per copy, {gbs:.1f} GB/s (read + write)")


def _launches(text):
    """{storage: ((bx, by, waves) even, (bx, by, waves) odd)} from 'fp64:256x1c3/0x256c3,fp16:...'."""
    from .autotune import parse_launch
    out = {}
    for item in (text or "").split(","):
        if item:
            st, _, spec = item.partition(":")
            out[st] = parse_launch(spec)
    return out


def run_traffic(storages, copy_gib=4.0, launches=None):
    """The launches whose memory counters calibrate and measure HBM traffic: the streaming copy
    of `copy_gib` GiB twice, then for each storage one sampled step of each AA form of the
    1536x1024x64 lattice (direct launches, so every sweep is one dispatch), with the sweep
    launch of `launches` {storage: (even, odd)} where given."""
    import torch
    from . import cases
    from .autotune import apply_launch
    from .solver import Solver
    ms, gbs = copy_bandwidth(copy_gib, 2)
    _log(f"copy {int(copy_gib * 2**30)} bytes x 2: {ms:.4f} ms per copy")
    case = cases.NACA0012_C128_1536x1024x64
    section, flow, grid, plan = case.build()
    for storage in storages:
        s = Solver(section, flow, grid, storage=storage, plan=plan, graphs=False)
        if launches and storage in launches:
            apply_launch(s.lattice, launches[storage])
        sweep_ms, _ = s.lattice.run_profiled(2, 0, 1)
        s.lattice.read_history()
        _log(f"{case.name} {storage}: 2 sampled steps (one per AA form), sweep {sweep_ms / 2:.4f} ms per launch, "
             f"{int(s._fluid.sum())} fluid cells")
        s.lattice.close()
        del s
        torch.cuda.empty_cache()


def run_serve(storage, cases, rate, chunk, window, max_steps):
    from . import serving as S
    reqs = S.polar_requests(cases, seed=2026, rate=rate, chord=(16, 24), span=(8, 8), reynolds=(1e4, 3e4))
    t0 = time.perf_counter()
    rep = S.Server(storage=storage, chunk=chunk, window=window, max_steps=max_steps).serve(reqs)
    s = rep.summary()
    _log(f"serve {cases} requests {storage}: {time.perf_counter() - t0:.2f} s, {s['cases_per_hour']:.0f} cases/h, "
         f"p50 {s['latency_p50_s']:.2f} s, p99 {s['latency_p99_s']:.2f} s, peak members {s['peak_members']}, "
         f"batch steps {rep.gpu_steps}, serving-thread seconds {s['host_s']}")


def run_batch(storage, ncases, steps, warmup, chunk=50, sweep_grid=0, chord=(16, 24), span=(8, 8),
              reynolds=(1e4, 3e4)):
    """A fixed batch of `ncases` two-level requests (coarse chord, span and Re ranges as given,
    by default 16-24, 8, 1e4-3e4; seed 2026) stepped together in chunks of `chunk` coarse steps;
    wall time per batch step."""
    import torch
    from . import serving as S
    from .batch import RefinedBatch
    from .refined import RefinedSolver
    reqs = S.polar_requests(ncases, seed=2026, chord=chord, span=span, reynolds=reynolds)
    members = []
    for r in reqs:
        project, settings, plan = S.project_of(r)
        members.append(RefinedSolver(project, settings, plan=plan, storage=storage))
    b = RefinedBatch(storage, sweep_grid=sweep_grid)
    for m in members:
        b.add(m)
    done = 0
    while done < warmup:
        b.advance(min(chunk, warmup - done))
        done += min(chunk, warmup - done)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        b.advance(min(chunk, steps - done))
        done += min(chunk, steps - done)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _log(f"batch of {ncases} {storage} (sweep grid {sweep_grid or 'one block per tile'}): "
         f"{dt / steps * 1e3:.4f} ms per batch step over {steps} steps")
    b.close()
    for m in members:
        m.close()


def run_slab(shape, storage, steps, warmup):
    """The sweep of one GPU's share of the 1536x1024x64 lattice at 8 GPUs: a plain fluid
    slab (no body, seeded field, legacy farfield), the sweep kernel timed with its own events."""
    import math
    import numpy as np
    from .lattice import Lattice
    nx, ny, nz = shape
    lat = Lattice(nx, ny, nz, storage=storage, omega_shear=1.99, omega_bulk=1.0)
    lat.set_solid(np.zeros((nx, ny), dtype=np.uint8))
    seed = {"U": 0.0577, "amp": 1e-4, "phase_y": 0.3, "phase_z": 0.1, "ky": 2 * math.pi, "kz": 2 * math.pi,
            "lex": 1.0, "ley": 1.0}
    lat.init_seeded(seed, (0.0, 0.0), 1.0 / 128)
    lat.set_farfield("legacy", 0.0577)
    lat.run_profiled(warmup, 0, 0)
    sweep_ms, _ = lat.run_profiled(steps, warmup, 0)
    ms = sweep_ms / steps
    cells = nx * ny * nz
    gbs = cells * 54 * {"fp64": 8, "fp32": 4, "fp16": 2}[storage] / (ms * 1e-3) / 1e9
    _log(f"slab {nx}x{ny}x{nz} storage {storage}: sweep {ms:.4f} ms, {cells / (ms * 1e3):.0f} MLUPS, "
         f"{gbs:.0f} GB/s ({gbs / 80:.1f} % of 8 TB/s)")
    lat.close()


def main(argv=None):
    from . import cases
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=("uniform-c32", "exterior-c32", "uniform-100m", "two-level", "copy", "serve",
                                           "batch", "slab-y", "slab-x", "traffic"), required=True)
    ap.add_argument("--storages", default="fp64,fp16", help="traffic: storages of the 1536x1024x64 sweeps")
    ap.add_argument("--storage", choices=("fp64", "fp32", "fp16"), default="fp64")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sample-every", type=int, default=1)
    ap.add_argument("--copy-gib", type=float, default=4.0)
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel directly")
    ap.add_argument("--launch", default="", help="uniform and traffic workloads: sweep launch per storage, "
                    "'fp64:BXxBYcW/BXxBYcW,...' (even form / odd form tile and wave cap, as the autotuner picks)")
    ap.add_argument("--cases", type=int, default=24, help="serve / batch: number of requests")
    ap.add_argument("--rate", type=float, default=16.0, help="serve: Poisson arrival rate (1/s)")
    ap.add_argument("--chunk", type=int, default=100, help="serve: coarse steps per batch call")
    ap.add_argument("--sweep-grid", type=int, default=0, help="batch: persistent sweep grid (0: one block per tile)")
    ap.add_argument("--chord", default="16,24", help="batch: coarse chord range in cells (the polar sweep: 32,96)")
    ap.add_argument("--span", default="8,8", help="batch: span range in cells (the polar sweep: 8,32)")
    ap.add_argument("--reynolds", default="1e4,3e4", help="batch: Reynolds number range (the polar sweep: 1e4,1e5)")
    a = ap.parse_args(argv)
    if a.workload == "copy":
        run_copy(a.copy_gib, a.steps)
    elif a.workload == "traffic":
        run_traffic([x for x in a.storages.split(",") if x], a.copy_gib, _launches(a.launch))
    elif a.workload == "batch":
        pair = lambda v, f: tuple(f(x) for x in v.split(","))
        run_batch(a.storage, a.cases, a.steps, a.warmup, sweep_grid=a.sweep_grid, chord=pair(a.chord, int),
                  span=pair(a.span, int), reynolds=pair(a.reynolds, float))
    elif a.workload == "serve":
        run_serve(a.storage, a.cases, a.rate, a.chunk, 400, a.steps)
    elif a.workload in ("slab-y", "slab-x"):
        run_slab((1536, 128, 64) if a.workload == "slab-y" else (192, 1024, 64), a.storage, a.steps, a.warmup)
    elif a.workload == "two-level":
        run_two_level(cases.NACA0012_REFINED_256x192x8, a.storage, a.steps, a.warmup, a.sample_every, not a.no_graphs)
    else:
        case = {"uniform-c32": cases.NACA0012_C32_384x256x8, "exterior-c32": cases.NACA0012_C32_384x256x8_EXTERIOR,
                "uniform-100m": cases.NACA0012_C128_1536x1024x64}[a.workload]
        run_uniform(case, a.storage, a.steps, a.warmup, a.sample_every, not a.no_graphs,
                    _launches(a.launch).get(a.storage))
    return 0


if __name__ == "__main__":
    sys.exit(main())
