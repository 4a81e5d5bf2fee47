"""Continuous-batching service for two-level airfoil cases on one GPU.

Requests (airfoil geometry, angle, Reynolds number, grid) arrive over time.  The
service keeps every admitted case resident in HBM and steps all of them together in
one `RefinedBatch`:

* **Admission.** A request is admitted once it has arrived, a slot is free and its
  footprint (population bytes per cell x cells on both levels, plus the transfer
  buffers) fits the memory budget.  The host part of a case (geometry, masks, wall
  links) is prepared on worker threads while the GPU keeps stepping the others; the
  device part (buffers, tables, initial field) is created on the serving thread, which
  makes every GPU call, and the case joins the batch at the next chunk boundary.
* **Stepping.** The batch advances in chunks of coarse steps, with loads recorded every
  coarse step.  Members of different sizes and ages share each batched sweep launch,
  and their transfer kernels overlap.
* **Retirement.** A case retires when its lift coefficient over the last `window`
  coarse steps varies by less than `tol` (max - min relative to max(|mean|, cl_floor)),
  when it reaches `max_steps`, or when its transfer guards fire.  Its slot and memory
  are then free for the next request.

`serve_sequential` runs the same requests one at a time with the plain single-case
solver, as the baseline.  The report gives cases per hour and latency percentiles
(arrival to retirement).
"""
from __future__ import annotations

from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import asdict, dataclass, field
import math
import time

import numpy as np

BYTES = {"fp64": 8, "fp32": 4, "fp16": 2}
U_LATTICE = 0.1 / math.sqrt(3.0)  # Ma 0.1


@dataclass(frozen=True)
class CaseRequest:
    """One airfoil case: NACA 4-digit main element (real-valued camber, position, thickness),
    optional plain flap, angle, Reynolds number and coarse grid (cells per chord, span cells)."""
    id: int
    arrival: float
    camber: float
    position: float
    thickness: float
    alpha_deg: float
    reynolds: float
    chord_cells: int
    span_cells: int
    flap_deflection_deg: float | None = None
    flap_gap: float = 0.02
    flap_overlap: float = 0.02
    flap_chord: float = 0.25


def polar_requests(n, seed=2026, rate=None, chord=(32, 96), span=(8, 32), reynolds=(1e4, 1e5),
                   alpha=(-4.0, 12.0), camber=(0.0, 0.04), thickness=(0.09, 0.15), flap=None):
    """Seeded requests of the batched polar sweep: camber, thickness and angle uniform,
    Reynolds number log-uniform, chord and span cells uniform integers; `rate` (requests per
    second) gives Poisson arrivals, None makes every request available at t = 0.  flap =
    (min, max) deflection in degrees adds a 25 %-chord flap with 1-3 % chord gap and overlap."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        if rate:
            t += float(rng.exponential(1.0 / rate))
        cam = float(rng.uniform(*camber))
        thk = float(rng.uniform(*thickness))
        a = float(rng.uniform(*alpha))
        re = float(np.exp(rng.uniform(np.log(reynolds[0]), np.log(reynolds[1]))))
        c = int(rng.integers(chord[0], chord[1] + 1))
        nz = int(rng.integers(span[0], span[1] + 1))
        kw = {}
        if flap:
            kw = dict(flap_deflection_deg=float(rng.uniform(*flap)), flap_gap=float(rng.uniform(0.01, 0.03)),
                      flap_overlap=float(rng.uniform(0.01, 0.03)))
        out.append(CaseRequest(i, t, cam, 0.4, thk, a, re, c, nz, **kw))
    return out


def case_cells(req: CaseRequest):
    """Lattice cells of a request's two-level case: coarse cells plus fine cells of the box."""
    plan = project_of(req)[2]
    return int(np.prod(plan.shape)) + 8 * int(np.prod(plan.box.shape))


def serving_requests(n, seed=2026, rate=None, cells=(2e6, 2e8), span_ratio=0.25, reynolds=(1e4, 1e5),
                     alpha=(-4.0, 12.0), camber=(0.0, 0.04), thickness=(0.09, 0.18), flap=(0.0, 30.0)):
    """Seeded requests of the serving workload: a NACA 4-digit main element (camber and
    thickness uniform) with a 25 %-chord flap (deflection uniform in `flap`, gap and overlap
    1-3 % chord), angle uniform, Reynolds number log-uniform, and a lattice size log-uniform in
    `cells` (coarse + fine cells): the coarse cells per chord are the integer whose case comes
    closest to the drawn size, with span = max(8, span_ratio x chord) cells."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        if rate:
            t += float(rng.exponential(1.0 / rate))
        target = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        kw = dict(camber=float(rng.uniform(*camber)), position=0.4, thickness=float(rng.uniform(*thickness)),
                  alpha_deg=float(rng.uniform(*alpha)),
                  reynolds=float(np.exp(rng.uniform(np.log(reynolds[0]), np.log(reynolds[1])))),
                  flap_deflection_deg=float(rng.uniform(*flap)), flap_gap=float(rng.uniform(0.01, 0.03)),
                  flap_overlap=float(rng.uniform(0.01, 0.03)))
        # the case size grows like the chord cubed (~28 c^3 for this domain and span ratio)
        c0 = max(16, int(round((target / 28.0) ** (1.0 / 3.0))))
        best = None
        for c in range(max(16, c0 - 3), c0 + 4):
            req = CaseRequest(i, t, chord_cells=c, span_cells=max(8, int(round(span_ratio * c))), **kw)
            err = abs(math.log(case_cells(req) / target))
            if best is None or err < best[0]:
                best = (err, req)
        out.append(best[1])
    return out


def section_of(req: CaseRequest):
    """Main element (chord 1) and, if requested, a flap below and behind its trailing edge."""
    from .geometry import Contour, Section, naca4_profile
    main = naca4_profile(req.camber, req.position, req.thickness, 161, True, "main")
    if req.flap_deflection_deg is None:
        return Section((main,), 1.0, (0.25, 0.0))
    cf, d = req.flap_chord, math.radians(req.flap_deflection_deg)
    flap = naca4_profile(0.0, 0.4, 0.15, 81, True, "flap").xy * cf
    rot = np.array([[math.cos(d), math.sin(d)], [-math.sin(d), math.cos(d)]])  # trailing edge down
    flap = flap @ rot.T
    lower = main.xy[int(np.argmin(main.xy[:, 0])):]  # leading edge to trailing edge along the lower side
    x_le = 1.0 - req.flap_overlap
    y_main = float(np.interp(x_le, lower[:, 0], lower[:, 1]))
    top = flap[:, 1].max()
    flap = flap + np.array([x_le, y_main - req.flap_gap - top])
    return Section((main, Contour(flap, "flap")), 1.0, (0.25, 0.0))


def project_of(req: CaseRequest):
    """(project, refinement settings, plan): domain 12 chords x 8 chords, legacy farfield."""
    from .config import Flow, Grid
    from .project import Project, RunSettings
    from .refined_plan import RefinedPlan, RefinementSettings
    section = section_of(req)
    flow = Flow(reynolds=req.reynolds, alpha_deg=req.alpha_deg, lattice_speed=U_LATTICE, perturbation=1e-4)
    grid = Grid(cells_per_chord=req.chord_cells, upstream_chords=4.0, downstream_chords=7.0, transverse_chords=3.5,
                span_chords=req.span_cells / req.chord_cells, memory_limit_gib=1e6, farfield="legacy")
    project = Project(section, flow, grid, RunSettings(), f"request {req.id}")
    settings = RefinementSettings()
    return project, settings, RefinedPlan.create(project, settings)


def footprint_bytes(plan, storage):
    """Device bytes of one two-level case: populations of both lattices (with ghost rows and
    pitch) and the transfer, face and history buffers."""
    nx, ny, nz = plan.shape
    fx, fy, fz = (2 * n for n in plan.box.shape)
    cells = (nx + 2) * (ny + 2) * nz + (fx + 6) * (fy + 6) * fz
    bx, by, bz = plan.box.shape
    transfer = 8 * 10 * ((bx + 6) * (by + 6) * (bz + 6) * 3 + (2 * bx + 4) * (2 * by + 4) * (bz + 6) * 2)
    return int(cells * 27 * BYTES[storage] * 1.1 + transfer + (8 << 20))


@dataclass
class CaseResult:
    id: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    reason: str
    cl: float
    cd: float
    cm: float
    cl_window_variation: float
    cells: int

    @property
    def latency(self):
        return self.finished - self.arrival


@dataclass
class ServeReport:
    storage: str
    mode: str
    results: list = field(default_factory=list)
    makespan_s: float = 0.0
    peak_members: int = 0
    gpu_steps: int = 0
    host_s: dict = field(default_factory=dict)  # serving-thread time by activity
    tile: tuple | None = None  # sweep tile of the batch (None: library defaults)
    tile_timings: dict = field(default_factory=dict)  # tile -> ms per batch step, when tuned

    def summary(self):
        lat = np.array([r.latency for r in self.results]) if self.results else np.zeros(1)
        done = len(self.results)
        return {"mode": self.mode, "storage": self.storage, "cases": done,
                "cases_per_hour": done / self.makespan_s * 3600.0 if self.makespan_s > 0 else 0.0,
                "makespan_s": self.makespan_s, "latency_p50_s": float(np.percentile(lat, 50)),
                "latency_p99_s": float(np.percentile(lat, 99)), "latency_mean_s": float(lat.mean()),
                "peak_members": self.peak_members, "coarse_steps": int(sum(r.steps for r in self.results)),
                "host_s": {k: round(v, 3) for k, v in self.host_s.items()},
                "reasons": {k: sum(r.reason == k for r in self.results) for k in ("converged", "max_steps", "failed")},
                "tile": list(self.tile) if self.tile else None,
                "tile_ms_per_batch_step": {f"{t[0]}x{t[1]}": round(v, 4) for t, v in self.tile_timings.items()}}


def _window_variation(cl, window, cl_floor):
    w = np.asarray(cl[-window:])
    return float((w.max() - w.min()) / max(abs(w.mean()), cl_floor))


class Server:
    """Continuous-batching service (see the module docstring)."""

    def __init__(self, storage="fp16", device=0, memory_bytes=None, max_members=64, chunk=200, window=2000,
                 tol=0.005, cl_floor=0.1, max_steps=20000, builders=4, graphs=True, tile=None, tune_members=4,
                 tune_steps=10):
        if chunk < 4:
            raise ValueError("chunk must be at least 4 coarse steps")
        self.storage, self.device = storage, device
        self.memory_bytes = memory_bytes
        self.max_members, self.chunk = max_members, chunk
        self.window, self.tol, self.cl_floor, self.max_steps = window, tol, cl_floor, max_steps
        self.builders, self.graphs = builders, graphs
        # sweep tile of the batch: library default (None), fixed (bx, by), or "auto": the fastest
        # of batch.BATCH_TILES, measured on real steps once tune_members cases have joined
        self.tile, self.tune_members, self.tune_steps = tile, tune_members, tune_steps

    def _budget(self):
        if self.memory_bytes:
            return self.memory_bytes
        import torch
        free, _ = torch.cuda.mem_get_info(self.device)
        return int(free * 0.9)

    def _prepare(self, req):
        """Host-only part of a case (worker thread): project, plan, footprint, geometry."""
        from .refined import host_setup
        bundle = project_of(req)
        return bundle, footprint_bytes(bundle[2], self.storage), host_setup(bundle[0], bundle[2])

    def _build(self, bundle, host):
        """Device part of a case (serving thread)."""
        from .refined import RefinedSolver
        project, settings, plan = bundle
        return RefinedSolver(project, settings, plan=plan, storage=self.storage, device=self.device,
                             graphs=self.graphs, host=host)

    def _retire_reason(self, st):
        steps = len(st["cl"])
        if steps >= self.window and _window_variation(st["cl"], self.window, self.cl_floor) <= self.tol:
            return "converged"
        if steps >= self.max_steps:
            return "max_steps"
        return None

    def _result(self, st, now, reason):
        req, cl = st["req"], st["cl"]
        w = min(len(cl), self.window) or 1
        loads = np.concatenate(st["loads"])[-w:] if st["loads"] else np.zeros((1, 3))
        return CaseResult(req.id, req.arrival, st["admitted"], now, len(cl), reason, *map(float, loads.mean(axis=0)),
                          _window_variation(cl, w, self.cl_floor) if cl else float("nan"), st["cells"])

    def serve(self, requests, log=None):
        """Serve `requests` (arrival times relative to the start); returns a ServeReport."""
        from .batch import RefinedBatch
        log = log or (lambda m: None)
        report = ServeReport(self.storage, "batched")
        budget = self._budget()
        queue = deque(sorted(requests, key=lambda r: r.arrival))
        batch = RefinedBatch(self.storage, self.device, graphs=self.graphs)
        if self.tile not in (None, "auto"):
            batch.tile = tuple(self.tile)
        building, ready, active = {}, deque(), {}  # prepare futures; prepared; in the batch
        reserved = 0
        t0 = time.perf_counter()
        staged = []  # cases whose device part was built while the last chunk ran
        clock = {"build": 0.0, "launch": 0.0, "collect": 0.0, "retire": 0.0, "admit": 0.0, "idle": 0.0, "tune": 0.0}

        def timed(key, fn, *a, **k):
            t = time.perf_counter()
            out = fn(*a, **k)
            clock[key] += time.perf_counter() - t
            return out

        with ThreadPoolExecutor(self.builders) as pool:
            while queue or building or ready or active or staged:
                if ready and not (active or staged or building) and reserved + ready[0][2] > budget:
                    raise MemoryError(f"request {ready[0][0].id} needs {ready[0][2]} bytes, over the budget {budget}")
                now = time.perf_counter() - t0
                # arrived requests are prepared on the host ahead of admission
                while queue and queue[0].arrival <= now and len(building) + len(ready) < 2 * self.builders:
                    req = queue.popleft()
                    building[pool.submit(self._prepare, req)] = req
                for fut in [f for f in building if f.done()]:
                    req = building.pop(fut)
                    ready.append((req, *fut.result()))
                # the batch is idle here: admitted cases join it
                for solver, st in staged:
                    timed("admit", batch.add, solver)
                    st["admitted"] = time.perf_counter() - t0
                    active[solver] = st
                    log(f"admit {st['req'].id} ({st['cells']} cells) t={st['admitted']:.2f}s members={len(batch)}")
                staged = []
                report.peak_members = max(report.peak_members, len(batch))
                if self.tile == "auto" and batch.tile is None and len(active) >= self.tune_members:
                    _, report.tile_timings, rows = timed("tune", batch.tune, self.tune_steps)
                    report.gpu_steps += (self.tune_steps + 2) * len(report.tile_timings)
                    for solver, r in rows.items():
                        active[solver]["cl"].extend(r[:, 1].tolist())
                        active[solver]["loads"].append(r[:, 1:4].copy())
                if active:
                    timed("launch", batch.launch, self.chunk, sample=True)
                # admission: a free slot and the footprint within the budget; the device part is
                # built while the chunk runs
                while ready and len(active) + len(staged) < self.max_members and reserved + ready[0][2] <= budget:
                    req, bundle, need, host = ready.popleft()
                    reserved += need
                    plan = bundle[2]
                    cells = int(np.prod(plan.shape)) + 8 * int(np.prod(plan.box.shape))
                    staged.append((timed("build", self._build, bundle, host),
                                   {"req": req, "need": need, "admitted": None, "cl": [], "loads": [], "cells": cells}))
                if not active:
                    if staged:
                        continue
                    if building:
                        timed("idle", wait, list(building), timeout=1.0, return_when=FIRST_COMPLETED)
                    elif queue:
                        timed("idle", time.sleep, max(0.0, min(queue[0].arrival - (time.perf_counter() - t0), 1.0)))
                    continue
                rows, failed = timed("collect", batch.collect, raw=True)
                report.gpu_steps += self.chunk
                now = time.perf_counter() - t0
                for solver, r in rows.items():
                    st = active[solver]
                    st["cl"].extend(r[:, 1].tolist())
                    st["loads"].append(r[:, 1:4].copy())
                    reason = "failed" if solver in failed else self._retire_reason(st)
                    if reason:
                        timed("retire", batch.remove, solver)
                        timed("retire", solver.close)
                        del active[solver]
                        reserved -= st["need"]
                        report.results.append(self._result(st, now, reason))
                        log(f"retire {st['req'].id} {reason} after {len(st['cl'])} steps, t={now:.2f}s")
        report.makespan_s = time.perf_counter() - t0
        report.host_s = clock
        report.tile = batch.tile
        batch.close()
        return report


def serve_sequential(requests, storage="fp64", device=0, chunk=200, window=2000, tol=0.005, cl_floor=0.1,
                     max_steps=20000, log=None):
    """Baseline: the same requests one at a time, in arrival order, each with the plain
    single-case solver (graph replay of its own step) and the same retirement rule."""
    from .refined import RefinedSolver, TransferError
    log = log or (lambda m: None)
    report = ServeReport(storage, "sequential")
    rule = Server(storage, device, chunk=chunk, window=window, tol=tol, cl_floor=cl_floor, max_steps=max_steps)
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        now = time.perf_counter() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        project, settings, plan = project_of(req)
        solver = RefinedSolver(project, settings, plan=plan, storage=storage, device=device)
        cells = int(np.prod(plan.shape)) + 8 * int(np.prod(plan.box.shape))
        st = {"req": req, "admitted": time.perf_counter() - t0, "cl": [], "loads": [], "cells": cells}
        report.peak_members = 1
        while True:
            try:
                rs = solver.advance(chunk, 1)
                reason = None
            except TransferError:
                rs, reason = [], "failed"
            report.gpu_steps += chunk
            st["cl"].extend(r["cl"] for r in rs)
            st["loads"].append(np.array([[r["cl"], r["cd"], r["cm"]] for r in rs]).reshape(-1, 3))
            reason = reason or rule._retire_reason(st)
            if reason:
                break
        solver.close()
        now = time.perf_counter() - t0
        report.results.append(rule._result(st, now, reason))
        log(f"sequential {req.id} {reason} after {len(st['cl'])} steps, t={now:.2f}s")
    report.makespan_s = time.perf_counter() - t0
    return report


def report_dict(report):
    out = report.summary()
    out["cases_detail"] = [dict(asdict(r), latency=r.latency) for r in report.results]
    return out
