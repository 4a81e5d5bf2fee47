"""Two-level 2:1 refined solver on one MI355X: the drop-in `RefinedSolver`.

Public surface of the reference (`src/aerokinetic/refined_solver.py:115-379`,
`refined_run.py:19-157`): `RefinedSolver(project, refinement=RefinementSettings(),
progress=None)` with `step() -> record` (step, t_star, cl, cd, cm[, time_s]),
`diagnostics()` and `run(output, callback=None, stop_event=None)` writing
`history.csv` (step,t_star,cl,cd,cm,time_s), `metadata.json`, `summary.json`.

The coarse lattice (legacy farfield, z periodic) and the fine lattice (box plus two
donor layers in x and y, z periodic, Bouzidi walls) are native AA lattices; the
whole coarse step -- donors, sub-cycled fine substeps, face fluxes, reflux,
restriction and loads -- runs in the HIP library without host round trips.
"""
from __future__ import annotations

from dataclasses import asdict
import csv
import ctypes
import json
import math
import time
from pathlib import Path

import numpy as np

from . import _lib
from ._io import write_json
from .project import Project
from .refined_plan import RefinedPlan, RefinementSettings, load_refined_project, refined_project_dict, virtual_core
from .solver import NumericalFailure, seed_parameters
from .walls import labels, wall_links

_vp, _dp, _ip = _lib._vp, _lib._dp, _lib._ip


def _bind():
    return _lib.load()


class TransferError(ValueError):
    pass


def host_setup(project: Project, plan, pad=2):
    """Host-side preparation of a two-level case (no GPU work): contours, level masks, the
    fine wall links and the seeded-field parameters.  `RefinedSolver(..., host=...)` takes it,
    so it can be computed on a worker thread while the GPU steps other cases."""
    section, flow = project.section, project.flow
    contours = section.normalized(flow.alpha_deg)
    box = plan.box
    nx, ny, _ = plan.shape
    fine_shape = tuple(2 * n for n in box.shape)
    fnx, fny = fine_shape[0] + 2 * pad, fine_shape[1] + 2 * pad
    # cell masks: coarse covered cells are computed donors (2); fine solid (1), donor layers (2)
    cmask = np.zeros((nx, ny), dtype=np.uint8)
    cmask[box.lower[0]:box.upper[0], box.lower[1]:box.upper[1]] = 2
    dxf = plan.fine_level.dx
    origin = (plan.origin[0] + box.lower[0] * plan.coarse_level.dx - pad * dxf,
              plan.origin[1] + box.lower[1] * plan.coarse_level.dx - pad * dxf)
    lab = labels(contours, (fnx, fny), origin, dxf)[1:-1, 1:-1]
    solid = lab > 0
    core = np.repeat(np.repeat(virtual_core(box.shape)[:, :, 0], 2, 0), 2, 1)
    owned_solid = solid[pad:-pad, pad:-pad]
    if np.any(solid & ~np.pad(core, pad, constant_values=False)):
        raise ValueError("A solid intersects the fluid donor band")
    for k, e in enumerate(contours):
        if not (lab == k + 1).any():
            raise ValueError(f"Element {e.name} has no fine solid cells; increase resolution")
    fmask = np.full((fnx, fny), 2, dtype=np.uint8)
    fmask[pad:-pad, pad:-pad] = np.where(owned_solid, 1, 0)
    links, fallback = wall_links(contours, (fnx, fny), origin, dxf)

    class _P:  # uniform-plan view (shape, dx, span) used by the seeded field
        pass
    q = _P()
    q.shape, q.dx, q.span = plan.shape, plan.coarse_level.dx, plan.span
    return {"contours": contours, "fine_shape": fine_shape, "cmask": cmask, "fmask": fmask,
            "owned_solid": owned_solid, "fine_origin_padded": origin, "links": links, "fallback": fallback,
            "seed": seed_parameters(section, flow, q)}


class RefinedSolver:
    HISTORY_COLUMNS = ["step", "t_star", "cl", "cd", "cm", "time_s"]
    pad = 2

    def __init__(self, project: Project, refinement=RefinementSettings(), progress=None, *, storage="fp64",
                 device=0, base_plan=None, plan=None, fp16_scale=0.0, graphs=True, host=None):
        import torch
        if not torch.cuda.is_available():
            raise _lib.NativeLibraryError("no HIP device visible to PyTorch-ROCm")
        self.project, self.refinement = project, refinement
        self.section, self.flow, self.grid = project.section, project.flow, project.grid
        self.plan = plan or RefinedPlan.create(project, refinement, base=base_plan)
        if self.plan.estimated_memory_gib > self.grid.memory_limit_gib:
            raise MemoryError(f"Estimated {self.plan.estimated_memory_gib:.2f} GiB exceeds the explicit limit "
                              f"{self.grid.memory_limit_gib:.2f} GiB")
        if storage not in _lib.STORE:
            raise ValueError("storage must be fp64, fp32 or fp16")
        self.storage = storage
        hs = host or host_setup(project, self.plan, self.pad)
        self.contours = hs["contours"]
        self.L = _bind()
        p, box = self.plan, self.plan.box
        nx, ny, nz = p.shape
        self.fine_shape = hs["fine_shape"]
        h = _vp()
        lo = np.array(box.lower, dtype=np.int32)
        hi = np.array(box.upper, dtype=np.int32)
        lev = np.array([p.coarse_level.dx, p.coarse_level.dt, p.coarse_level.tau_shear, p.coarse_level.tau_bulk])
        stream = torch.cuda.current_stream(torch.device("cuda", device))
        _lib.check(self.L.ak_refined_create(ctypes.byref(h), device, nx, ny, nz, _lib.iptr(lo), _lib.iptr(hi),
                                            _lib.STORE[storage], _lib.dptr(lev), self.flow.lattice_speed, fp16_scale,
                                            4096, _vp(stream.cuda_stream)), None, "ak_refined_create")
        self.h = h
        _lib.check(self.L.ak_refined_set_graph_mode(h, 1 if graphs else 0), None, "ak_refined_set_graph_mode")
        self._levels = [self.L.ak_refined_level(h, 0), self.L.ak_refined_level(h, 1)]
        self._bufs = []
        for lvl in self._levels:
            n = int(self.L.ak_population_bytes(lvl))
            buf = torch.empty(n + 128, dtype=torch.uint8, device=torch.device("cuda", device))
            off = (-buf.data_ptr()) % 128
            _lib.check(self.L.ak_bind_populations(lvl, _vp(buf.data_ptr() + off), n), lvl, "ak_bind_populations")
            self._bufs.append(buf)
        _lib.check(self.L.ak_set_solid(self._levels[0], _lib.u8ptr(hs["cmask"])), self._levels[0], "ak_set_solid")
        self.fine_origin_padded = hs["fine_origin_padded"]
        _lib.check(self.L.ak_set_solid(self._levels[1], _lib.u8ptr(hs["fmask"])), self._levels[1], "ak_set_solid")
        self._fine_solid_xy = hs["owned_solid"]
        links = hs["links"]
        self.links = links
        self.wall_fallback_links = hs["fallback"] * 2 * nz
        args = [_lib.iptr(np.ascontiguousarray(links[k], dtype=np.int32)) for k in ("x", "y", "d")]
        q = np.ascontiguousarray(links["q"])
        el = np.ascontiguousarray(links["element"], dtype=np.int32)
        _lib.check(self.L.ak_refined_set_walls(h, len(q), *args, _lib.dptr(q), _lib.iptr(el), len(self.contours),
                                               self.fine_origin_padded[0], self.fine_origin_padded[1]), None,
                   "ak_refined_set_walls")
        self._keep = (links, q, el)
        self.hist_width = int(self.L.ak_history_width(self._levels[0]))
        self._seed = hs["seed"]
        self.initialize_flow()

    def close(self):
        if getattr(self, "h", None):
            self.L.ak_refined_destroy(self.h)
            self.h = None
            self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def case_fingerprint(self):
        import hashlib
        data = {"project_case": self.project.case_fingerprint, "refinement": asdict(self.refinement)}
        return hashlib.sha256(json.dumps(data, sort_keys=True, separators=(",", ":")).encode()).hexdigest()

    def initialize_flow(self):
        sp = self._seed
        arr = np.array([sp["U"], sp["amp"], sp["phase_y"], sp["phase_z"], sp["ky"], sp["kz"], sp["lex"], sp["ley"],
                        self.plan.origin[0], self.plan.origin[1]])
        _lib.check(self.L.ak_refined_init_seeded(self.h, _lib.dptr(arr)), None, "ak_refined_init_seeded")
        self._check_transfer()
        self.steps = 0
        self.last_loads = None

    def plan_uniform(self):
        """Uniform-plan view (shape, dx, span) used by the seeded field."""
        class _P:
            pass
        q = _P()
        q.shape, q.dx, q.span = self.plan.shape, self.plan.coarse_level.dx, self.plan.span
        return q

    def _check_transfer(self):
        e = self.L.ak_refined_transfer_errors(self.h)
        if e < 0:
            raise _lib.HipError(_lib.last_error(None))
        if e:
            what = [m for b, m in ((1, "nonpositive density"), (2, "negative physical populations"),
                                   (4, "local lattice Mach > 0.3"), (8, "non-finite values")) if e & b]
            raise TransferError("Transfer failed: " + ", ".join(what))

    def _record(self, row):
        step = int(row[0])
        rec = {"step": step, "t_star": (step - 0.5) * self.plan.coarse_level.dt,
               "cl": float(row[1]), "cd": float(row[2]), "cm": float(row[3]), "wall_mass_exchange": float(row[8])}
        if self.plan.dt_seconds is not None:
            rec["time_s"] = (step - 0.5) * self.plan.dt_seconds
        return rec

    def _read(self):
        out = np.empty((4096, self.hist_width))
        n = _lib.check(self.L.ak_read_history(self._levels[0], _lib.dptr(out), 4096), self._levels[0], "ak_read_history")
        return out[:n]

    def advance(self, nsteps, sample_every=1):
        if nsteps <= 0:
            return []
        recs = []
        done = 0
        while done < nsteps:
            chunk = min(nsteps - done, 4000 * max(sample_every, 1))
            _lib.check(self.L.ak_refined_run(self.h, chunk, self.steps, sample_every), None, "ak_refined_run")
            self.steps += chunk
            done += chunk
            if sample_every > 0:
                recs.extend(self._record(r) for r in self._read())
        self._check_transfer()
        return recs

    def run_profiled(self, nsteps, sample_every=1):
        """Advance with HIP events; returns (total, coarse sweep, fine sweeps) milliseconds."""
        out = np.zeros(4)
        _lib.check(self.L.ak_refined_run_profiled(self.h, int(nsteps), self.steps, int(sample_every), _lib.dptr(out)),
                   None, "ak_refined_run_profiled")
        self.steps += nsteps
        self._read()
        self._check_transfer()
        return float(out[0]), float(out[1]), float(out[2])

    def step(self):
        """One coarse step (two fine substeps); returns its load record (refined_solver.py:311-345)."""
        rec = self.advance(1, 1)[-1]
        self.last_loads = rec
        return rec

    def populations(self, level):
        """Pre-collision shifted populations of a level (0 coarse, 1 padded fine), FP64."""
        lvl = self._levels[level]
        nx, ny, nz = (self.plan.shape if level == 0 else
                      (self.fine_shape[0] + 2 * self.pad, self.fine_shape[1] + 2 * self.pad, self.fine_shape[2]))
        g = np.zeros((nx, ny, nz, 27))
        _lib.check(self.L.ak_refined_get_populations(self.h, level, _lib.dptr(g)), None, "ak_refined_get_populations")
        return g

    def diagnostics(self):
        mach, dens = 0.0, 0.0
        for lvl in self._levels:
            out = np.zeros(5)
            _lib.check(self.L.ak_diagnostics(lvl, _lib.dptr(out)), lvl, "ak_diagnostics")
            if out[3]:
                raise NumericalFailure("Non-finite fluid state; the run is invalid")
            mach, dens = max(mach, out[0]), max(dens, out[1])
        if mach > 0.3 or dens > 0.1:
            raise RuntimeError(f"Refined low-Mach state invalid: max Ma={mach:.5g}, density deviation={dens:.5g}")
        return {"local_mach_max": float(mach), "density_error_max": float(dens),
                "wall_fallback_links_including_donors": self.wall_fallback_links}

    @property
    def residual_dt(self):
        """Coarse convective time per coarse step (the residual's time unit)."""
        return self.plan.coarse_level.dt

    def residual_norms(self, reset=False, dt=1.0):
        """Volume-weighted L2 and L-infinity of d(rho, j/U)/dt over the owned fluid cells of both
        levels (pre-collision state; convergence.py:55-96)."""
        from .convergence import combine_levels
        parts = []
        for lvl, dx in zip(self._levels, (self.plan.coarse_level.dx, self.plan.fine_level.dx)):
            out = np.zeros(9)
            _lib.check(self.L.ak_residual(lvl, 1, self.flow.lattice_speed, float(dt), 1 if reset else 0,
                                          _lib.dptr(out)), lvl, "ak_residual")
            parts.append((out, dx ** 3))
        return (np.zeros(4), np.zeros(4)) if reset else combine_levels(parts)

    def save_checkpoint(self, path, context=None):
        """Restartable device state of both levels + run context (refined_checkpoint.py:54-142)."""
        from .checkpoint import save_checkpoint
        return save_checkpoint(self, path, context)

    def run(self, output, callback=None, stop_event=None, resume=False, residual_callback=None):
        """Run project.run.duration_convective coarse time (refined_run.py:19-157); with resume=True
        continue an interrupted run in `output` from the checkpoint this solver was restored from."""
        from .convergence import RunControl
        from .timing import duration_steps
        from ._io import file_hash
        from . import __version__
        import os
        settings = self.project.run
        output = Path(output)
        history = output / "history.csv"
        if not resume and output.exists() and any(output.iterdir()):
            raise FileExistsError("Use a new/empty output directory to avoid mixing histories")
        output.mkdir(parents=True, exist_ok=True)
        end_step = self.steps + duration_steps(settings.duration_convective, self.plan.coarse_level.dt)
        rows = []
        if resume:
            old = json.loads((output / "metadata.json").read_text(encoding="utf-8"))
            context = getattr(self, "_checkpoint_context", {})
            if (old.get("case_fingerprint") != self.case_fingerprint
                    or context.get("case_fingerprint") != self.case_fingerprint):
                raise ValueError("Resume requires the original project, settings and checkpoint")
            if old["status"] == "completed":
                raise ValueError("This run has already completed")
            end_step, prefix = context["end_step"], context["history_bytes"]
            if file_hash(history, prefix) != context["history_sha256"]:
                raise ValueError("History before the checkpoint has changed; refusing to mix results")
            if history.stat().st_size > prefix:
                import shutil
                import uuid
                recovery = output / "recovery" / uuid.uuid4().hex
                recovery.mkdir(parents=True)
                shutil.copy2(history, recovery / "history_before_resume.csv")
                with history.open("r+b") as handle:
                    handle.truncate(prefix)
            with history.open(newline="", encoding="utf-8") as handle:
                for row in csv.DictReader(handle):
                    rows.append({"step": int(row["step"]), "t_star": float(row["t_star"]),
                                 **{k: float(row[k]) for k in ("cl", "cd", "cm")}})
        else:
            write_json(output / "project.json", refined_project_dict(self.project, self.refinement))
        control = RunControl(self, settings, output, resume=resume, callback=residual_callback)
        metadata = {"version": __version__, "refined_plan": json.loads(json.dumps(self.plan.as_dict(), default=list)),
                    "case_fingerprint": self.case_fingerprint, "status": "running", "end_step": end_step,
                    "resume_supported": True, "device": {"backend": "hip-gfx950", "storage": self.storage}}
        write_json(output / "metadata.json", metadata)
        state, reason, start = "completed", "duration_reached", time.monotonic()
        try:
            with history.open("a" if resume else "w", newline="", encoding="utf-8") as handle:
                writer = csv.DictWriter(handle, self.HISTORY_COLUMNS)
                if not resume:
                    writer.writeheader()

                def checkpoint():
                    handle.flush()
                    os.fsync(handle.fileno())
                    context = {"case_fingerprint": self.case_fingerprint, "end_step": end_step,
                               "history_bytes": history.stat().st_size, "history_sha256": file_hash(history)}
                    context.update(control.checkpoint_context())
                    self.save_checkpoint(output / "checkpoint.npz", context)

                se, ce, ke = settings.sample_every, settings.checkpoint_every, settings.check_every
                while self.steps < end_step:
                    if stop_event is not None and stop_event.is_set():
                        state, reason = "stopped", "user_stopped"
                        break
                    nxt = min(end_step, (self.steps // se + 1) * se if callback else self.steps + 64 * se)
                    for every in (ce, ke):
                        if every:
                            nxt = min(nxt, (self.steps // every + 1) * every)
                    recs = self.advance(nxt - self.steps, se)
                    stop = False
                    for rec in recs:
                        if not all(math.isfinite(rec[k]) for k in ("cl", "cd", "cm")):
                            raise NumericalFailure("Non-finite loads; the run is invalid")
                        writer.writerow({k: rec.get(k, "") for k in self.HISTORY_COLUMNS})
                        rows.append(rec)
                        if callback is not None and callback(self, rec) is False:
                            state, reason, stop = "stopped", "user_stopped", True
                            break
                    handle.flush()
                    if ke and self.steps % ke == 0:
                        self.diagnostics()
                        control.observe()
                    if stop:
                        break
                    selected = control.stop_reason()
                    if selected:
                        reason = selected
                        if reason == "wall_time_limit":
                            state = "stopped"
                        break
                    if ce and self.steps % ce == 0 and self.steps < end_step:
                        checkpoint()
                self.diagnostics()
                checkpoint()
        except BaseException as error:
            state, reason = "failed", "failed"
            metadata["failure"] = f"{type(error).__name__}: {error}"
            raise
        finally:
            sel = np.array([[r["cl"], r["cd"], r["cm"]] for r in rows if r["t_star"] >= settings.average_after])
            ok = len(sel) and state != "failed"
            summary = {"status": state, "steps": self.steps, "t_star": self.steps * self.plan.coarse_level.dt,
                       "wall_clock_seconds": time.monotonic() - start, "samples": len(rows),
                       "mean": dict(zip(("cl", "cd", "cm"), map(float, sel.mean(axis=0)))) if ok else None,
                       "rms_fluctuation": dict(zip(("cl", "cd", "cm"), map(float, sel.std(axis=0)))) if ok else None}
            summary.update(control.summary(reason))
            control.close()
            metadata["termination_reason"] = reason
            metadata["status"] = state
            write_json(output / "metadata.json", metadata)
            write_json(output / "summary.json", summary)
        return summary


def load(path, **kwargs):
    project, settings = load_refined_project(path)
    return RefinedSolver(project, settings, **kwargs)
