"""Two-level 2:1 refined solver on one MI355X: the drop-in `RefinedSolver`.

Public surface of the reference (`src/aerokinetic/refined_solver.py:115-379`,
`refined_run.py:19-157`): `RefinedSolver(project, refinement=RefinementSettings(),
progress=None)` with `step() -> record` (step, t_star, cl, cd, cm[, time_s]),
`diagnostics()` (low-Mach guards and the FP64 conservation budget kept on the device),
`run(output, **kwargs)` writing `history.csv` (step,t_star,cl,cd,cm[,time_s]),
`residuals.csv`, `metadata.json`, `summary.json`, and `from_checkpoint` / `resume_run`.

The coarse lattice (legacy farfield, z periodic) and the fine lattice (box plus two
donor layers in x and y, z periodic, Bouzidi walls) are native AA lattices; the
whole coarse step -- donors, sub-cycled fine substeps, face fluxes, reflux,
restriction and loads -- runs in the HIP library without host round trips.
"""
from __future__ import annotations

from dataclasses import asdict
import ctypes
import json
import math
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


BUDGET_WIDTH = 32  # doubles returned by ak_refined_budget


class RefinedSolver:
    pad = 2

    def __init__(self, project: Project, refinement=RefinementSettings(), progress=None, *, storage="fp64",
                 device=0, base_plan=None, plan=None, fp16_scale=0.0, graphs=True, host=None, tile=None):
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
        # refined_run.py:31-33: time_s only when the physical time step is known
        self.HISTORY_COLUMNS = ["step", "t_star", "cl", "cd", "cm"] + (
            ["time_s"] if self.plan.dt_seconds is not None else [])
        fine_fluid_cells = int((~self._fine_solid_xy).sum()) * self.fine_shape[2]
        coarse_owned_cells = int(np.prod(p.shape)) - int(np.prod(box.shape))
        self._reference_volume = (coarse_owned_cells * p.coarse_level.dx ** 3
                                  + fine_fluid_cells * p.fine_level.dx ** 3)
        self.initialize_flow()
        # sweep tiles of the two levels: library defaults, explicit ((bx, by), (bx, by)), or
        # "auto" (aerokinetic_amd.autotune.tune_refined: fastest measured, state kept)
        self.tile_timings = {}
        if tile == "auto":
            from .autotune import tune_refined
            self.tile, self.tile_timings = tune_refined(self)
        elif tile:
            self.set_tiles(*tile)
        else:
            from .autotune import default_tile
            self.tile = (default_tile(nx), default_tile(self.fine_shape[0] + 2 * self.pad))

    def set_tiles(self, coarse, fine):
        """Sweep tiles (bx, by) of the coarse and fine lattices (bx = 0: row-linear blocks)."""
        for lvl, (bx, by) in zip(self._levels, (coarse, fine)):
            _lib.check(self.L.ak_set_sweep_tile(lvl, int(bx), int(by)), lvl, "ak_set_sweep_tile")
        self.tile = (tuple(coarse), tuple(fine))

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
        """Every recorded history row (read in blocks until the device history is empty)."""
        rows = []
        while True:
            out = np.empty((4096, self.hist_width))
            n = _lib.check(self.L.ak_read_history(self._levels[0], _lib.dptr(out), 4096), self._levels[0],
                           "ak_read_history")
            rows.append(out[:n])
            if n < 4096:
                return np.concatenate(rows)

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

    def budget(self):
        """The device's FP64 budget accumulators (ak_refined_budget layout)."""
        out = np.zeros(BUDGET_WIDTH)
        _lib.check(self.L.ak_refined_budget(self.h, _lib.dptr(out)), None, "ak_refined_budget")
        return out

    @staticmethod
    def _last_budget(b, volume):
        """Budget of the last coarse step from the accumulators (refined_solver.py:330-340):
        mass and momentum residuals per owned volume, wall mass exchange and impulse, and what
        entered through the outer faces."""
        if b[30] <= 0:
            return {}
        before, after = b[21:25], b[25:29]
        wall_impulse, wall_mass, external = b[0:3], b[4], b[5:9]
        mass_residual = after[0] - before[0] - external[0] - wall_mass
        momentum_residual = after[1:] - before[1:] - external[1:] + wall_impulse
        return {"mass_residual_per_volume": float(mass_residual / volume),
                "momentum_residual_per_volume": (momentum_residual / volume).tolist(),
                "wall_mass_exchange": float(wall_mass), "external_mass_in": float(external[0]),
                "wall_impulse": wall_impulse.tolist(), "external_impulse_in": external[1:].tolist()}

    @property
    def last_budget(self):
        """Mass/momentum budget of the last coarse step ({} before the first step)."""
        return self._last_budget(self.budget(), self._reference_volume)

    @property
    def coarse_boundary_calls(self):
        return int(self.budget()[29])

    @property
    def fine_boundary_calls(self):
        return 2 * self.coarse_boundary_calls

    def diagnostics(self):
        """Low-Mach guards over the leaf cells and the conservation budget
        (refined_solver.py:347-361)."""
        mach, dens = 0.0, 0.0
        for lvl in self._levels:
            out = np.zeros(5)
            _lib.check(self.L.ak_diagnostics(lvl, _lib.dptr(out)), lvl, "ak_diagnostics")
            if out[3]:
                raise NumericalFailure("Non-finite fluid state; the run is invalid")
            mach, dens = max(mach, out[0]), max(dens, out[1])
        if mach > 0.3 or dens > 0.1:
            raise RuntimeError(f"Refined low-Mach state invalid: max Ma={mach:.5g}, density deviation={dens:.5g}")
        b = self.budget()
        out = {"local_mach_max": float(mach), "density_error_max": float(dens),
               "wall_fallback_links_including_donors": self.wall_fallback_links,
               "coarse_boundary_calls": int(b[29]), "fine_boundary_calls": 2 * int(b[29]),
               "cumulative_wall_mass_exchange": float(b[9]),
               "cumulative_wall_mass_exchange_over_initial_mass": float(b[9] / b[17]),
               "last_budget": self._last_budget(b, self._reference_volume)}
        if not b[31]:
            out["budget_complete"] = False  # batched steps keep no budget
        return out

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

    @classmethod
    def from_checkpoint(cls, path, progress=None, **kwargs):
        """Rebuild a case from its checkpoint and continue from it (refined_solver.py:367-370)."""
        from .checkpoint import restore_checkpoint
        return restore_checkpoint(cls, path, progress, **kwargs)

    @classmethod
    def resume_run(cls, output, **kwargs):
        """Continue the interrupted run in `output` from its checkpoint (refined_run.py:160-163)."""
        output = Path(output)
        solver = cls.from_checkpoint(output / "checkpoint.npz", progress=kwargs.pop("progress", None))
        return solver.run(output, resume=True, **kwargs)

    def run(self, output, *, callback=None, stop_event=None, resume=False, residual_callback=None):
        """Run project.run.duration_convective of coarse time into a new or empty directory
        (refined_run.py:19-157); resume=True continues an interrupted run in `output` from the
        checkpoint this solver was restored from.  callback(solver, record) may return False
        to stop; history.csv holds every sample_every-th step's loads, residuals.csv every
        check_every-th step's residuals and balance defects."""
        from ._io import CsvLog, Recovery
        from .convergence import RunControl
        from .timing import duration_steps
        from . import __version__
        settings = self.project.run
        output = Path(output)
        if not resume and output.exists() and any(output.iterdir()):
            raise FileExistsError("Use a new/empty refined result directory")
        output.mkdir(parents=True, exist_ok=True)
        recovery = Recovery(output)
        if resume:
            metadata = json.loads((output / "metadata.json").read_text(encoding="utf-8"))
            context = getattr(self, "_checkpoint_context", {})
            saved_project, saved_refinement = load_refined_project(output / "project.json")
            if (saved_project.case_fingerprint != self.project.case_fingerprint or saved_refinement != self.refinement
                    or saved_project.run != settings or metadata.get("case_fingerprint") != self.case_fingerprint
                    or context.get("case_fingerprint") != self.case_fingerprint):
                raise ValueError("Refined project/run settings changed; cannot mix histories")
            if metadata["status"] == "completed":
                raise ValueError("The requested refined run has already completed")
            end_step = context["end_step"]
            history = CsvLog(output / "history.csv", self.HISTORY_COLUMNS,
                             resume=(context["history_bytes"], context["history_sha256"]), recovery=recovery)
            metadata["resumed_with_version"] = __version__
        else:
            end_step = self.steps + duration_steps(settings.duration_convective, self.plan.coarse_level.dt)
            write_json(output / "project.json", refined_project_dict(self.project, self.refinement))
            metadata = {"format": "aerokinetic_refined_run", "version": __version__,
                        "case_fingerprint": self.case_fingerprint, "flow": asdict(self.flow), "grid": asdict(self.grid),
                        "refinement": asdict(self.refinement),
                        "plan": json.loads(json.dumps(self.plan.as_dict(), default=list)),
                        "run": asdict(settings), "start_step": self.steps,
                        "contours": [{"name": e.name, "xy": e.xy.tolist()} for e in self.section.elements],
                        "coefficient_convention": "Cl up; Cd downstream; Cm nose-up about the chosen reference; "
                                                  "area=c*periodic_span",
                        "force_time": "average of both fine momentum-transfer intervals; timestamp at coarse "
                                      "interval midpoint",
                        "experimental_validation": "not_performed", "resume_supported": True,
                        "reference_chord_input_units": self.section.chord,
                        "moment_reference_input_units": self.section.moment_reference,
                        "device": {"backend": "hip-gfx950", "storage": self.storage,
                                   "library": self.L.ak_version().decode()}}
            history = CsvLog(output / "history.csv", self.HISTORY_COLUMNS)
        control = RunControl.open(self, settings, output, resume=resume, callback=residual_callback,
                                  start_step=metadata.get("start_step", self.steps), recovery=recovery)
        metadata.update(status="running", target_steps=end_step,
                        residual_definition="Volume-weighted L2/Linf of d(rho,j_common)/dt*, owned fluid only; "
                                            "integral balance defects of the last coarse step are a separate "
                                            "diagnostic")
        write_json(output / "metadata.json", metadata)

        def checkpoint():
            size, digest = history.durable()
            context = {"case_fingerprint": self.case_fingerprint, "end_step": end_step,
                       "history_bytes": size, "history_sha256": digest}
            context.update(control.checkpoint_context())
            self.save_checkpoint(output / "checkpoint.npz", context)

        status, reason, error = "completed", "duration_reached", None
        se, ke, ce = settings.sample_every, settings.check_every, settings.checkpoint_every
        try:
            while self.steps < end_step:
                if stop_event is not None and stop_event.is_set():
                    status, reason = "stopped", "user_stopped"
                    break
                # one device call up to the next event: a sample a callback must see, a check,
                # a checkpoint, or the end
                events = [end_step, self.steps + 64 * se] + [(self.steps // e + 1) * e for e in (ke, ce) if e]
                if callback is not None:
                    events.append((self.steps // se + 1) * se)
                recs = self.advance(min(events) - self.steps, se)
                for rec in recs:
                    if not all(math.isfinite(rec[k]) for k in ("cl", "cd", "cm")):
                        raise NumericalFailure("Non-finite loads; the run is invalid")
                    history.append(rec)
                    if callback is not None and callback(self, rec) is False:
                        status, reason = "stopped", "user_stopped"
                        break
                history.flush()
                if ke and self.steps % ke == 0:
                    self.diagnostics()
                    control.observe()
                if status == "stopped":
                    break
                selected = control.stop_reason()
                if selected:
                    reason = selected
                    status = "stopped" if selected == "wall_time_limit" else status
                    break
                if ce and self.steps % ce == 0 and self.steps < end_step:
                    checkpoint()
            self.diagnostics()
            checkpoint()
        except Exception as failure:
            status, reason, error = "failed", "failed", f"{type(failure).__name__}: {failure}"
        rows = history.read()
        history.close()
        selected_rows = [r for r in rows if float(r["t_star"]) >= settings.average_after]
        coef = np.array([[float(r[k]) for k in ("cl", "cd", "cm")] for r in selected_rows]).reshape(-1, 3)
        final = {"state": "invalid"}
        if error is None:
            try:
                final = self.diagnostics()
            except Exception as failure:
                status, reason, error = "failed", "failed", f"{type(failure).__name__}: {failure}"
        summary = {"status": status, "steps": self.steps, "fine_steps": 2 * self.steps,
                   "t_star": self.steps * self.plan.coarse_level.dt, "average_after_t_star": settings.average_after,
                   "samples": len(rows), "averaging_samples": len(selected_rows),
                   "mean": dict(zip(("cl", "cd", "cm"), map(float, coef.mean(axis=0)))) if len(coef) else None,
                   "rms": dict(zip(("cl", "cd", "cm"), map(float, coef.std(axis=0)))) if len(coef) else None,
                   "experimental_validation": "not_performed", "statistical_convergence": "not_established",
                   "diagnostics": final}
        summary.update(control.summary(reason))
        control.close()
        if error:
            summary["error"] = error
        metadata.update(status=status, completed_steps=self.steps, termination_reason=reason)
        write_json(output / "metadata.json", metadata)
        write_json(output / "summary.json", summary)
        if error:
            raise RuntimeError(error)
        return summary


def load(path, **kwargs):
    project, settings = load_refined_project(path)
    return RefinedSolver(project, settings, **kwargs)
