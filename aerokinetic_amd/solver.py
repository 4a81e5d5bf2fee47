"""Uniform-grid solver: the drop-in `Solver` of the reference on one MI355X.

Public surface of the reference (`src/aerokinetic/solver.py:27-537`):

* `coefficients(force, moment_z, speed, chord_cells, span_cells)` (solver.py:27-30);
* `Solver(section, flow, grid, progress=None)` with `step(sample)`,
  `diagnostics()`, `run(...)`, `save_frame(...)` and the `NumericalFailure`
  error (solver.py:23-24, 252-537);
* outputs `history.csv` (step,t_star,time_s,cl,cd,cm,local_mach_max,
  density_error_max; with the exterior farfield also the wall mass exchange and
  the exterior source of mass, per step and accumulated), `metadata.json`,
  `summary.json`, `project.json` and mid-span `frames/*.npz`.

The `exterior_regularized_fd_v1` farfield (solver.py:36-41, 81-90, 252-279,
298-310) runs on a lattice padded by one cell in x and y (the exterior ring);
loads, frames, residuals and the density mean cover the physical box only.

All numerics run in the native HIP library: one AA-pattern lattice in HBM,
boundary kernels before each sweep, loads and low-Mach guards reduced on the
device into a history buffer read back in batches.  Extra keyword-only
arguments select the storage precision ("fp64", "fp32", "fp16"), the device, an
explicit plan (fixed lattice shapes) and the sweep tile (bx, by), or "auto" for the
population buffer (of a few allocations), tile and resident-wave cap of each AA form that the
autotuner (aerokinetic_amd.autotune) measures fastest on this lattice.
"""
from __future__ import annotations

from dataclasses import asdict
import json
import math
import time
from pathlib import Path

import numpy as np

from .config import EXTERIOR_FARFIELD, Flow, Grid, Plan
from .geometry import Section
from ._io import PublishedFile, write_json
from .lattice import Lattice
from .project import Project, RunSettings
from .walls import labels, wall_links
from . import _lib


class NumericalFailure(RuntimeError):
    pass


def coefficients(force, moment_z, speed, chord_cells, span_cells):
    """Force on the solid: +Cl up, +Cd downstream, +Cm nose up."""
    q_area = 0.5 * speed**2 * chord_cells * span_cells
    return float(force[1] / q_area), float(force[0] / q_area), float(-moment_z / (q_area * chord_cells))


def seed_parameters(section, flow, plan):
    """Scalars of the seeded initial field (solver.py:167-197): phases from the flow seed,
    leading edge in the normalized frame, wave numbers of the transverse/spanwise modes."""
    rng = np.random.default_rng(flow.seed)
    phase_y, phase_z = rng.uniform(0, 2 * np.pi, 2)
    le = section.leading_edge_normalized(flow.alpha_deg)
    return {"U": flow.lattice_speed, "amp": flow.perturbation * flow.lattice_speed,
            "phase_y": float(phase_y), "phase_z": float(phase_z),
            "ky": 2 * np.pi / (plan.shape[1] * plan.dx), "kz": 2 * np.pi / plan.span,
            "lex": float(le[0]), "ley": float(le[1])}


EXTERIOR_COLUMNS = ["wall_mass_exchange_lattice", "wall_mass_exchange_accumulated",
                    "exterior_source_mass_lattice", "exterior_source_mass_accumulated"]


class Solver:
    HISTORY_COLUMNS = ["step", "t_star", "time_s", "cl", "cd", "cm", "local_mach_max", "density_error_max"]

    def __init__(self, section: Section, flow: Flow = Flow(), grid: Grid = Grid(), progress=None, *,
                 storage="fp64", device=0, plan: Plan | None = None, fp16_scale=0.0, tile=None, graphs=True):
        self.section, self.flow, self.grid = section, flow, grid
        if storage not in _lib.STORE:
            raise ValueError("storage must be fp64, fp32 or fp16")
        self.plan = plan or Plan.create(section, flow, grid)
        self.storage = storage
        self.contours = section.normalized(flow.alpha_deg)
        if self.plan.estimated_memory_gib > grid.memory_limit_gib:
            raise MemoryError(f"Estimated {self.plan.estimated_memory_gib:.2f} GiB exceeds the explicit "
                              f"limit {grid.memory_limit_gib:.2f} GiB. Inspect Plan before choosing a larger limit.")
        self._progress = progress or (lambda message: None)
        nx, ny, nz = self.plan.shape
        # the exterior farfield pads the physical box by one cell in x and y (solver.py:36-41)
        self._exterior = grid.farfield == EXTERIOR_FARFIELD
        pad = 1 if self._exterior else 0
        self.storage_shape = (nx + 2 * pad, ny + 2 * pad, nz)
        self.storage_origin = (self.plan.origin[0] - pad * self.plan.dx, self.plan.origin[1] - pad * self.plan.dx)
        self._core = (slice(pad, pad + nx), slice(pad, pad + ny))
        if self._exterior:
            self.HISTORY_COLUMNS = Solver.HISTORY_COLUMNS + EXTERIOR_COLUMNS
        X, Y = self.storage_shape[:2]
        lab = labels(self.contours, (X, Y), self.storage_origin, self.plan.dx)
        storage_solid = lab[1:-1, 1:-1] > 0
        for k, e in enumerate(self.contours):
            if not (lab == k + 1).any():
                raise ValueError(f"Element {e.name} has no solid cells; increase resolution")
        self._solid_xy = storage_solid[self._core]
        self._fluid = np.broadcast_to(~self._solid_xy[:, :, None], self.plan.shape)
        self._progress("Building the D3Q27 cumulant lattice on the GPU")
        self.lattice = Lattice(X, Y, nz, storage=storage, omega_shear=self.plan.relaxation_rate,
                               omega_bulk=1.0, fp16_scale=fp16_scale, device=device)
        if tile and tile != "auto":
            self.lattice.set_tile(*tile)
        self.lattice.set_graphs(graphs)
        self.lattice.set_solid(storage_solid)
        self.links, fallback = wall_links(self.contours, (X, Y), self.storage_origin, self.plan.dx)
        self.wall_fallback_links = fallback * nz
        self.lattice.set_walls(self.links, len(self.contours), self.storage_origin[0] / self.plan.dx,
                               self.storage_origin[1] / self.plan.dx, grid.cells_per_chord, flow.lattice_speed)
        self._initialize()
        self.tile, self.tile_timings = tile, {}
        self.placement_ms = []
        if tile == "auto":
            # keep the fastest of a few population buffers, time every candidate tile and wave cap
            # of each AA form on it (cached per shape and storage), then start again from the
            # seeded field: the tuning steps advanced the state
            from .autotune import choose_placement, tune
            self.placement_ms = choose_placement(self.lattice, self._initialize)
            self.tile, self.tile_timings = tune(self.lattice)
            self._initialize()
        self.last_loads = None

    # ------------------------------------------------------------------ state
    def _initialize(self):
        self.lattice.init_seeded(seed_parameters(self.section, self.flow, self.plan), self.storage_origin,
                                 self.plan.dx)
        self.lattice.set_farfield(self.grid.farfield, self.flow.lattice_speed)
        self.steps = 0

    def _record(self, row):
        nel = len(self.contours)
        rec = {"cl": float(row[1]), "cd": float(row[2]), "cm": float(row[3]),
               "element_lattice_loads": np.array(row[_lib.HIST_FIXED:_lib.HIST_FIXED + 4 * nel]).reshape(nel, 4),
               "step": int(row[0]), "t_star": (int(row[0]) - 0.5) * self.plan.dt_convective,
               "local_mach_max": float(row[4]), "density_error_max": float(row[5]),
               "density_min": float(row[6]), "nonfinite": float(row[7])}
        if self.plan.dt_seconds is not None:
            rec["time_s"] = (int(row[0]) - 0.5) * self.plan.dt_seconds
        if self._exterior:
            e = _lib.HIST_EXT
            rec.update(wall_mass_exchange_lattice=float(row[8]), wall_mass_exchange_accumulated=float(row[e]),
                       exterior_source_mass_lattice=float(row[e + 1]),
                       exterior_source_mass_accumulated=float(row[e + 2]))
        return rec

    def _exterior_state(self):
        out = np.zeros(13)
        h = self.lattice.h
        _lib.check(_lib.load().ak_exterior_state(h, _lib.dptr(out)), h, "ak_exterior_state")
        return out

    def boundary_diagnostics(self):
        """Exterior-farfield bookkeeping in lattice mass per step (solver.py:298-310); None
        for the other farfields."""
        if not self._exterior:
            return None
        st = self._exterior_state()
        last = None
        if st[11] > 0:
            last = {"step": int(st[11]), "reconstruction_moment_error": float(st[10]),
                    "exterior_source_mass_momentum": [float(v) for v in st[2:6]]}
        return {"model": EXTERIOR_FARFIELD, "pdf_stage": "post_collision",
                "physical_shape": list(self.plan.shape), "storage_shape": list(self.storage_shape),
                "storage_origin": list(self.storage_origin),
                "wall_mass_exchange_lattice": float(st[0]), "wall_mass_exchange_accumulated": float(st[1]),
                "exterior_source_mass_momentum_accumulated": [float(v) for v in st[6:10]],
                "last_exterior_step": last,
                "definition": "Wall incoming-minus-outgoing mass and explicit auxiliary PDF source; a closed "
                              "bookkeeping budget does not prove zero wall mass flux or accuracy."}

    def advance(self, nsteps, sample_every=0):
        """Advance nsteps; returns the records of the sampled steps (device-batched)."""
        if nsteps <= 0:
            return []
        self.lattice.run(nsteps, self.steps, sample_every)
        self.steps += nsteps
        if self._exterior and self._exterior_state()[12]:
            raise NumericalFailure("Non-finite exterior macro predictor")
        if sample_every <= 0:
            return []
        return [self._record(r) for r in self.lattice.read_history()]

    def step(self, sample=False):
        """One time step; with sample=True returns the loads of this step (solver.py:252-279)."""
        recs = self.advance(1, 1 if sample else 0)
        if not sample:
            return None
        rec = recs[-1]
        loads = {k: rec[k] for k in ("cl", "cd", "cm", "element_lattice_loads", "step", "t_star")}
        for k in ["time_s"] + EXTERIOR_COLUMNS:
            if k in rec:
                loads[k] = rec[k]
        self.last_loads = loads
        self._last_diag = rec
        return loads

    def diagnostics(self):
        """Low-Mach validity guards of the current state (solver.py:281-297)."""
        mach, derr, rmin, bad, rsum = self.lattice.diagnostics()
        if bad or not (math.isfinite(mach) and math.isfinite(derr)):
            raise NumericalFailure("Non-finite fluid state; the run is invalid")
        if mach > 0.3 or derr > 0.1 or rmin <= 0:
            raise NumericalFailure(f"Low-Mach validity check failed: local Ma={mach:.3g}, |rho-1|max={derr:.3g}")
        nfluid = int(self._fluid.sum())
        return {"local_mach_max": float(mach), "density_mean": float(rsum / nfluid),
                "density_error_max": float(derr), "wall_fallback_links": self.wall_fallback_links}

    def populations(self):
        """Post-collision shifted populations g[x, y, z, 27] (FP64 copy)."""
        return self.lattice.populations()

    def set_populations(self, g):
        self.lattice.set_populations(g)

    # ------------------------------------------------------------------ outputs
    @property
    def residual_dt(self):
        """Convective time per step (the residual's time unit)."""
        return self.plan.dt_convective

    def residual_norms(self, reset=False, dt=1.0):
        """L2 and L-infinity of d(rho, j/U)/dt over owned fluid cells since the last call
        (convergence.py:55-96); reset only takes the snapshot."""
        from .convergence import combine_levels
        out = np.zeros(9)
        h = self.lattice.h
        _lib.check(_lib.load().ak_residual(h, 0, self.flow.lattice_speed, float(dt), 1 if reset else 0,
                                           _lib.dptr(out)), h, "ak_residual")
        return (np.zeros(4), np.zeros(4)) if reset else combine_levels([(out, self.plan.dx ** 3)])

    @staticmethod
    def _failed_safe(report, failed):
        """The boundary report of a run; after a failure non-finite numbers become null
        (and are flagged) so that the failed run's JSON stays valid (solver.py:519-530)."""
        try:
            value = report()
        except Exception:
            if not failed:
                raise
            return None
        if not failed:
            return value
        replaced = []

        def invalid(constant):
            replaced.append(constant)
            return None
        value = json.loads(json.dumps(value, allow_nan=True), parse_constant=invalid)
        if replaced:
            value["nonfinite_diagnostics_replaced_with_null"] = True
        return value

    def save_checkpoint(self, path, context=None):
        """Restartable device state + run context (checkpoint.py:32-70)."""
        from .checkpoint import save_checkpoint
        return save_checkpoint(self, path, context)

    @classmethod
    def from_checkpoint(cls, path, progress=None, **kwargs):
        """Rebuild a case from its checkpoint and continue from it (solver.py:340-343)."""
        from .checkpoint import restore_checkpoint
        return restore_checkpoint(cls, path, progress, **kwargs)

    @classmethod
    def resume_run(cls, output, callback=None, stop_event=None, progress=None, residual_callback=None):
        """Continue the interrupted run in `output` with the run settings it recorded
        (solver.py:345-351)."""
        output = Path(output)
        metadata = json.loads((output / "metadata.json").read_text(encoding="utf-8"))
        solver = cls.from_checkpoint(output / "checkpoint.npz", progress)
        run = dict(metadata["run"])
        if isinstance(run.get("termination"), dict):
            from .convergence import StopCriteria
            run["termination"] = StopCriteria(**run["termination"])
        return solver.run(output=output, callback=callback, stop_event=stop_event, residual_callback=residual_callback,
                          resume=True, **run)

    def save_frame(self, path, max_pixels=800):
        """Mid-span speed, vorticity and Cp planes (solver.py:312-334)."""
        nx, ny, nz = self.plan.shape
        stride = max(1, math.ceil(max(nx, ny) / max_pixels))
        rho, vel = self.lattice.plane(nz // 2)
        rho, vel = rho[self._core][::stride, ::stride], vel[self._core][::stride, ::stride]
        U = self.flow.lattice_speed
        speed = np.linalg.norm(vel, axis=-1) / U
        h = stride * self.plan.dx
        omega_z = (np.gradient(vel[..., 1], h, axis=0) - np.gradient(vel[..., 0], h, axis=1)) / U
        cp = (rho - 1) / (1.5 * U**2)
        mask = self._solid_xy[::stride, ::stride]
        for a in (speed, omega_z, cp):
            a[mask] = np.nan
        with PublishedFile(path) as handle:
            np.savez_compressed(handle, speed=speed, omega_z=omega_z, cp=cp, mask=mask,
                                origin=self.plan.origin, dx=self.plan.dx * stride,
                                t_star=self.steps * self.plan.dt_convective)

    def run(self, duration_convective, output, sample_every=20, frame_every=200, average_after=0.0,
            check_every=100, callback=None, stop_event=None, checkpoint_every=1000, resume=False,
            termination=None, residual_callback=None):
        """Run a further t* = t U / c interval into a new/empty directory (solver.py:353-537).

        callback(solver, record) may return False to stop.  Mean/RMS coefficients
        are plain sample statistics over t_star >= average_after.
        """
        from ._io import CsvLog, Recovery
        from .convergence import RunControl, StopCriteria
        from .timing import duration_steps
        from . import __version__
        import shutil
        termination = termination or StopCriteria()
        settings = RunSettings(duration_convective, sample_every, frame_every, check_every, average_after,
                               checkpoint_every, termination)
        project = Project(self.section, self.flow, self.grid, settings, "+".join(e.name for e in self.section.elements))
        output = Path(output)
        if not resume and output.exists() and any(output.iterdir()):
            raise FileExistsError("Use a new/empty output directory to avoid mixing histories")
        output.mkdir(parents=True, exist_ok=True)
        (output / "frames").mkdir(exist_ok=resume)
        end_step = self.steps + duration_steps(duration_convective, self.plan.dt_convective)
        recovery = Recovery(output)
        start_step = self.steps
        if resume:
            # checkpoint.py / solver.py:389-425: same project, settings and durable history prefix
            old = json.loads((output / "metadata.json").read_text(encoding="utf-8"))
            context = getattr(self, "_checkpoint_context", {})
            saved = Project.load(output / "project.json")
            if (saved.case_fingerprint != project.case_fingerprint or asdict(saved.run) != asdict(settings)
                    or old.get("case_fingerprint") != project.case_fingerprint
                    or context.get("case_fingerprint") != project.case_fingerprint):
                raise ValueError("Resume requires the original project, settings and checkpoint")
            if old["status"] == "completed":
                raise ValueError("This run has already completed")
            start_step = old.get("start_step", 0)
            end_step = context["end_step"]
            history = CsvLog(output / "history.csv", self.HISTORY_COLUMNS,
                             resume=(context["history_bytes"], context["history_sha256"]), recovery=recovery)
            # frames written after the last durable checkpoint are kept aside with the rows
            for f in (output / "frames").glob("frame_*.npz"):
                if int(f.stem.split("_")[-1]) > self.steps:
                    shutil.move(str(f), str(recovery.directory() / f.name))
            rows = [{"step": int(r["step"]), "t_star": float(r["t_star"]),
                     **{k: float(r[k]) for k in ("cl", "cd", "cm")}} for r in history.read()]
            if rows and rows[-1]["step"] > self.steps:
                raise ValueError("History extends beyond the checkpoint")
        else:
            history = CsvLog(output / "history.csv", self.HISTORY_COLUMNS)
            rows = []
        if self.steps > end_step:
            raise ValueError("Checkpoint is past the requested end time")
        metadata = {"version": __version__, "flow": asdict(self.flow), "grid": asdict(self.grid),
                    "plan": self.plan.as_dict(), "reference_chord_input_units": self.section.chord,
                    "moment_reference_input_units": self.section.moment_reference,
                    "contours": [{"name": e.name, "xy": e.xy.tolist()} for e in self.section.elements],
                    "coefficient_convention": "Cl up; Cd downstream; Cm nose-up about moment_reference; area=c*periodic_span",
                    "force_time": "midpoint of boundary momentum-transfer interval",
                    "status": "running", "experimental_validation": "not_performed",
                    "case_fingerprint": project.case_fingerprint, "run": asdict(settings),
                    "start_step": start_step, "start_t_star": start_step * self.plan.dt_convective,
                    "resume_supported": True, "end_step": end_step,
                    "device": {"backend": "hip-gfx950", "storage": self.storage,
                               "library": _lib.load().ak_version().decode()}}
        if not resume:
            project.save(output / "project.json")
        control = RunControl.open(self, settings, output, resume=resume, callback=residual_callback,
                                  recovery=recovery)
        metadata["residual_definition"] = ("Volume-weighted L2/Linf of d(rho,j_common)/dt*, owned fluid only; "
                                           "stationary-flow residual, not a turbulence-accuracy measure")
        if self._exterior:
            metadata["boundary_model"] = self.boundary_diagnostics()
        write_json(output / "metadata.json", metadata)
        state, reason = "completed", "duration_reached"
        start = time.monotonic()
        def checkpoint():
            size, digest = history.durable()
            context = {"case_fingerprint": project.case_fingerprint, "end_step": end_step,
                       "history_bytes": size, "history_sha256": digest}
            context.update(control.checkpoint_context())
            self.save_checkpoint(output / "checkpoint.npz", context)

        try:
            self.diagnostics()
            self.save_frame(output / "frames" / f"frame_{self.steps:010d}.npz")
            stop = False
            while self.steps < end_step and not stop:
                if stop_event is not None and stop_event.is_set():
                    state, reason = "stopped", "user_stopped"
                    break
                # next event: a sample (when a callback must see it), a check, a frame, the end
                nxt = end_step
                for every in (check_every, frame_every, checkpoint_every) + ((sample_every,) if callback else ()):
                    if every:
                        nxt = min(nxt, (self.steps // every + 1) * every)
                if not callback:
                    nxt = min(nxt, self.steps + 256 * max(sample_every, 1))
                recs = self.advance(nxt - self.steps, sample_every)
                for rec in recs:
                    if rec["nonfinite"] or not all(math.isfinite(rec[k]) for k in ("cl", "cd", "cm")):
                        raise NumericalFailure("Non-finite loads; the run is invalid")
                    if rec["local_mach_max"] > 0.3 or rec["density_error_max"] > 0.1 or rec["density_min"] <= 0:
                        raise NumericalFailure(f"Low-Mach validity check failed: local Ma={rec['local_mach_max']:.3g}, "
                                               f"|rho-1|max={rec['density_error_max']:.3g}")
                    history.append(rec)
                    rows.append(rec)
                    self.last_loads = rec
                    if callback is not None and callback(self, rec) is False:
                        state, reason, stop = "stopped", "user_stopped", True
                        break
                history.flush()
                if check_every and self.steps % check_every == 0:
                    self.diagnostics()
                    control.observe()
                if (frame_every and self.steps % frame_every == 0) or self.steps == end_step:
                    self.save_frame(output / "frames" / f"frame_{self.steps:010d}.npz")
                if stop:
                    break
                selected = control.stop_reason()
                if selected:
                    reason = selected
                    if reason == "wall_time_limit":
                        state = "stopped"
                    break
                if checkpoint_every and self.steps % checkpoint_every == 0 and self.steps < end_step:
                    checkpoint()
            self.diagnostics()
            self.save_frame(output / "frames" / f"frame_{self.steps:010d}.npz")
            checkpoint()
        except BaseException as error:
            state, reason = "failed", "failed"
            metadata["failure"] = f"{type(error).__name__}: {error}"
            raise
        finally:
            sel = np.array([[r["cl"], r["cd"], r["cm"]] for r in rows if r["t_star"] >= average_after])
            ok = len(sel) and state != "failed"
            summary = {"status": state, "steps": self.steps, "t_star": self.steps * self.plan.dt_convective,
                       "wall_clock_seconds": time.monotonic() - start, "samples": len(rows),
                       "averaging_samples": len(sel), "average_after_t_star": average_after,
                       "mean": dict(zip(("cl", "cd", "cm"), map(float, sel.mean(axis=0)))) if ok else None,
                       "rms_fluctuation": dict(zip(("cl", "cd", "cm"), map(float, sel.std(axis=0)))) if ok else None,
                       "wall_fallback_links": self.wall_fallback_links,
                       "experimental_validation": "not_performed"}
            summary.update(control.summary(reason))
            control.close()
            history.close()
            if self._exterior:
                boundary = self._failed_safe(self.boundary_diagnostics, state == "failed")
                summary["boundary_diagnostics"] = metadata["boundary_model"] = boundary
            metadata["termination_reason"] = reason
            metadata["status"] = state
            write_json(output / "metadata.json", metadata)
            write_json(output / "summary.json", summary)
        return summary
