"""Conservative stationarity residuals and the explicit stop criteria.

Restates src/aerokinetic/convergence.py:
* `StopCriteria` (convergence.py:26-53): the reference's fields and validation, so project
  files round-trip;
* `RunControl` (convergence.py:104-176): every check, the residuals d(rho, j/U)/dt* of the
  owned fluid cells since the previous check (volume-weighted L2 over both levels of a
  refined case, and L-infinity), `residuals.csv`, the steady-mode streak and the stop
  reasons (residual targets met, max_steps, wall time).

The residuals are reduced on the GPU (`ak_residual`): the previous snapshot of
(rho, j/U) stays in device memory and travels with device checkpoints, so only the
eight norms and the cell count cross to the host.
"""
from dataclasses import asdict, dataclass
import csv
import math
import os
from pathlib import Path
import shutil
import time
import uuid

import numpy as np

from ._io import file_hash

CHANNELS = ("rho", "jx", "jy", "jz")
RESIDUAL_COLUMNS = ("step", "t_star", *(f"l2_{c}" for c in CHANNELS), *(f"linf_{c}" for c in CHANNELS),
                    *(f"balance_{c}" for c in CHANNELS), "streak")


@dataclass(frozen=True)
class StopCriteria:
    mode: str = "time"
    mass_tolerance: float = 1e-5
    momentum_x_tolerance: float = 1e-5
    momentum_y_tolerance: float = 1e-5
    momentum_z_tolerance: float = 1e-5
    minimum_t_star: float = 1.0
    consecutive_checks: int = 10
    max_steps: int = 0
    wall_time_limit_s: float = 0.0

    def __post_init__(self):
        if self.mode not in ("time", "steady"):
            raise ValueError("Stop mode must be time or steady")
        for name in ("mass_tolerance", "momentum_x_tolerance", "momentum_y_tolerance", "momentum_z_tolerance"):
            v = getattr(self, name)
            if not math.isfinite(v) or v <= 0:
                raise ValueError(f"{name} must be finite and positive")
        for name in ("minimum_t_star", "wall_time_limit_s"):
            v = getattr(self, name)
            if not math.isfinite(v) or v < 0:
                raise ValueError(f"{name} must be finite and nonnegative")
        for name, lo in (("consecutive_checks", 1), ("max_steps", 0)):
            v = getattr(self, name)
            if type(v) is not int or v < lo:
                raise ValueError(f"{name} must be an integer >= {lo}")

    @property
    def thresholds(self):
        return (self.mass_tolerance, self.momentum_x_tolerance,
                self.momentum_y_tolerance, self.momentum_z_tolerance)


def validate_history(output, context):
    if "residuals_bytes" in context:
        path = Path(output) / "residuals.csv"
        if file_hash(path, context["residuals_bytes"]) != context["residuals_sha256"]:
            raise ValueError("Residual history before checkpoint changed")


class RunControl:
    """Residual observation and stop control of one run (convergence.py:104-176)."""

    def __init__(self, solver, settings, output, *, resume=False, callback=None, start_step=None):
        self.solver, self.criteria, self.callback = solver, settings.termination, callback
        self.path = Path(output) / "residuals.csv"
        self.started = time.monotonic()
        self.dt = solver.residual_dt
        self.run_start_step = solver.steps if start_step is None else start_step
        self.previous_step, self.streak, self.last_record = solver.steps, 0, None
        restored = getattr(solver, "_restored_control", None) if resume else None
        if restored is not None:
            if (type(restored["previous_step"]) is not int or not 0 <= restored["previous_step"] <= solver.steps
                    or type(restored["streak"]) is not int or restored["streak"] < 0):
                raise ValueError("Invalid residual checkpoint clock/streak")
            self.previous_step, self.streak = restored["previous_step"], restored["streak"]
            self.last_record, self.run_start_step = restored["last_record"], restored["run_start_step"]
        else:
            solver.residual_norms(reset=True)  # snapshot of the current state
        context = getattr(solver, "_checkpoint_context", {}) if resume else {}
        validate_history(output, context)
        if resume and self.path.exists():
            prefix = context.get("residuals_bytes", 0)
            if self.path.stat().st_size > prefix:
                recovery = Path(output) / "recovery" / uuid.uuid4().hex
                recovery.mkdir(parents=True)
                shutil.copy2(self.path, recovery / "residuals_before_resume.csv")
                with self.path.open("r+b") as handle:
                    handle.truncate(prefix)
        append = resume and "residuals_bytes" in context
        self.handle = self.path.open("a" if append else "w", newline="", encoding="utf-8")
        self.writer = csv.DictWriter(self.handle, RESIDUAL_COLUMNS)
        if not append:
            self.writer.writeheader()
            self.handle.flush()
        solver._run_control = self

    def observe(self):
        step = self.solver.steps
        if step <= self.previous_step:
            return None
        l2, linf = self.solver.residual_norms(dt=(step - self.previous_step) * self.dt)
        record = {"step": step, "t_star": step * self.dt}
        record.update(zip((f"l2_{c}" for c in CHANNELS), map(float, l2)))
        record.update(zip((f"linf_{c}" for c in CHANNELS), map(float, linf)))
        eligible = (self.criteria.mode == "steady" and record["t_star"] >= self.criteria.minimum_t_star
                    and all(v <= lim for v, lim in zip(l2, self.criteria.thresholds)))
        self.streak = self.streak + 1 if eligible else 0
        record["streak"] = self.streak
        self.previous_step, self.last_record = step, record
        self.writer.writerow(record)
        self.handle.flush()
        if self.callback:
            self.callback(self.solver, dict(record))
        return record

    def stop_reason(self):
        if self.criteria.mode == "steady" and self.streak >= self.criteria.consecutive_checks:
            return "residual_targets_met"
        if self.criteria.max_steps and self.solver.steps - self.run_start_step >= self.criteria.max_steps:
            return "max_steps_reached"
        if self.criteria.wall_time_limit_s and time.monotonic() - self.started >= self.criteria.wall_time_limit_s:
            return "wall_time_limit"
        return None

    def checkpoint_context(self):
        self.handle.flush()
        os.fsync(self.handle.fileno())
        return {"residuals_bytes": self.path.stat().st_size, "residuals_sha256": file_hash(self.path),
                "run_control": {"previous_step": self.previous_step, "streak": self.streak,
                                "last_record": self.last_record, "run_start_step": self.run_start_step}}

    def summary(self, reason):
        return {"termination_reason": reason, "termination": asdict(self.criteria), "last_residuals": self.last_record,
                "steady_residual_criterion_met": reason == "residual_targets_met",
                "statistical_convergence": "not_established"}

    def close(self):
        self.handle.close()


def combine_levels(parts):
    """Volume-weighted L2 and L-infinity over levels: parts = [(out9, cell_volume), ...]."""
    square, infinity, volume = np.zeros(4), np.zeros(4), 0.0
    for out, v in parts:
        square += v * np.asarray(out[:4])
        infinity = np.maximum(infinity, np.asarray(out[4:8]))
        volume += v * out[8]
    if volume <= 0:
        raise ValueError("Residual domain has no owned fluid volume")
    return np.sqrt(square / volume), infinity
