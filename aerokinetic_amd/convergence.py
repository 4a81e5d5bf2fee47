"""Conservative stationarity residuals and the explicit stop criteria.

Restates src/aerokinetic/convergence.py:
* `StopCriteria` (convergence.py:26-53): the reference's fields and validation, so project
  files round-trip;
* `RunControl` (convergence.py:97-176): every check, the residuals d(rho, j/U)/dt* of the
  owned fluid cells since the previous check (volume-weighted L2 over both levels of a
  refined case, and L-infinity), the balance defects of the last coarse step of a
  two-level case, `residuals.csv`, the steady-mode streak and the stop reasons (residual
  targets met, max_steps, wall time).

The residuals are reduced on the GPU (`ak_residual`): the previous snapshot of
(rho, j/U) stays in device memory and travels with device checkpoints, so only the
eight norms and the cell count cross to the host.
"""
from dataclasses import asdict, dataclass
import math
from pathlib import Path
import shutil
import time

import numpy as np

from ._io import CsvLog, Recovery

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


class RunControl:
    """Residual checks and stop decisions of one run (convergence.py:97-176).

    A check asks the solver for the norms of d(rho, j/U)/dt* since the previous check, which
    the device reduces against a snapshot of the conserved fields it keeps (so only eight
    norms and a cell count cross to the host), adds the balance defects of the last coarse
    step when the solver keeps a conservation budget (two-level cases), appends one
    residuals.csv row and updates the steady-mode streak.  The clock of the checks (previous
    check step, streak, last record, first step of the run) travels in the checkpoint
    context, beside the durable prefix of residuals.csv.
    """

    def __init__(self, solver, criteria, log, dt, *, callback=None, start_step=None, clock=None):
        self.solver, self.criteria, self.log, self.dt, self.callback = solver, criteria, log, dt, callback
        self.started = time.monotonic()
        if clock is None:
            self.previous_step, self.streak, self.last_record = solver.steps, 0, None
            self.start_step = solver.steps if start_step is None else start_step
            solver.residual_norms(reset=True)  # snapshot of the current state on the device
        else:
            prev, streak = clock.get("previous_step"), clock.get("streak")
            if type(prev) is not int or not 0 <= prev <= solver.steps or type(streak) is not int or streak < 0:
                raise ValueError("Invalid residual checkpoint clock/streak")
            self.previous_step, self.streak = prev, streak
            self.last_record, self.start_step = clock["last_record"], clock["run_start_step"]
        solver._run_control = self

    @classmethod
    def open(cls, solver, settings, output, *, resume=False, callback=None, start_step=None, recovery=None):
        """The control of a run writing into `output`; on resume, residuals.csv continues from
        the prefix the checkpoint recorded and the clock from the checkpoint context."""
        context = getattr(solver, "_checkpoint_context", {}) if resume else {}
        path = Path(output) / "residuals.csv"
        prefix = None
        if "residuals_bytes" in context:
            prefix = (context["residuals_bytes"], context["residuals_sha256"])
        elif resume and path.exists() and path.stat().st_size:
            # no durable residual rows: keep what is there aside and start the file again
            shutil.copy2(path, (recovery or Recovery(output)).directory() / "residuals_before_resume.csv")
        log = CsvLog(path, RESIDUAL_COLUMNS, resume=prefix, recovery=recovery)
        clock = getattr(solver, "_restored_control", None) if resume else None
        return cls(solver, settings.termination, log, solver.residual_dt, callback=callback, start_step=start_step,
                   clock=clock)

    def observe(self):
        step = self.solver.steps
        if step <= self.previous_step:
            return None
        l2, linf = self.solver.residual_norms(dt=(step - self.previous_step) * self.dt)
        record = {"step": step, "t_star": step * self.dt}
        record.update((f"l2_{c}", float(v)) for c, v in zip(CHANNELS, l2))
        record.update((f"linf_{c}", float(v)) for c, v in zip(CHANNELS, linf))
        budget = getattr(self.solver, "last_budget", None)
        if budget:
            defects = (budget["mass_residual_per_volume"], *budget["momentum_residual_per_volume"])
            record.update((f"balance_{c}", abs(float(v)) / self.dt) for c, v in zip(CHANNELS, defects))
        steady = self.criteria.mode == "steady" and record["t_star"] >= self.criteria.minimum_t_star
        self.streak = self.streak + 1 if steady and all(v <= t for v, t in zip(l2, self.criteria.thresholds)) else 0
        record["streak"] = self.streak
        self.previous_step, self.last_record = step, record
        self.log.append(record)
        self.log.flush()
        if self.callback:
            self.callback(self.solver, dict(record))
        return record

    def stop_reason(self):
        c = self.criteria
        if c.mode == "steady" and self.streak >= c.consecutive_checks:
            return "residual_targets_met"
        if c.max_steps and self.solver.steps - self.start_step >= c.max_steps:
            return "max_steps_reached"
        if c.wall_time_limit_s and time.monotonic() - self.started >= c.wall_time_limit_s:
            return "wall_time_limit"
        return None

    def checkpoint_context(self):
        size, digest = self.log.durable()
        return {"residuals_bytes": size, "residuals_sha256": digest,
                "run_control": {"previous_step": self.previous_step, "streak": self.streak,
                                "last_record": self.last_record, "run_start_step": self.start_step}}

    def summary(self, reason):
        return {"termination_reason": reason, "termination": asdict(self.criteria), "last_residuals": self.last_record,
                "steady_residual_criterion_met": reason == "residual_targets_met",
                "statistical_convergence": "not_established"}

    def close(self):
        self.log.close()


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
