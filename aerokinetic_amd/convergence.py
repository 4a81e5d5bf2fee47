"""Stop criteria carried by run settings.

`StopCriteria` keeps the reference's fields and validation
(`src/aerokinetic/convergence.py:26-53`) so project files round-trip; the
run driver here honours the `max_steps` and `wall_time_limit_s` limits.
"""
from dataclasses import dataclass
import math


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
