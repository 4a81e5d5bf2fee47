"""Whole-update step counts for a requested convective duration.

Same rounding contract as the reference (`src/aerokinetic/timing.py:5-21`):
ceil(duration/dt), except that a quotient within a few ULPs (capped at 1e-8
step) of a positive integer snaps to it, and a positive request is >= 1 step.
"""
import math


def duration_steps(duration, dt):
    if not (math.isfinite(duration) and math.isfinite(dt)) or duration <= 0 or dt <= 0:
        raise ValueError("Duration and time step must be finite and positive")
    q = duration / dt
    if not math.isfinite(q):
        raise ValueError("Requested step count overflows")
    k = round(q)
    if k >= 1 and abs(q - k) <= min(4 * math.ulp(q), 1e-8):
        return int(k)
    return max(1, math.ceil(q))
