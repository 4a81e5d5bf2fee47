"""Sweep tile autotuner.

Times the fused AA sweep for each candidate thread tile (bx threads along x, by rows
along y; bx a multiple of the 64-lane wavefront, bx*by <= 1024; bx = 0: row-linear
blocks of `by` consecutive cells of a plane) with HIP events on
the lattice's own stream and keeps the fastest.  A tile is tried on both sweep
parities (odd sweeps gather from neighbours, even sweeps read in place) by timing an
even number of steps.  The lattice state advances while tuning, so tune on a
scratch copy (or before the initial state is set).  Results are cached per
(shape, storage) for the process.
"""
from __future__ import annotations

CANDIDATES = ((64, 1), (64, 2), (64, 4), (128, 1), (128, 2), (256, 1), (64, 8), (128, 4), (256, 2), (512, 1),
              (64, 16), (128, 8), (256, 4), (512, 2), (1024, 1), (0, 256), (0, 512))
_cache: dict = {}


def tune(lattice, steps=8, candidates=CANDIDATES, first_step=0):
    """Pick the fastest sweep tile for `lattice` (an aerokinetic_amd.lattice.Lattice)."""
    key = (lattice.shape, lattice.storage)
    if key in _cache:
        lattice.set_tile(*_cache[key])
        return _cache[key], {}
    timings = {}
    for bx, by in candidates:
        if bx > 64 and lattice.shape[0] < bx:
            continue
        lattice.set_tile(bx, by)
        lattice.run_profiled(2, first_step, 0)  # warm-up of this tile
        sweep_ms, _ = lattice.run_profiled(steps, first_step, 0)
        timings[(bx, by)] = sweep_ms / steps
    best = min(timings, key=timings.get)
    lattice.set_tile(*best)
    _cache[key] = best
    return best, timings
