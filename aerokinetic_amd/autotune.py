"""Sweep tile autotuner.

Times the fused AA sweep for each candidate thread tile (bx threads along x, by rows
along y; bx a multiple of the 64-lane wavefront, bx*by <= 1024; bx = 0: row-linear
blocks of `by` consecutive cells of a plane) with HIP events on
the lattice's own stream and keeps the fastest.  A tile is tried on both sweep
parities (odd sweeps gather from neighbours, even sweeps read in place) by timing an
even number of steps.  The lattice state advances while tuning, so tune on a
scratch copy (or before the initial state is set).  Results are cached per
(shape, storage) for the process.

`tune_refined` does the same for both levels of a two-level case at once (the coarse and
fine sweeps are timed separately inside each coarse step); the case's state is saved before
and restored after, so tuning leaves it as it was.  `tune_batch` picks one 256-thread tile
for every member sweep of a batch (the batched sweep launches 256-thread blocks).
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


def default_tile(nx):
    """The library's tile when none is set: 64 x 4 on whole-wave rows, else row-linear 256."""
    return (64, 4) if nx % 64 == 0 else (0, 256)


def _fits(tile, nx):
    return not (tile[0] > 64 and nx < tile[0])


def tune_refined(solver, steps=4, candidates=CANDIDATES):
    """Fastest sweep tile of each level of a RefinedSolver: ((bx, by) coarse, (bx, by) fine),
    timings {level: {tile: ms per sweep}}."""
    import numpy as np
    from . import _lib
    nxs = (solver.plan.shape[0], solver.fine_shape[0] + 2 * solver.pad)
    key = ("refined", tuple(solver.plan.shape), tuple(solver.fine_shape), solver.storage)
    if key in _cache:
        solver.set_tiles(*_cache[key])
        return _cache[key], {}
    L, h = solver.L, solver.h
    n = int(L.ak_refined_state_bytes(h))
    state = np.empty(n, dtype=np.uint8)
    _lib.check(L.ak_refined_save_state(h, _lib._vp(state.ctypes.data), n), None, "ak_refined_save_state")
    steps0 = solver.steps
    timings = ({}, {})
    for tile in candidates:
        use = [tile if _fits(tile, nx) else default_tile(nx) for nx in nxs]
        solver.set_tiles(*use)
        solver.run_profiled(2, 0)
        _, cs, fs = solver.run_profiled(steps, 0)
        if use[0] == tile:
            timings[0][tile] = cs / steps
        if use[1] == tile:
            timings[1][tile] = fs / (2 * steps)
    best = tuple(min(t, key=t.get) for t in timings)
    solver.set_tiles(*best)
    _lib.check(L.ak_refined_load_state(h, _lib._vp(state.ctypes.data), n), None, "ak_refined_load_state")
    solver.steps = steps0
    _cache[key] = best
    return best, {"coarse": timings[0], "fine": timings[1]}
