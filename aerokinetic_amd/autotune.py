"""Sweep tile autotuner.

Times the fused AA sweep for each candidate thread tile (bx threads along x, by rows
along y; bx a multiple of the 64-lane wavefront, bx*by <= 1024; bx = 0: row-linear
blocks of `by` consecutive cells of a plane) with HIP events on
the lattice's own stream and keeps the fastest for each sweep parity (odd sweeps gather
from neighbours, even sweeps read in place; each parity keeps its own tile).  The lattice
state advances while tuning, so tune on a scratch copy (or before the initial state is set).  Results are cached per
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


WAVE_CAPS = (6, 5, 4, 3)
PLACEMENTS = 32          # at most this many candidate population buffers
PLACEMENT_MEMORY = 0.6   # ... holding at most this fraction of the free device memory at once


def choose_placement(lattice, initialize, candidates=None, steps=4):
    """Keep the fastest of several population buffers of `lattice`.

    The sweep's 27 in-place population rows per cell meet HBM channels through the physical
    pages behind the buffer, and the same launch on a buffer of other pages has measured
    up to 13 % faster or slower; a given buffer keeps its rate.  Each candidate (the
    lattice's own buffer first, then fresh allocations while the earlier ones are still
    held) gets the initial state from `initialize()` and is timed over `steps` single steps;
    the fastest stays bound, the others are released, and `initialize()` runs again on it.
    `candidates` defaults to as many buffers as PLACEMENT_MEMORY of the free memory holds, at
    most PLACEMENTS; allocation also stops at the first failure.  Returns the mean sweep
    milliseconds (even and odd form averaged) of every candidate, in order."""
    import torch
    if candidates is None:
        free, _ = torch.cuda.mem_get_info(lattice.torch_device)
        candidates = min(PLACEMENTS, 1 + int(free * PLACEMENT_MEMORY // max(lattice.nbytes, 1)))
    held, times = [], []
    for k in range(max(1, candidates)):
        if k:
            try:
                lattice._bind(lattice.new_buffer())
            except torch.cuda.OutOfMemoryError:
                break
        held.append(lattice._buf)
        initialize()
        times.append(sum(_time_forms(lattice, steps, 0)) / 2)
    best = min(range(len(times)), key=times.__getitem__)
    if lattice._buf is not held[best]:
        lattice._bind(held[best])
    del held
    torch.cuda.empty_cache()
    initialize()
    return times


def _time_forms(lattice, steps, first_step):
    """Mean sweep milliseconds (even form, odd form) over `steps` single timed steps after two
    warm-up steps; each step is credited to the AA form it ran.  The steps are sampled (loads
    and diagnostics, as a run recording every step has them); their history rows are
    discarded."""
    lattice.run_profiled(2, first_step, 1)
    per_form = {False: [], True: []}
    for _ in range(max(steps, 2)):
        odd = lattice.next_sweep_odd()
        sweep_ms, _ = lattice.run_profiled(1, first_step, 1)
        per_form[odd].append(sweep_ms)
    lattice.read_history()
    return tuple(sum(v) / len(v) for v in (per_form[False], per_form[True]))


def tune(lattice, steps=8, candidates=CANDIDATES, first_step=0, caps=WAVE_CAPS, finalists=3):
    """Pick the fastest sweep launch of each AA form for `lattice` (an aerokinetic_amd.lattice.Lattice).

    The two forms access memory differently (the odd form gathers from x-shifted rows whose
    cache lines neighbouring waves share; the even form reads in place), so each keeps its own
    tile and its own cap on resident waves per SIMD (fewer concurrent waves, each streaming
    27 population rows in and out, can move more bytes per second than a full CU).  Every
    candidate tile is timed without a cap; the `finalists` fastest tiles of each form are then
    timed under each cap in `caps`.  Returns ((bx, by, waves) of the even form, (bx, by, waves)
    of the odd form; waves 0 = no cap) and the timings {(bx, by, waves): (even ms, odd ms)}."""
    key = (lattice.shape, lattice.storage)
    if key in _cache:
        _apply(lattice, _cache[key])
        return _cache[key], {}
    timings = {}
    for bx, by in candidates:
        if bx > 64 and lattice.shape[0] < bx:
            continue
        lattice.set_tile(bx, by)
        timings[(bx, by, 0)] = _time_forms(lattice, steps, first_step)
    ranked = [sorted(timings, key=lambda t: timings[t][form])[:finalists] for form in (0, 1)]
    for even, odd in zip(*ranked):
        for cap in caps:
            tried = []
            for form, tile in ((0, even), (1, odd)):
                lattice.set_tile(*tile[:2], odd=bool(form))
                try:
                    lattice.set_waves(cap, odd=bool(form))
                    tried.append((form, tile[:2] + (cap,)))
                except Exception:  # the cap does not fit this tile's blocks
                    lattice.set_tile(*tile[:2], odd=bool(form))
            if not tried:
                continue
            t = _time_forms(lattice, steps, first_step)
            for form, tile in tried:
                prev = timings.get(tile, (float("inf"), float("inf")))
                timings[tile] = tuple(t[f] if f == form else prev[f] for f in (0, 1))
    best = tuple(min(timings, key=lambda t: timings[t][form]) for form in (0, 1))
    _apply(lattice, best)
    _cache[key] = best
    return best, timings


def _apply(lattice, tiles):
    for form, tile in ((0, tiles[0]), (1, tiles[1])):
        lattice.set_tile(*tile[:2], odd=bool(form))
        lattice.set_waves(tile[2] if len(tile) > 2 else 0, odd=bool(form))


apply_launch = _apply


def format_launch(tiles):
    """'BXxBYcW/BXxBYcW' (even form / odd form; W = wave cap, 0 none) of tune()'s result."""
    return "/".join(f"{t[0]}x{t[1]}c{t[2] if len(t) > 2 else 0}" for t in tiles)


def parse_launch(text):
    """Inverse of format_launch: ((bx, by, waves) even, (bx, by, waves) odd)."""
    forms = []
    for part in text.split("/"):
        tile, _, cap = part.partition("c")
        bx, by = tile.split("x")
        forms.append((int(bx), int(by), int(cap or 0)))
    if len(forms) != 2:
        raise ValueError(f"sweep launch {text!r}: expected 'even/odd' such as 256x1c3/0x256c3")
    return tuple(forms)


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
