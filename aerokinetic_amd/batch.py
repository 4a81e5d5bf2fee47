"""Several two-level cases stepped together on one GPU.

`RefinedBatch` wraps the library's batch entry points (`ak_batch_*`).  The members are
ordinary `RefinedSolver` objects, each with its own populations, transfer buffers, history,
step count and stream.  Per coarse step of the batch:

* the members' coarse sweeps run as one launch over their concatenated tile lists, and
  likewise each fine substep's sweeps;
* every transfer, farfield, wall and load kernel also runs once for all members, over the
  members' concatenated work lists (a block finds its member in a table of first blocks);
* the whole step is a captured HIP graph (re-captured when the membership changes).

A member gives the same history alone and in any mix: the same kernels do the same
arithmetic per member, and the load reductions have a fixed order.  The sweep tile changes
which thread updates a cell, never what it computes, so `tune` (one 256-thread tile for
every member sweep, the fastest of BATCH_TILES over real steps of the batch) leaves the
histories as they are.
"""
from __future__ import annotations

import ctypes

from . import _lib

_vp = _lib._vp

# 256-thread tiles of the batched sweep: (bx, by), bx = 0 for row-linear blocks
BATCH_TILES = ((64, 4), (128, 2), (256, 1), (0, 256))


class RefinedBatch:
    def __init__(self, storage="fp32", device=0, graphs=True, stream=None, sweep_grid=0):
        import torch
        if not torch.cuda.is_available():
            raise _lib.NativeLibraryError("no HIP device visible to PyTorch-ROCm")
        if storage not in _lib.STORE:
            raise ValueError("storage must be fp64, fp32 or fp16")
        self.L = _lib.load()
        self.storage = storage
        h = _vp()
        st = stream.cuda_stream if stream is not None else None
        _lib.check(self.L.ak_batch_create(ctypes.byref(h), int(device), _lib.STORE[storage], _vp(st)), None,
                   "ak_batch_create")
        self.h = h
        _lib.check(self.L.ak_batch_set_graph_mode(h, 1 if graphs else 0), None, "ak_batch_set_graph_mode")
        # batched sweep launch: 0 one block per tile, n > 0 a persistent grid of n blocks
        _lib.check(self.L.ak_batch_set_sweep_grid(h, int(sweep_grid)), None, "ak_batch_set_sweep_grid")
        self.members = []
        self.tile, self.tile_timings = None, {}

    def add(self, solver):
        """Admit a RefinedSolver (same storage; at a coarse-step boundary)."""
        if solver.storage != self.storage:
            raise ValueError(f"a {solver.storage} case cannot join a {self.storage} batch")
        if self.tile is not None:
            self._apply_tile(solver, self.tile)
        _lib.check(self.L.ak_batch_add(self.h, solver.h), None, "ak_batch_add")
        self.members.append(solver)

    def remove(self, solver):
        _lib.check(self.L.ak_batch_remove(self.h, solver.h), None, "ak_batch_remove")
        self.members.remove(solver)

    def __len__(self):
        return len(self.members)

    def launch(self, nsteps, sample=True):
        """Enqueue nsteps coarse steps of every member and return at once (the GPU runs them
        while the host prepares other work); `collect` waits and reads the results."""
        if nsteps <= 0 or not self.members:
            return
        _lib.check(self.L.ak_batch_run(self.h, int(nsteps), 1 if sample else 0), None, "ak_batch_run")
        for m in self.members:
            m.steps += nsteps

    def collect(self, raw=False):
        """Wait for the enqueued steps.  Returns ({member: records}, {member: TransferError});
        records are dicts, or with raw=True the history rows as an array (step, t*, cl, cd, cm
        in columns 0-3 as recorded).  A member whose transfer guards fired is reported in the
        second mapping and stays in the batch until removed."""
        from .refined import TransferError
        _lib.check(self.L.ak_batch_synchronize(self.h), None, "ak_batch_synchronize")
        out, failed = {}, {}
        for m in self.members:
            rows = m._read()
            out[m] = rows if raw else [m._record(r) for r in rows]
            try:
                m._check_transfer()
            except TransferError as e:
                failed[m] = e
        return out, failed

    def advance(self, nsteps, sample=True):
        """nsteps coarse steps of every member: launch + collect."""
        if nsteps <= 0 or not self.members:
            return {m: [] for m in self.members}, {}
        self.launch(nsteps, sample)
        return self.collect()

    @staticmethod
    def _apply_tile(solver, tile):
        from .autotune import default_tile
        nxs = (solver.plan.shape[0], solver.fine_shape[0] + 2 * solver.pad)
        solver.set_tiles(*[tile if not (tile[0] > 64 and nx < tile[0]) else default_tile(nx) for nx in nxs])

    def tune(self, steps=20, tiles=BATCH_TILES):
        """Time `steps` real coarse steps of the batch with each tile applied to every member
        and keep the fastest (members joining later get it too).  The steps are the members'
        own (sampled every step).  Returns (tile, {tile: ms per batch step}, {member: history
        rows of those steps, as collect(raw=True)})."""
        import time
        import numpy as np
        import torch
        if not self.members:
            return self.tile, {}, {}
        timings, rows = {}, {m: [] for m in self.members}

        def run(n):
            self.launch(n, sample=True)
            got, _ = self.collect(raw=True)
            for m, r in got.items():
                rows[m].append(r)

        for tile in tiles:
            for m in self.members:
                self._apply_tile(m, tile)
            run(2)  # descriptors and graphs of this tile
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(steps)
            torch.cuda.synchronize()
            timings[tile] = (time.perf_counter() - t0) / steps * 1e3
        self.tile = min(timings, key=timings.get)
        for m in self.members:
            self._apply_tile(m, self.tile)
        self.tile_timings = timings
        return self.tile, timings, {m: np.concatenate(r) for m, r in rows.items()}

    def close(self):
        if getattr(self, "h", None):
            self.L.ak_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
