"""One D3Q27 lattice on one GPU: a thin owner of a native handle and its device buffer.

PyTorch-ROCm supplies the population buffer (a caller-owned uint8 tensor whose
data_ptr() is bound to the native lattice) and the HIP stream; every kernel is
in the native library.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class Lattice:
    def __init__(self, nx, ny, nz, *, storage="fp64", omega_shear=1.0, omega_bulk=1.0,
                 fp16_scale=0.0, device=0, history_capacity=4096, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.NativeLibraryError("no HIP device visible to PyTorch-ROCm")
        lib = _lib.load()
        self.lib = lib
        self.shape = (int(nx), int(ny), int(nz))
        self.storage = storage
        self.device = int(device)
        self.torch_device = torch.device("cuda", self.device)
        if stream is None:
            stream = torch.cuda.current_stream(self.torch_device)
        self.stream = stream
        h = _lib._vp()
        rc = lib.ak_lattice_create(_lib.ctypes.byref(h), self.device, *self.shape, _lib.STORE[storage],
                                   float(omega_shear), float(omega_bulk), float(fp16_scale),
                                   int(history_capacity), _lib._vp(stream.cuda_stream))
        _lib.check(rc, None, "ak_lattice_create")
        self.h = h
        self.nbytes = int(lib.ak_population_bytes(h))
        # caller-owned population buffer; 128-byte aligned view
        self._bind(self.new_buffer())
        self.hist_width = int(lib.ak_history_width(h))

    def new_buffer(self):
        """A fresh device buffer that can hold this lattice's populations (128-byte aligned view)."""
        return _torch().empty(self.nbytes + 128, dtype=_torch().uint8, device=self.torch_device)

    def _bind(self, buf):
        """Make `buf` (from new_buffer) the population buffer; its contents start at zero."""
        base = buf.data_ptr()
        offset = (-base) % 128
        _lib.check(self.lib.ak_bind_populations(self.h, _lib._vp(base + offset), self.nbytes), self.h,
                   "ak_bind_populations")
        self._buf, self._offset = buf, offset

    def close(self):
        if getattr(self, "h", None):
            self.lib.ak_lattice_destroy(self.h)
            self.h = None
            self._buf = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup
    def set_solid(self, mask_xy):
        m = np.ascontiguousarray(mask_xy, dtype=np.uint8)
        assert m.shape == self.shape[:2]
        _lib.check(self.lib.ak_set_solid(self.h, _lib.u8ptr(m)), self.h, "ak_set_solid")

    def set_walls(self, links, nelem, rx_off, ry_off, chord_cells, lattice_speed):
        n = len(links["x"])
        x = np.ascontiguousarray(links["x"], dtype=np.int32)
        y = np.ascontiguousarray(links["y"], dtype=np.int32)
        d = np.ascontiguousarray(links["d"], dtype=np.int32)
        q = np.ascontiguousarray(links["q"], dtype=np.float64)
        e = np.ascontiguousarray(links["element"], dtype=np.int32)
        _lib.check(self.lib.ak_set_walls(self.h, n, _lib.iptr(x), _lib.iptr(y), _lib.iptr(d), _lib.dptr(q),
                                         _lib.iptr(e), int(nelem), float(rx_off), float(ry_off),
                                         float(chord_cells), float(lattice_speed)), self.h, "ak_set_walls")

    def set_farfield(self, kind, lattice_speed):
        _lib.check(self.lib.ak_set_farfield(self.h, _lib.FARFIELD[kind], float(lattice_speed)), self.h,
                   "ak_set_farfield")

    def init_seeded(self, p, origin, dx):
        arr = np.array([p["U"], p["amp"], p["phase_y"], p["phase_z"], p["ky"], p["kz"], p["lex"], p["ley"],
                        origin[0], origin[1], dx], dtype=np.float64)
        _lib.check(self.lib.ak_init_seeded(self.h, _lib.dptr(arr)), self.h, "ak_init_seeded")

    def set_graphs(self, enable):
        """Replay each step variant from a captured HIP graph (default) or launch directly."""
        _lib.check(self.lib.ak_set_graph_mode(self.h, 1 if enable else 0), self.h, "ak_set_graph_mode")

    def set_tile(self, bx, by, odd=None):
        """Sweep tile of both AA forms, or with odd=True / False of the odd / even form only."""
        if odd is None:
            _lib.check(self.lib.ak_set_sweep_tile(self.h, int(bx), int(by)), self.h, "ak_set_sweep_tile")
        else:
            _lib.check(self.lib.ak_set_sweep_tile_form(self.h, 1 if odd else 0, int(bx), int(by)), self.h,
                       "ak_set_sweep_tile_form")

    def set_waves(self, waves, odd):
        """Cap the resident waves per SIMD of the odd (odd=True) or even sweep (0: no cap)."""
        _lib.check(self.lib.ak_set_sweep_waves(self.h, 1 if odd else 0, int(waves)), self.h, "ak_set_sweep_waves")

    def populations_written(self):
        """Declare that the bound population buffer was written directly (not through this
        object): derived state kept between run calls is rebuilt at the next run call."""
        _lib.check(self.lib.ak_populations_written(self.h), self.h, "ak_populations_written")

    def next_sweep_odd(self):
        """True when the next sweep is the odd (gathering) AA form."""
        return bool(_lib.check(self.lib.ak_next_sweep_odd(self.h), self.h, "ak_next_sweep_odd"))

    # ---- stepping
    def run(self, nsteps, first_step, sample_every):
        _lib.check(self.lib.ak_run(self.h, int(nsteps), int(first_step), int(sample_every)), self.h, "ak_run")

    def run_profiled(self, nsteps, first_step, sample_every, total=False):
        """Advance like run(); returns (sweep ms, boundary ms) summed over the steps, and with
        total=True also the milliseconds of all the steps (direct launches)."""
        out = np.zeros(4)
        _lib.check(self.lib.ak_run_profiled(self.h, int(nsteps), int(first_step), int(sample_every), _lib.dptr(out)),
                   self.h, "ak_run_profiled")
        return (float(out[0]), float(out[1]), float(out[3])) if total else (float(out[0]), float(out[1]))

    def read_history(self, max_rows=1 << 20):
        out = np.zeros((max_rows, self.hist_width)) if max_rows <= 4096 else None
        if out is None:
            rows = []
            while True:
                chunk = np.zeros((4096, self.hist_width))
                n = _lib.check(self.lib.ak_read_history(self.h, _lib.dptr(chunk), 4096), self.h, "ak_read_history")
                rows.append(chunk[:n])
                if n < 4096:
                    break
            return np.concatenate(rows) if rows else np.zeros((0, self.hist_width))
        n = _lib.check(self.lib.ak_read_history(self.h, _lib.dptr(out), max_rows), self.h, "ak_read_history")
        return out[:n]

    def synchronize(self):
        _lib.check(self.lib.ak_synchronize(self.h), self.h, "ak_synchronize")

    # ---- state access
    def populations(self):
        g = np.zeros((*self.shape, 27))
        _lib.check(self.lib.ak_get_populations(self.h, _lib.dptr(g)), self.h, "ak_get_populations")
        return g

    def set_populations(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.shape == (*self.shape, 27)
        _lib.check(self.lib.ak_set_populations(self.h, _lib.dptr(g)), self.h, "ak_set_populations")

    def diagnostics(self):
        out = np.zeros(5)
        _lib.check(self.lib.ak_diagnostics(self.h, _lib.dptr(out)), self.h, "ak_diagnostics")
        return out

    def macroscopic(self):
        """rho[x][y][z] and u[x][y][z][3] of the interior (FP64 host arrays)."""
        rho = np.zeros(self.shape)
        u = np.zeros((*self.shape, 3))
        _lib.check(self.lib.ak_get_macroscopic(self.h, _lib.dptr(rho), _lib.dptr(u)), self.h, "ak_get_macroscopic")
        return rho, u

    def plane(self, z):
        nx, ny, _ = self.shape
        rho = np.zeros((nx, ny))
        u = np.zeros((nx, ny, 3))
        _lib.check(self.lib.ak_get_plane(self.h, int(z), _lib.dptr(rho), _lib.dptr(u)), self.h, "ak_get_plane")
        return rho, u
