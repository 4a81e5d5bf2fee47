"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module.  It builds an oracle case from the same inputs the reference's
`Solver` takes (section, flow, grid, plan) and evaluates the reference's
seeded initial field parameters itself (src/aerokinetic/solver.py:157-202).
"""
from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path
import subprocess

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libak_oracle.so"

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.ako_create.restype = ctypes.c_void_p
        L.ako_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.ako_destroy.argtypes = [ctypes.c_void_p]
        L.ako_set_geometry.restype = ctypes.c_int
        L.ako_set_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_double_p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.ako_link_count.restype = ctypes.c_int
        L.ako_link_count.argtypes = [ctypes.c_void_p]
        L.ako_fallback_links.restype = ctypes.c_long
        L.ako_fallback_links.argtypes = [ctypes.c_void_p]
        L.ako_get_links.argtypes = [ctypes.c_void_p] + [_c_int_p] * 4 + [_c_double_p, _c_int_p]
        L.ako_get_solid.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8)]
        L.ako_set_load_frame.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4
        L.ako_init_seeded.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 11
        L.ako_set_pdf.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ako_get_pdf.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ako_set_farfield.restype = ctypes.c_int
        L.ako_set_farfield.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ako_step.argtypes = [ctypes.c_void_p, _c_double_p, _c_double_p]
        L.ako_diagnostics.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ako_threads.restype = ctypes.c_int
        L.ako_set_threads.argtypes = [ctypes.c_int]
        for name in ("ako_collide", "ako_boundaries", "ako_stream"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.ako_measure.argtypes = [ctypes.c_void_p, _c_int_p, _c_int_p, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _c_double_p]
        L.ako_get_cells.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_double_p]
        L.ako_set_cells.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_double_p]
        L.ako_collided.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ako_fill_all.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ako_get_padded.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ako_wall_loads.argtypes = [ctypes.c_void_p, _c_double_p, _c_double_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_c_double_p)


def _ip(a):
    return a.ctypes.data_as(_c_int_p)


FARFIELD_KIND = {"periodic": 0, "legacy": 1, "characteristic_transport_v1": 2, "reservoir": 3}


def seed_parameters(section, flow, plan):
    """Scalars of the reference's seeded initial velocity field."""
    rng = np.random.default_rng(flow.seed)
    phase_y, phase_z = rng.uniform(0, 2 * np.pi, 2)
    pts = np.concatenate([e.xy for e in section.elements])
    lead = pts[np.argmin(pts[:, 0])]
    ang = -np.deg2rad(flow.alpha_deg)
    rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    lead = rot @ ((lead - np.asarray(section.moment_reference)) / section.chord)
    ky = 2 * np.pi / (plan.shape[1] * plan.dx)
    kz = 2 * np.pi / plan.span
    amp = flow.perturbation * flow.lattice_speed
    return dict(U=flow.lattice_speed, amp=amp, phase_y=float(phase_y), phase_z=float(phase_z),
                ky=ky, kz=kz, lex=float(lead[0]), ley=float(lead[1]))


class Oracle:
    """One uniform FP64 lattice advanced with the reference's step order."""

    def __init__(self, nx, ny, nz, omega_s, omega_b=1.0):
        self.L = lib()
        self.shape = (nx, ny, nz)
        self.h = self.L.ako_create(nx, ny, nz, omega_s, omega_b)
        if not self.h:
            raise ValueError("oracle lattice too small")
        self.nelem = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ako_destroy(self.h)
            self.h = None

    def set_geometry(self, contours, origin, dx):
        npts = np.array([len(c.xy) for c in contours], dtype=np.int32)
        xy = np.ascontiguousarray(np.concatenate([c.xy for c in contours]), dtype=np.float64)
        n = self.L.ako_set_geometry(self.h, len(contours), _ip(npts), _dp(xy), origin[0], origin[1], dx)
        if n < 0:
            raise ValueError("A fluid-to-solid link has no geometric intersection")
        self.nelem = len(contours)
        return n

    def links(self):
        n = self.L.ako_link_count(self.h)
        cols = [np.zeros(n, dtype=np.int32) for _ in range(4)]
        q = np.zeros(n)
        el = np.zeros(n, dtype=np.int32)
        self.L.ako_get_links(self.h, *[_ip(c) for c in cols], _dp(q), _ip(el))
        return dict(x=cols[0], y=cols[1], z=cols[2], d=cols[3], q=q, element=el)

    def fallback_links(self):
        return int(self.L.ako_fallback_links(self.h))

    def solid(self):
        nx, ny, _ = self.shape
        m = np.zeros((nx, ny), dtype=np.uint8)
        self.L.ako_get_solid(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return m.astype(bool)

    def set_load_frame(self, rx_off, ry_off, chord_cells, U):
        self.L.ako_set_load_frame(self.h, rx_off, ry_off, chord_cells, U)

    def init_seeded(self, p, origin, dx):
        self.L.ako_init_seeded(self.h, p["U"], p["amp"], p["phase_y"], p["phase_z"], p["ky"], p["kz"],
                               p["lex"], p["ley"], origin[0], origin[1], dx)

    def set_pdf(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.shape == (*self.shape, 27)
        self.L.ako_set_pdf(self.h, _dp(g))

    def pdf(self):
        g = np.zeros((*self.shape, 27))
        self.L.ako_get_pdf(self.h, _dp(g))
        return g

    def fill_all(self, g27):
        """Every cell of both arrays, ghosts included, to one population vector."""
        g27 = np.ascontiguousarray(g27, dtype=np.float64)
        assert g27.shape == (27,)
        self.L.ako_fill_all(self.h, _dp(g27))

    def padded_pdf(self):
        """The current array with its ghost layer, (nx+2, ny+2, nz+2, 27)."""
        g = np.zeros((*(n + 2 for n in self.shape), 27))
        self.L.ako_get_padded(self.h, _dp(g))
        return g

    def set_farfield(self, kind):
        return self.L.ako_set_farfield(self.h, FARFIELD_KIND[kind])

    def step(self, sample=False):
        if not sample:
            self.L.ako_step(self.h, None, None)
            return None
        loads = np.zeros((max(self.nelem, 1), 4))
        coef = np.zeros(3)
        self.L.ako_step(self.h, _dp(loads), _dp(coef))
        return {"cl": coef[0], "cd": coef[1], "cm": coef[2], "element_lattice_loads": loads[: self.nelem]}

    # collide_stream building blocks (exterior farfield, refined levels)
    def collide(self):
        self.L.ako_collide(self.h)

    def boundaries(self):
        self.L.ako_boundaries(self.h)

    def stream(self):
        self.L.ako_stream(self.h)

    def cells(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        g = np.zeros((len(xyz), 27))
        self.L.ako_get_cells(self.h, len(xyz), _ip(xyz), _dp(g))
        return g

    def set_cells(self, xyz, g):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.shape == (len(xyz), 27)
        self.L.ako_set_cells(self.h, len(xyz), _ip(xyz), _dp(g))

    def collided(self):
        """The state after one more collision (the stored state is left as it is)."""
        g = np.zeros((*self.shape, 27))
        self.L.ako_collided(self.h, _dp(g))
        return g

    def wall_loads(self):
        """Per-element [Fx, Fy, Fz, Mz] and the wall mass exchange of the current state."""
        loads = np.zeros((max(self.nelem, 1), 4))
        mass = np.zeros(1)
        self.L.ako_wall_loads(self.h, _dp(loads), _dp(mass))
        return loads[: self.nelem], float(mass[0])

    def diagnostics(self):
        out = np.zeros(4)
        self.L.ako_diagnostics(self.h, _dp(out))
        return {"local_mach_max": out[0], "density_error_max": out[1], "density_min": out[2],
                "finite": bool(out[3])}

    @staticmethod
    def threads():
        return lib().ako_threads()

    @staticmethod
    def set_threads(n):
        """OpenMP threads of the oracle's lattice kernels (process-wide)."""
        lib().ako_set_threads(int(n))


def from_case(section, flow, grid, plan, farfield=None):
    """Oracle lattice for a uniform airfoil case, initialised like the reference."""
    ff = farfield or grid.farfield
    nx, ny, nz = plan.shape
    o = Oracle(nx, ny, nz, plan.relaxation_rate, 1.0)
    contours = section.normalized(flow.alpha_deg)
    o.set_geometry(contours, plan.origin, plan.dx)
    o.set_load_frame(plan.origin[0] / plan.dx, plan.origin[1] / plan.dx, grid.cells_per_chord, flow.lattice_speed)
    o.init_seeded(seed_parameters(section, flow, plan), plan.origin, plan.dx)
    o.set_farfield(ff)
    return o
