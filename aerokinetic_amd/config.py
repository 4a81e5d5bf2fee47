"""Flow, grid and plan: the physical case mapped onto a uniform lattice.

Field names, defaults and validation follow the reference case model
(`src/aerokinetic/config.py:11-166`) so that project JSON files and their
sha256 case fingerprints are interchangeable. The lattice relations are the
reference's (config.py:119-163):

    shape  = ceil((extent + margins) * N),  nz = max(8, ceil(span * N))
    nu     = U * N / Re,  tau = 1/2 + 3 nu,  omega = 1/tau,  dt* = U / N

`Plan.explicit` builds a plan from an explicit lattice shape and leading-edge
cell (the form used by fixed benchmark workloads) instead of chord margins.
"""
from __future__ import annotations

import math
from dataclasses import asdict, dataclass, field

import numpy as np

from .geometry import Section

EXTERIOR_FARFIELD = "exterior_regularized_fd_v1"
CHARACTERISTIC_FARFIELD = "characteristic_transport_v1"
FARFIELDS = ("legacy", CHARACTERISTIC_FARFIELD, EXTERIOR_FARFIELD)


@dataclass(frozen=True)
class Atmosphere:
    """Dry ISA sea-level air; only needed for dimensional seconds."""

    temperature_k: float = 288.15
    pressure_pa: float = 101325.0
    gas_constant: float = 287.05287
    dynamic_viscosity: float = 1.7894e-5

    def __post_init__(self):
        for v in asdict(self).values():
            if not math.isfinite(v) or v <= 0:
                raise ValueError("Atmospheric quantities must be finite and positive")

    @property
    def rho(self):
        return self.pressure_pa / (self.gas_constant * self.temperature_k)

    @property
    def nu(self):
        return self.dynamic_viscosity / self.rho

    @property
    def speed_of_sound(self):
        return math.sqrt(1.4 * self.gas_constant * self.temperature_k)


@dataclass(frozen=True)
class Flow:
    reynolds: float = 10000.0
    alpha_deg: float = 0.0
    lattice_speed: float = 0.04
    physical_chord_m: float | None = None
    atmosphere: Atmosphere = field(default_factory=Atmosphere)
    sgs: str = "none"
    smagorinsky_constant: float = 0.10
    perturbation: float = 1e-4
    seed: int = 31

    def __post_init__(self):
        if not math.isfinite(self.reynolds) or not 10000 <= self.reynolds <= 100000:
            raise ValueError("This module targets 10000 <= Re <= 100000")
        if not math.isfinite(self.alpha_deg) or abs(self.alpha_deg) >= 90:
            raise ValueError("Use an angle strictly between -90 and 90 degrees")
        if not 0 < self.lattice_speed <= 0.08:
            raise ValueError("Lattice speed must be in (0,0.08]; local Mach is monitored")
        if self.sgs not in ("none", "smagorinsky"):
            raise ValueError("sgs must be none or smagorinsky")
        if not 0 < self.smagorinsky_constant < 0.3:
            raise ValueError("Smagorinsky constant must be in (0,0.3)")
        if not math.isfinite(self.perturbation) or not 0 <= self.perturbation <= 0.01:
            raise ValueError("Initial perturbation must be in [0,0.01] relative to U")
        if self.physical_chord_m is not None:
            if not math.isfinite(self.physical_chord_m) or self.physical_chord_m <= 0:
                raise ValueError("Physical chord must be positive")
            if self.physical_mach > 0.15:
                raise ValueError("Physical Mach exceeds 0.15: this low-Mach model is unsuitable")

    @property
    def physical_speed(self):
        if self.physical_chord_m is None:
            return None
        return self.reynolds * self.atmosphere.nu / self.physical_chord_m

    @property
    def physical_mach(self):
        v = self.physical_speed
        return None if v is None else v / self.atmosphere.speed_of_sound


@dataclass(frozen=True)
class Grid:
    cells_per_chord: int = 128
    upstream_chords: float = 6.0
    downstream_chords: float = 12.0
    transverse_chords: float = 6.0
    span_chords: float = 0.5
    memory_limit_gib: float = 8.0
    target: str = "cpu"
    openmp_threads: int = 1
    farfield: str = "legacy"

    def __post_init__(self):
        if not isinstance(self.cells_per_chord, int) or self.cells_per_chord < 16:
            raise ValueError("cells_per_chord must be an integer >=16")
        if self.target not in ("cpu", "gpu"):
            raise ValueError("target must be cpu or gpu")
        for key in ("upstream_chords", "downstream_chords", "transverse_chords",
                    "span_chords", "memory_limit_gib"):
            v = getattr(self, key)
            if not math.isfinite(v) or v <= 0:
                raise ValueError(f"{key} must be finite and positive")
        if type(self.openmp_threads) is not int or self.openmp_threads < 1:
            raise ValueError("openmp_threads must be a positive integer")
        if self.farfield not in FARFIELDS:
            raise ValueError("farfield must be one of " + ", ".join(FARFIELDS))
        if self.farfield != "legacy" and self.target != "cpu":
            raise ValueError("Experimental characteristic farfield is currently verified for CPU only")


@dataclass(frozen=True)
class Plan:
    shape: tuple
    origin: tuple
    dx: float
    span: float
    viscosity: float
    relaxation_rate: float
    dt_convective: float
    dt_seconds: float | None
    estimated_memory_gib: float
    notes: tuple

    @staticmethod
    def _lattice(flow: Flow, n: int):
        nu = flow.lattice_speed * n / flow.reynolds
        tau = 0.5 + 3 * nu
        dt = flow.lattice_speed / n
        dt_s = None if flow.physical_speed is None else dt * flow.physical_chord_m / flow.physical_speed
        return nu, tau, dt, dt_s

    @staticmethod
    def _memory(shape, farfield):
        # Same estimate the reference reports (two FP64 population copies,
        # macroscopic fields, flags and headroom); it is not an allocation.
        exterior = farfield == EXTERIOR_FARFIELD
        storage = tuple(v + (2 if a < 2 else 0) for a, v in enumerate(shape)) if exterior else shape
        extra = 4 * 4 * 8 if exterior else 0
        return math.prod(v + 2 for v in storage) * ((2 * 27 * 8 + 4 * 8 + 8 + extra) * 1.3 + 128) / 2**30

    @staticmethod
    def _storage(shape):
        """Lattice of the exterior farfield: the physical box padded by one cell in x and y."""
        return (shape[0] + 2, shape[1] + 2, shape[2])

    @staticmethod
    def _notes(flow, grid, n, tau, exterior_shape=None):
        notes = ["Uniform isotropic grid; no mesh refinement or wall model.",
                 "Accuracy is unvalidated until grid/domain/span/Mach/statistical convergence and tunnel comparison.",
                 "The initial spanwise perturbation is numerical seeding, not a specified tunnel turbulence spectrum."]
        if grid.farfield != "legacy":
            if min(grid.upstream_chords, grid.downstream_chords, grid.transverse_chords) * n < 2:
                raise ValueError("Characteristic farfield needs at least two fluid cells between body bounds and exterior")
            if grid.farfield == EXTERIOR_FARFIELD:
                if flow.sgs != "none":
                    raise ValueError("Exterior FD currently supports molecular viscosity with sgs=none only")
                notes.append(f"Research exterior_regularized_fd_v1; physical shape unchanged, auxiliary storage "
                             f"shape={exterior_shape}; explicit exterior source. Uniform grids only; no validated "
                             "airfoil accuracy or full viscous NSCBC.")
            else:
                notes.append("Experimental characteristic_transport_v1 farfield with linearized transverse transport.")
        if n < 256:
            notes.append("Fewer than 256 cells/chord: use for development; do not assume a resolved boundary layer.")
        if tau < 0.501:
            notes.append("Relaxation time is close to 0.5; low-viscosity stability and resolution require special checks.")
        if flow.physical_speed is None:
            notes.append("Physical chord unspecified; time is t*=t U/c only, with no invented dimensional seconds.")
        else:
            atm = flow.atmosphere
            mfp = atm.dynamic_viscosity / atm.pressure_pa * math.sqrt(math.pi * atm.gas_constant * atm.temperature_k / 2)
            if mfp / flow.physical_chord_m > 1e-3:
                raise ValueError("Kn > 1e-3: the no-slip low-Kn model needs reconsideration")
        return tuple(notes)

    @classmethod
    def create(cls, section: Section, flow: Flow, grid: Grid):
        pts = np.concatenate([e.xy for e in section.normalized(flow.alpha_deg)])
        xmin, ymin = pts.min(axis=0)
        xmax, ymax = pts.max(axis=0)
        n = grid.cells_per_chord
        shape = (math.ceil((xmax - xmin + grid.upstream_chords + grid.downstream_chords) * n),
                 math.ceil((ymax - ymin + 2 * grid.transverse_chords) * n),
                 max(8, math.ceil(grid.span_chords * n)))
        # Rounding slack is split evenly on both transverse sides.
        origin = (xmin - grid.upstream_chords, (ymin + ymax) / 2 - shape[1] / (2 * n))
        nu, tau, dt, dt_s = cls._lattice(flow, n)
        notes = cls._notes(flow, grid, n, tau, cls._storage(shape))
        return cls(shape, origin, 1 / n, shape[2] / n, nu, 1 / tau, dt, dt_s,
                   cls._memory(shape, grid.farfield), notes)

    @classmethod
    def explicit(cls, section: Section, flow: Flow, grid: Grid, shape, leading_edge_cell):
        """Plan on a given lattice shape with the rotated leading edge at a given cell corner.

        `leading_edge_cell` = (i, j) in lattice units from the domain's lower-left corner.
        """
        shape = tuple(int(v) for v in shape)
        if len(shape) != 3 or min(shape) < 4:
            raise ValueError("Explicit shape needs three sizes >= 4")
        n = grid.cells_per_chord
        le = section.leading_edge_normalized(flow.alpha_deg)
        origin = (float(le[0] - leading_edge_cell[0] / n), float(le[1] - leading_edge_cell[1] / n))
        nu, tau, dt, dt_s = cls._lattice(flow, n)
        notes = cls._notes(flow, grid, n, tau, cls._storage(shape)) + (
            f"Explicit lattice {shape} with leading edge at cell {tuple(leading_edge_cell)}.",)
        return cls(shape, origin, 1 / n, shape[2] / n, nu, 1 / tau, dt, dt_s,
                   cls._memory(shape, grid.farfield), notes)

    def as_dict(self):
        return asdict(self)

    @classmethod
    def from_dict(cls, d):
        """The plan that as_dict() described (JSON lists back to tuples)."""
        d = dict(d)
        d["shape"], d["origin"], d["notes"] = tuple(d["shape"]), tuple(d["origin"]), tuple(d["notes"])
        return cls(**d)
