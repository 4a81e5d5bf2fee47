"""Two-level 2:1 case model: refinement settings, acoustic levels, fine box and plan.

Restates the reference's refined case model (src/aerokinetic/refinement.py:23-65,
interface.py:20-41, refined_solver.py:23-112):

* `AcousticLevel`: dx, dt, shear and bulk relaxation times of one level;
  `refined()` halves dx and dt and doubles tau - 1/2 (acoustic scaling, fixed
  physical viscosities);
* `FineBox`: half-open coarse-cell bounds of the fine region;
* `RefinedPlan.create(project, settings)`: the fine box from chord margins around
  the rotated section (floor/ceil on the coarse grid), the wall-to-interface guard
  check, coarse level tau_shear = 1/omega of the uniform plan, tau_bulk from
  the settings;
* `load_refined_project` / `refined_project_dict`: the
  `aerokinetic_refined_project` schema v1.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass
import json
import math
from pathlib import Path

import numpy as np

from .project import Project

CS2 = 1.0 / 3.0


@dataclass(frozen=True)
class AcousticLevel:
    dx: float
    dt: float
    tau_shear: float
    tau_bulk: float

    def __post_init__(self):
        for name in ("dx", "dt", "tau_shear", "tau_bulk"):
            v = getattr(self, name)
            if isinstance(v, (bool, np.bool_)) or not math.isfinite(v) or v <= 0:
                raise ValueError(f"{name} must be finite and positive")
        if self.tau_shear <= 0.5 or self.tau_bulk <= 0.5:
            raise ValueError("Relaxation times must be strictly above 0.5")

    @property
    def lattice_velocity_unit(self):
        return self.dx / self.dt

    @property
    def shear_viscosity(self):
        return CS2 * (self.tau_shear - 0.5) * self.dx**2 / self.dt

    @property
    def bulk_viscosity(self):
        return (2.0 / 3.0) * CS2 * (self.tau_bulk - 0.5) * self.dx**2 / self.dt

    def refined(self):
        return AcousticLevel(self.dx / 2, self.dt / 2, 0.5 + 2 * (self.tau_shear - 0.5), 0.5 + 2 * (self.tau_bulk - 0.5))


@dataclass(frozen=True)
class FineBox:
    lower: tuple
    upper: tuple

    def __post_init__(self):
        for vals in (self.lower, self.upper):
            if len(vals) != 3 or any(isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)) for v in vals):
                raise ValueError("Fine-box bounds must be three integer indices")
        if any(lo < 0 or hi <= lo for lo, hi in zip(self.lower, self.upper)):
            raise ValueError("Fine-box bounds must have 0 <= lower < upper")
        object.__setattr__(self, "lower", tuple(int(v) for v in self.lower))
        object.__setattr__(self, "upper", tuple(int(v) for v in self.upper))

    @property
    def shape(self):
        return tuple(hi - lo for lo, hi in zip(self.lower, self.upper))

    @property
    def slices(self):
        return tuple(slice(lo, hi) for lo, hi in zip(self.lower, self.upper))


@dataclass(frozen=True)
class RefinementSettings:
    upstream_chords: float = 0.375
    downstream_chords: float = 0.75
    transverse_chords: float = 0.375
    guard_coarse_cells: int = 4
    coarse_bulk_relaxation_time: float = 1.0

    def __post_init__(self):
        for name in ("upstream_chords", "downstream_chords", "transverse_chords"):
            v = getattr(self, name)
            if not math.isfinite(v) or v <= 0:
                raise ValueError(f"Refinement {name} must be finite and positive")
        if type(self.guard_coarse_cells) is not int or self.guard_coarse_cells < 4:
            raise ValueError("At least four coarse cells must separate every wall from the fine interface")
        if not math.isfinite(self.coarse_bulk_relaxation_time) or self.coarse_bulk_relaxation_time <= .5:
            raise ValueError("Coarse bulk relaxation time must be finite and above 0.5")


@dataclass(frozen=True)
class RefinedPlan:
    shape: tuple
    origin: tuple
    box: FineBox
    coarse_level: AcousticLevel
    fine_level: AcousticLevel
    span: float
    dt_seconds: float | None
    estimated_memory_gib: float
    minimum_wall_interface_distance_coarse_cells: float
    notes: tuple

    @classmethod
    def create(cls, project: Project, refinement: RefinementSettings, base=None):
        if project.grid.farfield != "legacy":
            raise ValueError("Characteristic farfield is not yet verified on refined grids; use a uniform grid explicitly")
        if project.flow.sgs != "none":
            raise ValueError("Refined SGS is not implemented; select sgs=none explicitly")
        base = base or project.plan()
        pts = np.concatenate([e.xy for e in project.section.normalized(project.flow.alpha_deg)])
        lower, upper = pts.min(axis=0), pts.max(axis=0)
        m_lo = np.array([refinement.upstream_chords, refinement.transverse_chords])
        m_hi = np.array([refinement.downstream_chords, refinement.transverse_chords])
        lo = np.floor((lower - m_lo - base.origin) / base.dx + 1e-12).astype(int)
        hi = np.ceil((upper + m_hi - base.origin) / base.dx - 1e-12).astype(int)
        if np.any(lo < 1) or np.any(hi > np.array(base.shape[:2]) - 1):
            raise ValueError("The refined box must fit strictly inside the external x/y domain")
        dist = np.r_[(lower - (base.origin + lo * base.dx)) / base.dx, ((base.origin + hi * base.dx) - upper) / base.dx]
        if dist.min() < refinement.guard_coarse_cells - 1e-10:
            raise ValueError("Wall/interface guard too small; enlarge the refinement margins")
        box = FineBox((*map(int, lo), 0), (*map(int, hi), base.shape[2]))
        coarse = AcousticLevel(base.dx, base.dt_convective, 1 / base.relaxation_rate, refinement.coarse_bulk_relaxation_time)
        fine = coarse.refined()
        padded = tuple(2 * n + 4 for n in box.shape)
        cells = math.prod(n + 2 for n in base.shape) + math.prod(n + 2 for n in padded)
        memory = (cells * (8 * 27 * 8 + 4 * 10 * 8 + 4 * 8 + 128) + 256 * 2**20) / 2**30
        return cls(base.shape, base.origin, box, coarse, fine, base.span, base.dt_seconds, memory, float(dist.min()),
                   ("Static two-level grid; fine cells/chord = 2 * coarse cells/chord.",
                    "All walls on the fine level; legacy farfield on the coarse level."))

    @classmethod
    def explicit(cls, project: Project, refinement: RefinementSettings, base, lower, upper):
        """Plan with an explicit fine box (half-open coarse bounds in x, y; full span in z)."""
        box = FineBox((int(lower[0]), int(lower[1]), 0), (int(upper[0]), int(upper[1]), base.shape[2]))
        pts = np.concatenate([e.xy for e in project.section.normalized(project.flow.alpha_deg)])
        lo_pt, hi_pt = pts.min(axis=0), pts.max(axis=0)
        org = np.asarray(base.origin)
        dist = np.r_[(lo_pt - (org + np.array(box.lower[:2]) * base.dx)) / base.dx,
                     ((org + np.array(box.upper[:2]) * base.dx) - hi_pt) / base.dx]
        if dist.min() < refinement.guard_coarse_cells - 1e-10:
            raise ValueError("Wall/interface guard too small; enlarge the refinement box")
        coarse = AcousticLevel(base.dx, base.dt_convective, 1 / base.relaxation_rate, refinement.coarse_bulk_relaxation_time)
        padded = tuple(2 * n + 4 for n in box.shape)
        cells = math.prod(n + 2 for n in base.shape) + math.prod(n + 2 for n in padded)
        memory = (cells * (8 * 27 * 8 + 4 * 10 * 8 + 4 * 8 + 128) + 256 * 2**20) / 2**30
        return cls(base.shape, base.origin, box, coarse, coarse.refined(), base.span, base.dt_seconds, memory,
                   float(dist.min()), ("Static two-level grid with an explicit fine box.",))

    @property
    def fine_shape(self):
        return tuple(2 * n for n in self.box.shape)

    def as_dict(self):
        d = asdict(self)
        d.update(dx=self.coarse_level.dx, dt_convective=self.coarse_level.dt,
                 coarse_cells_per_chord=round(1 / self.coarse_level.dx),
                 fine_cells_per_chord=round(1 / self.fine_level.dx), fine_shape=self.fine_shape)
        return d

    @classmethod
    def from_dict(cls, d):
        """The plan that as_dict() described (the derived keys are ignored)."""
        return cls(tuple(d["shape"]), tuple(d["origin"]), FineBox(tuple(d["box"]["lower"]), tuple(d["box"]["upper"])),
                   AcousticLevel(**d["coarse_level"]), AcousticLevel(**d["fine_level"]), d["span"], d["dt_seconds"],
                   d["estimated_memory_gib"], d["minimum_wall_interface_distance_coarse_cells"], tuple(d["notes"]))


def refined_project_dict(project, refinement):
    return {"format": "aerokinetic_refined_project", "schema_version": 1,
            "project": project.as_dict(), "refinement": asdict(refinement)}


def load_refined_project(path):
    path = Path(path)
    data = json.loads(path.read_text(encoding="utf-8-sig"))
    if data.get("format") != "aerokinetic_refined_project" or data.get("schema_version") != 1:
        raise ValueError("This command requires an aerokinetic_refined_project, not a uniform project")
    return Project.from_dict(data["project"], path.resolve().parent), RefinementSettings(**data["refinement"])


def virtual_core(box_shape):
    """Covered coarse cells more than three rows inside the box in x and y (refined_solver.py:137-145)."""
    core = np.ones(box_shape, bool)
    for axis in (0, 1):
        band = np.zeros(box_shape[axis], bool)
        band[3:-3] = True
        shp = [1, 1, 1]
        shp[axis] = len(band)
        core &= band.reshape(shp)
    return core
