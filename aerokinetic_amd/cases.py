"""Fixed-lattice airfoil workloads (explicit shapes instead of chord margins).

A workload names its lattice directly: shape (nx, ny, nz), chord in cells, angle,
Reynolds and Mach numbers, and the cell of the rotated leading edge.  Mach sets
the lattice speed U = Ma * c_s = Ma / sqrt(3).  The flow, grid and plan objects
are the package's ordinary `Flow`, `Grid`, `Plan`, so every solver accepts them.
"""
from __future__ import annotations

from dataclasses import dataclass
import math

from .config import Flow, Grid, Plan
from .geometry import Section, naca4, naca4_profile


@dataclass(frozen=True)
class UniformCase:
    name: str
    shape: tuple
    chord_cells: int
    alpha_deg: float
    reynolds: float
    mach: float
    leading_edge_cell: tuple
    naca: str = "0012"
    closed_te: bool = True
    farfield: str = "characteristic_transport_v1"
    perturbation: float = 1e-4
    seed: int = 31
    n_side: int = 161

    @property
    def lattice_speed(self):
        return self.mach / math.sqrt(3.0)

    def build(self):
        """(section, flow, grid, plan) of this workload."""
        section = Section((naca4(self.naca, self.n_side, self.closed_te),), 1.0, (0.25, 0.0))
        flow = Flow(reynolds=self.reynolds, alpha_deg=self.alpha_deg, lattice_speed=self.lattice_speed,
                    perturbation=self.perturbation, seed=self.seed)
        nx, ny, nz = self.shape
        n = self.chord_cells
        grid = Grid(cells_per_chord=n, upstream_chords=self.leading_edge_cell[0] / n,
                    downstream_chords=max((nx - self.leading_edge_cell[0]) / n - 1.0, 0.5),
                    transverse_chords=max(self.leading_edge_cell[1] / n - 0.1, 0.5),
                    span_chords=nz / n, memory_limit_gib=1e6, farfield=self.farfield)
        plan = Plan.explicit(section, flow, grid, self.shape, self.leading_edge_cell)
        return section, flow, grid, plan

    def with_shape(self, shape, leading_edge_cell=None):
        return UniformCase(self.name + f"[{shape[0]}x{shape[1]}x{shape[2]}]", tuple(shape), self.chord_cells,
                           self.alpha_deg, self.reynolds, self.mach, leading_edge_cell or self.leading_edge_cell,
                           self.naca, self.closed_te, self.farfield, self.perturbation, self.seed, self.n_side)


# Plain uniform NACA0012 case: 384x256x8, chord 32 at 4 deg, leading edge at (128, 128),
# Re 1e4, Ma 0.1, closed trailing edge, characteristic farfield, seeded perturbation 1e-4.
NACA0012_C32_384x256x8 = UniformCase("naca0012 c32 384x256x8 a4 Re1e4 Ma0.1", (384, 256, 8), 32, 4.0, 1e4, 0.1,
                                     (128, 128))
# The same plain case with the exterior farfield (exterior_regularized_fd_v1): the lattice is the
# 384x256x8 box padded by one cell in x and y.
NACA0012_C32_384x256x8_EXTERIOR = UniformCase("naca0012 c32 384x256x8 a4 Re1e4 Ma0.1 exterior", (384, 256, 8), 32,
                                              4.0, 1e4, 0.1, (128, 128), farfield="exterior_regularized_fd_v1")
# Precision-study lattice: 1536x1024x64 (~100M cells), chord 128 at 8 deg, Re 1e5, Ma 0.1.
NACA0012_C128_1536x1024x64 = UniformCase("naca0012 c128 1536x1024x64 a8 Re1e5 Ma0.1", (1536, 1024, 64), 128, 8.0,
                                         1e5, 0.1, (512, 512))


def sampled_naca_section(camber, thickness, position=0.4, n_side=161, closed_te=True):
    """NACA 4-digit section with real-valued camber / thickness fractions."""
    return Section((naca4_profile(camber, position, thickness, n_side, closed_te, "NACA"),), 1.0, (0.25, 0.0))


@dataclass(frozen=True)
class RefinedCase:
    """Two-level workload: coarse lattice, explicit fine box, legacy coarse farfield."""
    name: str
    coarse: UniformCase
    box_lower: tuple
    box_upper: tuple
    coarse_bulk_tau: float = 1.0

    def build(self):
        """(project, settings, plan) of this workload."""
        from .project import Project, RunSettings
        from .refined_plan import RefinedPlan, RefinementSettings
        section, flow, grid, base = self.coarse.build()
        project = Project(section, flow, grid, RunSettings(), self.name)
        settings = RefinementSettings(coarse_bulk_relaxation_time=self.coarse_bulk_tau)
        plan = RefinedPlan.explicit(project, settings, base, self.box_lower, self.box_upper)
        return project, settings, plan


# Two-level case: coarse 256x192x8 (chord 32 coarse = 64 fine cells) with a fine patch of
# 160x96x16 fine cells (80x48x8 coarse) around a NACA0012 at 6 deg, Re 3e4, Ma 0.1.
NACA0012_REFINED_256x192x8 = RefinedCase(
    "naca0012 two-level coarse 256x192x8 fine 160x96x16 c64f a6 Re3e4 Ma0.1",
    UniformCase("coarse", (256, 192, 8), 32, 6.0, 3e4, 0.1, (80, 96), farfield="legacy"),
    (62, 72), (142, 120))
