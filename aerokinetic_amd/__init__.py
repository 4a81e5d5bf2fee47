"""MI355X-native D3Q27 cumulant lattice-Boltzmann step for extruded airfoil sections.

Drop-in for the reference package's case-JSON -> Solver step/run API and its
force-history outputs.  Numerics run in hand-written gfx950 HIP kernels
(`aerokinetic_amd/csrc`, loaded through a C ABI); PyTorch-ROCm provides device
buffers and streams only.
"""
from .geometry import Contour, Section, naca0012, naca4, naca4_profile
from .config import Atmosphere, Flow, Grid, Plan
from .project import Project, RunSettings
from .convergence import StopCriteria

__version__ = "0.5.3+mi355x.1"
__all__ = ["Contour", "Section", "naca0012", "naca4", "naca4_profile", "Atmosphere", "Flow", "Grid", "Plan",
           "Project", "RunSettings", "StopCriteria", "Solver", "NumericalFailure", "coefficients", "RefinedSolver",
           "RefinedPlan", "RefinementSettings"]


def __getattr__(name):
    # the solvers load the HIP library (and torch) on first use only
    if name in ("Solver", "NumericalFailure", "coefficients"):
        from . import solver
        return getattr(solver, name)
    if name == "RefinedSolver":
        from .refined import RefinedSolver
        return RefinedSolver
    if name in ("RefinedPlan", "RefinementSettings"):
        from . import refined_plan
        return getattr(refined_plan, name)
    raise AttributeError(name)
