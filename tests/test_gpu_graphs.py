"""Graph replay of the step: identical results to direct launches.

A step is captured as a HIP graph per launch variant (AA parity x sampled, and the
fine parity on the two-level step) and replayed; the step number and history row
live in device memory.  Replays must give bit-identical populations and history
rows to launching every kernel directly, across sampled and unsampled steps, a
history flush in the middle of a run, and a setter (tile change) that invalidates
the captured graphs.
"""
from pathlib import Path

import numpy as np
import pytest

from aerokinetic_amd.config import Flow, Grid, Plan
from aerokinetic_amd.geometry import Section, naca0012
from aerokinetic_amd.refined_plan import RefinedPlan, load_refined_project

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def small_case(farfield):
    section = Section((naca0012(41),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=10000, alpha_deg=4.0, lattice_speed=0.05, perturbation=1e-3)
    grid = Grid(cells_per_chord=16, upstream_chords=1.0, downstream_chords=2.0, transverse_chords=1.0,
                span_chords=0.5, memory_limit_gib=2, farfield=farfield)
    return section, flow, grid, Plan.create(section, flow, grid)


def _rows(recs):
    keys = ("step", "cl", "cd", "cm", "local_mach_max", "density_error_max", "density_min", "nonfinite")
    return np.array([[r[k] for k in keys] for r in recs])


@pytest.mark.parametrize("farfield", ["characteristic_transport_v1", "legacy"])
def test_uniform_graph_replay_matches_direct_launches(farfield):
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = small_case(farfield)
    print("case", farfield, plan.shape, flush=True)
    direct = Solver(section, flow, grid, plan=plan, graphs=False)
    graph = Solver(section, flow, grid, plan=plan, graphs=True)
    for n, every in ((7, 3), (10, 1), (6, 0), (9, 2)):
        a, b = direct.advance(n, every), graph.advance(n, every)
        assert len(a) == len(b)
        if a:
            np.testing.assert_array_equal(_rows(a), _rows(b))
            for ra, rb in zip(a, b):
                np.testing.assert_array_equal(ra["element_lattice_loads"], rb["element_lattice_loads"])
    np.testing.assert_array_equal(direct.populations(), graph.populations())
    # a tile change invalidates the captured graphs; the replay must follow it
    graph.lattice.set_tile(128, 2)
    direct.lattice.set_tile(128, 2)
    a, b = direct.advance(5, 1), graph.advance(5, 1)
    np.testing.assert_array_equal(_rows(a), _rows(b))
    np.testing.assert_array_equal(direct.populations(), graph.populations())
    assert [r["step"] for r in b] == list(range(33, 38))


def test_uniform_graph_history_flush_mid_run():
    """History capacity smaller than the sampled steps of one run: rows are flushed to
    the host mid-run and the device row cursor restarts at zero."""
    from aerokinetic_amd.lattice import Lattice
    from aerokinetic_amd.solver import Solver, seed_parameters
    section, flow, grid, plan = small_case("legacy")
    ref = Solver(section, flow, grid, plan=plan, graphs=False)
    recs_ref = ref.advance(11, 1)
    s = Solver(section, flow, grid, plan=plan, graphs=True)
    small = Lattice(*plan.shape, storage="fp64", omega_shear=plan.relaxation_rate, omega_bulk=1.0,
                    history_capacity=4)
    small.set_solid(s._solid_xy)
    small.set_walls(s.links, len(s.contours), plan.origin[0] / plan.dx, plan.origin[1] / plan.dx,
                    grid.cells_per_chord, flow.lattice_speed)
    small.init_seeded(seed_parameters(section, flow, plan), plan.origin, plan.dx)
    small.set_farfield(grid.farfield, flow.lattice_speed)
    small.run(11, 0, 1)
    rows = small.read_history()
    assert rows.shape[0] == 11
    np.testing.assert_array_equal(rows[:, 0], np.arange(1, 12))
    np.testing.assert_array_equal(rows[:, 1:4], np.array([[r["cl"], r["cd"], r["cm"]] for r in recs_ref]))
    small.close()


def test_refined_graph_replay_matches_direct_launches():
    from aerokinetic_amd.refined import RefinedSolver
    project, settings = load_refined_project(GOLDEN / "refined_fowler_n32" / "project.json")
    plan = RefinedPlan.create(project, settings)
    print("case fowler", plan.shape, plan.box, flush=True)
    direct = RefinedSolver(project, settings, plan=plan, graphs=False)
    graph = RefinedSolver(project, settings, plan=plan, graphs=True)
    for n, every in ((5, 2), (4, 1), (3, 0)):
        a, b = direct.advance(n, every), graph.advance(n, every)
        assert [r["step"] for r in a] == [r["step"] for r in b]
        for ra, rb in zip(a, b):
            for k in ("cl", "cd", "cm", "wall_mass_exchange"):
                assert ra[k] == rb[k], k
    for level in (0, 1):
        np.testing.assert_array_equal(direct.populations(level), graph.populations(level))


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_x_face_cache_single_steps_match_one_run(storage):
    """The characteristic farfield reads the x-face cells from a copy gathered at the start
    of its boundary phase.  Twelve one-step calls and one twelve-step call (graph replays)
    must give bit-identical populations and loads, also across a tile change to row-linear
    sweep blocks."""
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = small_case("characteristic_transport_v1")
    print("case x-face cache", storage, plan.shape, flush=True)
    one = Solver(section, flow, grid, plan=plan, storage=storage, graphs=True)
    many = Solver(section, flow, grid, plan=plan, storage=storage, graphs=True)
    a = one.advance(12, 1)
    b = [many.step(sample=True) for _ in range(12)]
    np.testing.assert_array_equal(_rows(a)[:, 1:4], np.array([[r["cl"], r["cd"], r["cm"]] for r in b]))
    np.testing.assert_array_equal(one.populations(), many.populations())
    one.lattice.set_tile(0, 256)
    many.lattice.set_tile(0, 256)
    one.advance(7, 0)
    for _ in range(7):
        many.step()
    np.testing.assert_array_equal(one.populations(), many.populations())


def test_autotuned_tile_gives_identical_results():
    """tile="auto" times every candidate sweep tile on the case's own lattice, keeps the
    fastest and restarts from the seeded field: the run must equal one on the default tile
    bit for bit (a tile changes only which threads sweep which cells)."""
    from aerokinetic_amd import autotune
    from aerokinetic_amd.solver import Solver
    section, flow, grid, plan = small_case("characteristic_transport_v1")
    print("case autotune", plan.shape, flush=True)
    tuned = Solver(section, flow, grid, plan=plan, tile="auto")
    plain = Solver(section, flow, grid, plan=plan)
    assert len(tuned.tile) == 2 and all(t[:2] in autotune.CANDIDATES for t in tuned.tile)  # (even, odd)
    assert tuned.steps == 0
    a, b = tuned.advance(12, 3), plain.advance(12, 3)
    np.testing.assert_array_equal(_rows(a), _rows(b))
    np.testing.assert_array_equal(tuned.populations(), plain.populations())
