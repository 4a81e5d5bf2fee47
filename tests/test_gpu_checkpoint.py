"""Device checkpoints: a restored case continues bit-identically (populations in every
storage precision, AA parity, boundary histories), and an interrupted run resumed from
its checkpoint writes the same history as an uninterrupted one."""
import csv

import numpy as np
import pytest

from aerokinetic_amd.config import Flow, Grid, Plan
from aerokinetic_amd.geometry import Section, naca0012

pytestmark = pytest.mark.gpu


def small_case(farfield):
    section = Section((naca0012(41),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=10000, alpha_deg=4.0, lattice_speed=0.05, perturbation=1e-3)
    grid = Grid(cells_per_chord=16, upstream_chords=1.0, downstream_chords=2.0, transverse_chords=1.0,
                span_chords=0.5, memory_limit_gib=2, farfield=farfield)
    return section, flow, grid


def _rows(recs):
    return np.array([[r["step"], r["cl"], r["cd"], r["cm"], r["local_mach_max"]] for r in recs])


@pytest.mark.parametrize("farfield,storage", [("characteristic_transport_v1", "fp64"), ("legacy", "fp16"),
                                              ("characteristic_transport_v1", "fp32")])
def test_uniform_state_roundtrip(tmp_path, farfield, storage):
    from aerokinetic_amd.checkpoint import restore_checkpoint
    from aerokinetic_amd.solver import Solver
    section, flow, grid = small_case(farfield)
    print("case", farfield, storage, flush=True)
    a = Solver(section, flow, grid, storage=storage)
    a.advance(7, 0)  # odd number of sweeps: the other AA form
    a.save_checkpoint(tmp_path / "c.npz", {"note": "x"})
    b = restore_checkpoint(Solver, tmp_path / "c.npz")
    assert b.steps == 7 and b._checkpoint_context == {"note": "x"}
    ra, rb = a.advance(9, 1), b.advance(9, 1)
    np.testing.assert_array_equal(_rows(ra), _rows(rb))
    np.testing.assert_array_equal(a.populations(), b.populations())


def test_refined_state_roundtrip(tmp_path):
    from aerokinetic_amd import serving as S
    from aerokinetic_amd.checkpoint import restore_checkpoint
    from aerokinetic_amd.refined import RefinedSolver
    project, settings, plan = S.project_of(S.CaseRequest(0, 0.0, 0.02, 0.4, 0.12, 5.0, 2e4, 16, 8))
    a = RefinedSolver(project, settings, plan=plan, storage="fp32")
    a.advance(5, 1)
    a.save_checkpoint(tmp_path / "r.npz")
    b = restore_checkpoint(RefinedSolver, tmp_path / "r.npz")
    assert b.steps == 5
    ra, rb = a.advance(6, 1), b.advance(6, 1)
    assert [(r["step"], r["cl"], r["cd"]) for r in ra] == [(r["step"], r["cl"], r["cd"]) for r in rb]
    for level in (0, 1):
        np.testing.assert_array_equal(a.populations(level), b.populations(level))


def test_interrupted_run_resumes_to_the_same_history(tmp_path):
    from aerokinetic_amd.checkpoint import restore_checkpoint
    from aerokinetic_amd.solver import Solver
    section, flow, grid = small_case("legacy")
    kw = dict(sample_every=5, frame_every=50, check_every=25, checkpoint_every=40)
    full = Solver(section, flow, grid)
    dur = 130 * full.plan.dt_convective
    s1 = full.run(dur, tmp_path / "full", **kw)
    part = Solver(section, flow, grid)
    stop_at = {"step": 0}

    def cb(solver, rec):
        return rec["step"] < 70  # stop after the checkpoint at step 40, mid-way between checkpoints

    s2 = part.run(dur, tmp_path / "part", callback=cb, **kw)
    assert s2["status"] == "stopped"
    resumed = restore_checkpoint(Solver, tmp_path / "part" / "checkpoint.npz")
    s3 = resumed.run(dur, tmp_path / "part", resume=True, **kw)
    assert s3["status"] == "completed" and s3["steps"] == s1["steps"]
    read = lambda p, f="history.csv": list(csv.reader(open(p / f)))
    assert read(tmp_path / "full") == read(tmp_path / "part")
    assert read(tmp_path / "full", "residuals.csv") == read(tmp_path / "part", "residuals.csv")
    assert s3["mean"] == s1["mean"]


def test_interrupted_refined_run_resumes_to_the_same_history(tmp_path):
    from dataclasses import replace
    from aerokinetic_amd import serving as S
    from aerokinetic_amd.checkpoint import restore_checkpoint
    from aerokinetic_amd.project import RunSettings
    from aerokinetic_amd.refined import RefinedSolver
    project, settings, plan = S.project_of(S.CaseRequest(0, 0.0, 0.0, 0.4, 0.12, 4.0, 1.5e4, 16, 8))
    dur = 60 * plan.coarse_level.dt
    project = replace(project, run=RunSettings(duration_convective=dur, sample_every=3, check_every=10,
                                               checkpoint_every=20))
    s1 = RefinedSolver(project, settings, plan=plan).run(tmp_path / "full")
    part = RefinedSolver(project, settings, plan=plan)
    s2 = part.run(tmp_path / "part", callback=lambda solver, rec: rec["step"] < 30)
    assert s2["status"] == "stopped"
    resumed = restore_checkpoint(RefinedSolver, tmp_path / "part" / "checkpoint.npz", plan=plan)
    assert resumed.steps == 30
    s3 = resumed.run(tmp_path / "part", resume=True)
    assert s3["status"] == "completed" and s3["steps"] == s1["steps"] == 60
    read = lambda p, f="history.csv": list(csv.reader(open(p / f)))
    assert read(tmp_path / "full") == read(tmp_path / "part")
    assert read(tmp_path / "full", "residuals.csv") == read(tmp_path / "part", "residuals.csv")


def test_residuals_match_host_restatement():
    """Device residual norms = the reference's definition on the exported state
    (convergence.py:55-96): d = (q1 - q0)/dt, q = (rho, j/U) on owned fluid cells."""
    from aerokinetic_amd.d3q27 import CX, CY, CZ
    from aerokinetic_amd.solver import Solver
    section, flow, grid = small_case("characteristic_transport_v1")
    s = Solver(section, flow, grid)
    s.advance(5, 0)
    fluid = ~s._solid_xy

    def q(g):
        out = np.stack([1.0 + g.sum(-1), g @ CX, g @ CY, g @ CZ], axis=-1)
        out[..., 1:] /= flow.lattice_speed
        return out[fluid]

    q0 = q(s.populations())
    s.residual_norms(reset=True)
    s.advance(7, 0)
    dt = 7 * s.residual_dt
    l2, linf = s.residual_norms(dt=dt)
    d = (q(s.populations()) - q0) / dt
    np.testing.assert_allclose(l2, np.sqrt((d * d).mean(axis=(0, 1))), rtol=1e-9)
    np.testing.assert_allclose(linf, np.abs(d).max(axis=(0, 1)), rtol=1e-9)
    assert np.all(l2 > 0)


def test_steady_mode_stops_on_residual_targets(tmp_path):
    from aerokinetic_amd.convergence import StopCriteria
    from aerokinetic_amd.solver import Solver
    section, flow, grid = small_case("legacy")
    s = Solver(section, flow, grid)
    crit = StopCriteria(mode="steady", mass_tolerance=1e3, momentum_x_tolerance=1e3, momentum_y_tolerance=1e3,
                        momentum_z_tolerance=1e3, minimum_t_star=0.0, consecutive_checks=3)
    summary = s.run(400 * s.plan.dt_convective, tmp_path / "r", sample_every=10, check_every=20, termination=crit)
    assert summary["termination_reason"] == "residual_targets_met" and summary["steps"] == 60
    rows = list(csv.DictReader(open(tmp_path / "r" / "residuals.csv")))
    assert [int(r["step"]) for r in rows] == [20, 40, 60] and rows[-1]["streak"] == "3"
    assert float(rows[0]["l2_jx"]) > 0


def test_direct_buffer_write_declared_equals_import():
    """A caller may write the bound population buffer directly between run calls and declare
    it (ak_populations_written): the characteristic farfield's perimeter moments, which the
    sweep keeps between calls, are then rebuilt, and the run equals one whose state came in
    through the library's import.  Left undeclared, the stale moments give another run."""
    from aerokinetic_amd.config import Flow, Grid, Plan
    from aerokinetic_amd.geometry import Section, naca0012
    from aerokinetic_amd.solver import Solver
    section = Section((naca0012(41),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=10000, alpha_deg=4.0, lattice_speed=0.05, perturbation=1e-3)
    grid = Grid(cells_per_chord=16, upstream_chords=1.0, downstream_chords=2.0, transverse_chords=1.0,
                span_chords=0.5, memory_limit_gib=2, farfield="characteristic_transport_v1")
    plan = Plan.create(section, flow, grid)
    print("direct write", plan.shape, flush=True)
    sv = [Solver(section, flow, grid, plan=plan) for _ in range(3)]
    for s in sv:
        s.advance(3, 1)
    written, stale, imported = sv
    g = imported.populations()
    g[0, :, :, :] *= 1.001  # the x- face (perimeter) and the interior change
    g[g.shape[0] // 2, :, :, :] *= 0.999
    imported.set_populations(g)
    imported.lattice.synchronize()
    for s in (written, stale):
        a, b = s.lattice, imported.lattice
        a._buf[a._offset:a._offset + a.nbytes].copy_(b._buf[b._offset:b._offset + b.nbytes])
    written.lattice.populations_written()
    rw, rs, ri = (s.advance(4, 1) for s in sv)
    keys = ("step", "cl", "cd", "cm", "local_mach_max", "density_error_max")
    rows = lambda recs: np.array([[r[k] for k in keys] for r in recs])
    np.testing.assert_array_equal(rows(rw), rows(ri))
    np.testing.assert_array_equal(written.populations(), imported.populations())
    assert not np.array_equal(stale.populations(), imported.populations())
    for s in sv:
        s.lattice.close()
