"""Two-level conservation budget, leaf guards and residuals on the GPU (FP64 storage).

* The reference's recorded multi-element runs (two- and three-element sections, 32 and 800
  coarse steps; docs/verification_multielement_050): `RefinedSolver.run` with each project's
  own run settings writes history.csv, residuals.csv and summary.json; the loads, the residual
  norms, the guards and the budget (cumulative wall mass exchange, last step's wall mass,
  wall impulse and outer in-flux, balance defects) are compared with the recorded files.
* Against the refined oracle on a small open-boundary case whose start disturbance reaches
  the outer faces within 16 coarse steps (so the outer in-flux is not a free-stream
  cancellation): every budget term after each of 30 coarse steps.
"""
import csv
import json
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

from test_oracle_golden import check_budget

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _csv(path):
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


@pytest.mark.parametrize("name", ["refined_fowler_n32", "refined_double_n32", "refined_slotted_n32"])
def test_recorded_multielement_runs(tmp_path, name):
    from aerokinetic_amd.refined import RefinedSolver
    from aerokinetic_amd.refined_plan import load_refined_project
    project, settings = load_refined_project(GOLDEN / name / "project.json")
    want = json.loads((GOLDEN / name / "summary_diagnostics.json").read_text())
    print("case", name, "steps", want["steps"], flush=True)
    s = RefinedSolver(project, settings)
    summary = s.run(tmp_path / "run")
    assert summary["status"] == "completed" and summary["steps"] == want["steps"]
    # loads of every coarse step, same columns as the reference's refined history
    got, ref = _csv(tmp_path / "run" / "history.csv"), _csv(GOLDEN / name / "history_head.csv")
    assert list(got[0]) == list(ref[0])
    ref = {int(r["step"]): r for r in ref}
    worst = 0.0
    for r in got:
        if int(r["step"]) in ref:
            for k in ("cl", "cd", "cm"):
                worst = max(worst, abs(float(r[k]) - float(ref[int(r["step"])][k])))
    print("largest load difference", worst, flush=True)
    assert worst < (5e-12 if want["steps"] <= 32 else 1e-9)
    # guards and budget after the last step
    check_budget(summary["diagnostics"], want["diagnostics"], rel=1e-9 if want["steps"] <= 32 else 1e-6)
    assert summary["diagnostics"]["coarse_boundary_calls"] == want["diagnostics"]["coarse_boundary_calls"]
    assert summary["diagnostics"]["fine_boundary_calls"] == want["diagnostics"]["fine_boundary_calls"]
    # residuals: the stationarity norms to 1e-8, the balance defects (roundoff of the integral
    # sums) only in size
    got, ref = _csv(tmp_path / "run" / "residuals.csv"), _csv(GOLDEN / name / "residuals.csv")
    assert [r["step"] for r in got] == [r["step"] for r in ref]
    for g, r in zip(got, ref):
        for k in r:
            if k.startswith(("l2_", "linf_")):
                assert float(g[k]) == pytest.approx(float(r[k]), rel=1e-8, abs=1e-12), (g["step"], k)
            elif k.startswith("balance_"):
                assert g[k] != "" and float(g[k]) < 1e-9, (g["step"], k)


def test_budget_matches_refined_oracle():
    from aerokinetic_amd.refined import RefinedSolver
    from oracle.oracle import Oracle
    from oracle.refined_oracle import RefinedOracle
    from test_oracle_golden import open_boundary_case
    Oracle.set_threads(16)
    project, settings, plan = open_boundary_case()
    s = RefinedSolver(project, settings, plan=plan)
    o = RefinedOracle(project, settings, plan)
    assert s.budget()[17] == pytest.approx(o.initial_integrals["mass"], rel=1e-14)
    for step in range(1, 31):
        a, b = s.step(), o.step()
        for k in ("cl", "cd", "cm"):
            assert a[k] == pytest.approx(b[k], abs=1e-11)
        ga, gb = s.last_budget, o.last_budget
        if step % 5 == 0:
            print("step", step, "outer mass in", ga["external_mass_in"], gb["external_mass_in"], flush=True)
        # the outer faces carry a net flux once the start disturbance reaches them (step ~16)
        tol = 1e-9 * max(abs(gb["external_mass_in"]), 1e-6)
        assert ga["external_mass_in"] == pytest.approx(gb["external_mass_in"], abs=tol)
        assert ga["external_impulse_in"] == pytest.approx(gb["external_impulse_in"], rel=1e-9, abs=1e-14)
        assert ga["wall_mass_exchange"] == pytest.approx(gb["wall_mass_exchange"], rel=1e-10)
        assert ga["wall_impulse"] == pytest.approx(gb["wall_impulse"], rel=1e-10, abs=1e-18)
        assert abs(ga["mass_residual_per_volume"]) < 1e-14
        assert max(map(abs, ga["momentum_residual_per_volume"])) < 1e-13
    assert abs(o.last_budget["external_mass_in"]) > 1e-6
    d, od = s.diagnostics(), o.diagnostics()
    for k in ("local_mach_max", "density_error_max", "cumulative_wall_mass_exchange",
              "cumulative_wall_mass_exchange_over_initial_mass"):
        assert d[k] == pytest.approx(od[k], rel=1e-10), k
    b = s.budget()
    assert b[13:17] == pytest.approx(o.cumulative_external_in, rel=1e-9, abs=1e-14)
    assert int(b[29]) == 30 and b[31] == 1.0


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_reduced_storage_budget_closes(storage):
    """FP64 mass check of reduced-precision storage: the budget of a step closes to the
    storage's own rounding (per owned volume)."""
    from aerokinetic_amd.refined import RefinedSolver
    from aerokinetic_amd.refined_plan import load_refined_project
    project, settings = load_refined_project(GOLDEN / "refined_fowler_n32" / "project.json")
    s = RefinedSolver(project, settings, storage=storage)
    s.advance(20, 1)
    lb = s.last_budget
    bar = {"fp32": 1e-7, "fp16": 1e-4}[storage]
    print(storage, lb, flush=True)
    assert abs(lb["mass_residual_per_volume"]) < bar
    assert np.isfinite(lb["momentum_residual_per_volume"]).all()
    assert s.diagnostics()["coarse_boundary_calls"] == 20
