"""Pins of the FP64 CPU oracle against the reference's own recorded outputs.

The reference's arithmetic is generated by lbmpy/pystencils, which are absent
here; these golden vectors are its recorded 17-digit load histories:

* legacy farfield, N32, 176x101x16 (docs/verification_stage1_20261106/corrected32_t12):
  collision, Bouzidi walls, seeded init, loads, velocity bounce-back and the
  extrapolation outflow (the outflow reaches the body after ~130 steps);
* characteristic farfield, same case (docs/verification_farfield_20261107/airfoil/
  characteristic_down3): the characteristic farfield from step 64 on;
* NACA0012 with 81 points, 272 steps (profile_threads1.json last_sample):
  the NACA generator and the full legacy step.
"""
import csv
import json
from pathlib import Path

import numpy as np
import pytest

from aerokinetic_amd.config import Flow, Grid, Plan
from aerokinetic_amd.geometry import Section, naca0012
from aerokinetic_amd.project import Project
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def _history(name):
    with open(GOLDEN / name / "history_head.csv", newline="") as fh:
        return {int(r["step"]): {k: float(r[k]) for k in ("cl", "cd", "cm", "local_mach_max", "density_error_max")}
                for r in csv.DictReader(fh)}


def _run_against(name, last_step):
    project = Project.load(GOLDEN / name / "project.json")
    plan = project.plan()
    assert plan.shape == (176, 101, 16)
    ref = _history(name)
    o = O.from_case(project.section, project.flow, project.grid, plan)
    worst = 0.0
    for step in range(1, last_step + 1):
        rec = o.step(sample=step in ref)
        if step in ref:
            d = o.diagnostics()
            for k in ("cl", "cd", "cm"):
                worst = max(worst, abs(rec[k] - ref[step][k]))
            assert d["local_mach_max"] == pytest.approx(ref[step]["local_mach_max"], rel=1e-12)
            assert d["density_error_max"] == pytest.approx(ref[step]["density_error_max"], rel=1e-11)
    return worst


def test_legacy_history_pins_collision_walls_outflow():
    worst = _run_against("n32_legacy_down3", 240)
    assert worst < 5e-12, worst


def test_characteristic_history_pins_farfield():
    worst = _run_against("n32_characteristic_down3", 400)
    assert worst < 5e-12, worst


def test_profile_case_last_sample():
    spec = json.loads((GOLDEN / "profile_n32_naca0012_81.json").read_text())
    c = spec["case"]
    section = Section((naca0012(c["naca0012_n_side"]),), c["chord"], tuple(c["moment_reference"]))
    flow, grid = Flow(**c["flow"]), Grid(**c["grid"])
    plan = Plan.create(section, flow, grid)
    assert list(plan.shape) == spec["shape"]
    o = O.from_case(section, flow, grid, plan)
    for i in range(c["steps"]):
        rec = o.step(sample=(i + 1) % c["sample_every"] == 0)
    for k, v in spec["last_sample"].items():
        assert rec[k] == pytest.approx(v, abs=1e-12)


def test_fingerprint_fixture():
    p = Project.load(GOLDEN / "legacy051_project.json")
    assert p.case_fingerprint == "9ae371cecf492f67494d9f6ac6eb70f2a33cc4e6ca680a7088b6ce146206f48f"
    assert "farfield" not in p.as_dict()["grid"]


def _refined_history(name):
    with open(GOLDEN / name / "history_head.csv", newline="") as fh:
        return {int(r["step"]): {k: float(r[k]) for k in ("cl", "cd", "cm")} for r in csv.DictReader(fh)}


@pytest.mark.parametrize("name,steps", [("refined_startup_n32", 12), ("refined_fowler_n32", 6)])
def test_refined_oracle_pins_transfer_interface_and_fine_walls(name, steps):
    """Two-level 2:1 step: prolongation in time, reflux, restriction, fine Bouzidi walls, loads."""
    from aerokinetic_amd.refined_plan import RefinedPlan, load_refined_project
    from oracle.refined_oracle import RefinedOracle
    project, settings = load_refined_project(GOLDEN / name / "project.json")
    ref = _refined_history(name)
    ro = RefinedOracle(project, settings, RefinedPlan.create(project, settings))
    for _ in range(steps):
        rec = ro.step()
        if rec["step"] in ref:
            for k in ("cl", "cd", "cm"):
                assert rec[k] == pytest.approx(ref[rec["step"]][k], abs=1e-12)
