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


@pytest.mark.parametrize("name,steps", [("refined_startup_n32", 12), ("refined_fowler_n32", 32)])
def test_refined_oracle_pins_transfer_interface_and_fine_walls(name, steps):
    """Two-level 2:1 step: prolongation in time, reflux, restriction, fine Bouzidi walls, loads;
    for the two-element Fowler case also the conservation budget and the leaf guards of the
    recorded summary after 32 coarse steps (refined_solver.py:291-361)."""
    from aerokinetic_amd.refined_plan import RefinedPlan, load_refined_project
    from oracle.refined_oracle import RefinedOracle
    O.Oracle.set_threads(8)
    project, settings = load_refined_project(GOLDEN / name / "project.json")
    ref = _refined_history(name)
    ro = RefinedOracle(project, settings, RefinedPlan.create(project, settings))
    for _ in range(steps):
        rec = ro.step()
        if rec["step"] in ref:
            for k in ("cl", "cd", "cm"):
                assert rec[k] == pytest.approx(ref[rec["step"]][k], abs=1e-12)
    budget_file = GOLDEN / name / "summary_diagnostics.json"
    if not budget_file.exists():
        return
    want = json.loads(budget_file.read_text())
    assert want["steps"] == steps
    check_budget(ro.diagnostics(), want["diagnostics"], rel=1e-12)


def open_boundary_case():
    """A small two-level case whose start disturbance reaches the outer faces within 16
    coarse steps: NACA0012 at 6 deg, 16 coarse cells per chord, 56x27x8 coarse cells."""
    from aerokinetic_amd.project import RunSettings
    from aerokinetic_amd.refined_plan import RefinedPlan, RefinementSettings
    section = Section((naca0012(81),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=10000, alpha_deg=6.0, lattice_speed=0.05, perturbation=0.01)
    grid = Grid(cells_per_chord=16, upstream_chords=1.0, downstream_chords=1.5, transverse_chords=0.75,
                span_chords=0.5, memory_limit_gib=2, farfield="legacy")
    project = Project(section, flow, grid, RunSettings())
    settings = RefinementSettings()
    return project, settings, RefinedPlan.create(project, settings)


def test_refined_oracle_budget_closes_through_open_faces():
    """Conservation pins the outer face flux of the oracle: the change of the owned cells'
    mass and momentum over a coarse step equals what entered through the outer faces plus
    the wall exchange, to roundoff, also once the disturbance crosses the faces (the flux of
    the ghost layer's never-written populations, the free stream of the initial fill,
    included)."""
    from oracle.refined_oracle import RefinedOracle
    O.Oracle.set_threads(8)
    project, settings, plan = open_boundary_case()
    ro = RefinedOracle(project, settings, plan)
    for _ in range(30):
        ro.step()
        b = ro.last_budget
        # roundoff: the integrals are ~1 per unit volume, summed over ~1e5 cells per level
        assert abs(b["mass_residual_per_volume"]) < 1e-15
        assert max(map(abs, b["momentum_residual_per_volume"])) < 1e-12
    assert abs(b["external_mass_in"]) > 1e-6 and abs(b["external_impulse_in"][1]) > 1e-3
    # a planted mistake in the restatement: the face flux taken before the farfield writes
    # this step's incoming values breaks the balance once the faces carry a disturbance
    bad = RefinedOracle(project, settings, plan)
    bad.outer_flux_before_boundaries = True
    worst = 0.0
    for _ in range(24):
        bad.step()
        worst = max(worst, abs(bad.last_budget["mass_residual_per_volume"]))
    assert worst > 1e-10


def check_budget(got, want, rel):
    """Guards and budget of a two-level run against the reference's recorded diagnostics.
    The residuals per volume are sums of ~1e6 cells that cancel to a few hundred ulps of the
    integrals: only their size is pinned (the recorded x-momentum defect, 2.9e-14 for the
    Fowler case, is a roundoff pattern of the summation order)."""
    for k in ("local_mach_max", "density_error_max", "cumulative_wall_mass_exchange",
              "cumulative_wall_mass_exchange_over_initial_mass"):
        assert got[k] == pytest.approx(want[k], rel=rel), k
    gb, wb = got["last_budget"], want["last_budget"]
    assert gb["wall_mass_exchange"] == pytest.approx(wb["wall_mass_exchange"], rel=rel)
    assert gb["wall_impulse"][:2] == pytest.approx(wb["wall_impulse"][:2], rel=rel)
    assert abs(gb["wall_impulse"][2]) < 1e-18
    # outer in-flux: the free stream enters and leaves; it cancels to roundoff while the start
    # disturbance is inside the domain, and is a small difference of large face fluxes later
    scale = max(abs(wb["external_mass_in"]), 1e-9)
    assert gb["external_mass_in"] == pytest.approx(wb["external_mass_in"], abs=2e-5 * scale + 1e-15)
    for g, w in zip(gb["external_impulse_in"][:2], wb["external_impulse_in"][:2]):
        assert g == pytest.approx(w, abs=2e-5 * max(abs(w), 1e-9) + 1e-15)
    assert abs(gb["mass_residual_per_volume"]) < 1e-14
    assert max(map(abs, gb["momentum_residual_per_volume"])) < 1e-12


# ---------------------------------------------------------------- exterior farfield
EXT_COLUMNS = ("wall_mass_exchange_lattice", "wall_mass_exchange_accumulated",
               "exterior_source_mass_lattice", "exterior_source_mass_accumulated")


def _exterior_history(name):
    with open(GOLDEN / name / "history_head.csv", newline="") as fh:
        return {int(r["step"]): {k: float(v) for k, v in r.items() if k != "time_s" and v != ""}
                for r in csv.DictReader(fh)}


@pytest.mark.parametrize("name, shape, last_step", [
    ("exterior_n16_smoke", (136, 51, 8), 48),   # every step from step 1
    ("exterior_n16_short", (88, 51, 8), 32),
    ("exterior_n32_down3", (176, 101, 16), 160),  # departs from the legacy history from step 64
])
def test_exterior_oracle_pins_ring_predictor_reconstruction_and_sources(name, shape, last_step):
    """The exterior restatement (oracle/exterior.py) against the reference's recorded
    exterior runs: loads, guards, wall mass exchange and the exterior source of mass."""
    from oracle.exterior import ExteriorOracle
    project = Project.load(GOLDEN / name / "project.json")
    plan = project.plan()
    assert plan.shape == shape and project.grid.farfield == "exterior_regularized_fd_v1"
    ref = _exterior_history(name)
    o = ExteriorOracle(project.section, project.flow, project.grid, plan)
    assert o.storage_shape == (shape[0] + 2, shape[1] + 2, shape[2])
    for step in range(1, last_step + 1):
        rec = o.step(sample=step in ref)
        if step not in ref:
            continue
        r = ref[step]
        for k in ("cl", "cd", "cm"):
            assert abs(rec[k] - r[k]) < 5e-12, (step, k, rec[k], r[k])
        d = o.diagnostics()
        assert d["local_mach_max"] == pytest.approx(r["local_mach_max"], rel=1e-12)
        assert d["density_error_max"] == pytest.approx(r["density_error_max"], rel=1e-11)
        for k in EXT_COLUMNS:
            # a step's source is a sum of ~1e5 ring populations' changes (1e-13 absolute
            # rounding floor per step, which the accumulated column adds up)
            floor = 1e-12 + (2e-13 * step if k.endswith("accumulated") else 0.0)
            assert abs(rec[k] - r[k]) <= floor + 1e-9 * abs(r[k]), (step, k, rec[k], r[k])


def _band_oracle():
    from oracle.exterior import ExteriorOracle
    project = Project.load(GOLDEN / "exterior_n16_short" / "project.json")
    return ExteriorOracle(project.section, project.flow, project.grid, project.plan())


def test_exterior_reconstruction_moments_closed_form():
    """For a linear velocity field the three-point differences are exact, so the
    reconstructed ring populations must carry rho, rho u and the second moment
    rho (u u + I/3) - rho/3 [(A + A^T)/omega + 2/3 tr(A) (1 - 1/omega) I] exactly."""
    from oracle.exterior import _C
    o = _band_oracle()
    X, Y, Z = o.storage_shape
    A = np.array([[1e-3, -2e-3, 0.0], [5e-4, -7e-4, 0.0], [0.0, 0.0, 0.0]])  # A[a][b] = d_a u_b
    xyz = o.band_xyz.reshape(len(o.band), Z, 3).astype(float)
    u = np.array([0.04, 0.01, -0.002]) + xyz[..., 0, None] * A[0] + xyz[..., 1, None] * A[1]
    lnrho = np.full(u.shape[:2], np.log(1.003))
    V = np.concatenate([lnrho[..., None], u], axis=-1)
    pdf, err = o.regularize(V)
    assert err < 1e-15
    rho = 1.003
    w = o.omega
    S = -rho / 3 * ((A + A.T) / w + np.eye(3) * (2 / 3 * np.trace(A) * (1 - 1 / w)))
    f = pdf + np.array([{0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}[int(np.abs(c).sum())] for c in _C])
    second = np.einsum("rzi,ia,ib->rzab", f, _C, _C)
    ur = u[: o.R]
    expect = rho * (ur[..., :, None] * ur[..., None, :] + np.eye(3) / 3) + S
    assert np.max(np.abs(second - expect)) < 1e-15


def test_exterior_predictor_keeps_free_stream():
    """The uniform free stream is a fixed point of the ring predictor (rates vanish, the
    imposed inflow and glancing invariants are the free stream's)."""
    o = _band_oracle()
    X, Y, Z = o.storage_shape
    V = np.zeros((len(o.band), Z, 4))
    V[..., 1] = o.U
    o.rows = np.stack([-V[o.g_row_index[0], :, 2] - V[o.g_row_index[0], :, 0] * np.sqrt(1 / 3),
                       V[o.g_row_index[1], :, 2] - V[o.g_row_index[1], :, 0] * np.sqrt(1 / 3)])
    y = o.predict(V, 7)
    assert np.max(np.abs(y - V[: o.R])) < 1e-17


def test_exterior_predictor_signs_fail_on_planted_mistake(monkeypatch):
    """A flipped sign of the density rate of the predictor breaks the pin above within
    48 steps (through the exterior source column; the loads feel it later)."""
    import oracle.exterior as E
    project = Project.load(GOLDEN / "exterior_n16_smoke" / "project.json")
    ref = _exterior_history("exterior_n16_smoke")
    good = E.ExteriorOracle(project.section, project.flow, project.grid, project.plan())
    bad = E.ExteriorOracle(project.section, project.flow, project.grid, project.plan())
    rhs = bad._rhs

    def flipped(y, cur):
        out = rhs(y, cur)
        out[..., 0] = -out[..., 0]
        return out
    bad._rhs = flipped

    def violation(rec, r):
        k = "exterior_source_mass_lattice"
        return abs(rec[k] - r[k]) - (1e-12 + 1e-9 * abs(r[k]))
    worst_good = worst_bad = -1.0
    for step in range(1, 49):
        a, b = good.step(sample=True), bad.step(sample=True)
        worst_good = max(worst_good, violation(a, ref[step]))
        worst_bad = max(worst_bad, violation(b, ref[step]))
    assert worst_good <= 0 < worst_bad
