"""GPU parity of the HIP path against the FP64 CPU oracle and the reference's recorded histories.

Bars (FP64 storage): per population |g_gpu - g_oracle| / |f| <= 1e-12 after one
step and <= 1e-9 after 1000 steps on laminar cases; loads to 1e-11 absolute
in the coefficients.  f = g + w is the absolute population.
"""
import csv
from pathlib import Path

import numpy as np
import pytest

from aerokinetic_amd.config import Flow, Grid, Plan
from aerokinetic_amd.d3q27 import W
from aerokinetic_amd.geometry import Section, naca0012
from aerokinetic_amd.project import Project
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def small_case(farfield="characteristic_transport_v1", alpha=4.0, perturbation=1e-3, reynolds=10000):
    section = Section((naca0012(41),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=reynolds, alpha_deg=alpha, lattice_speed=0.05, perturbation=perturbation)
    grid = Grid(cells_per_chord=16, upstream_chords=1.0, downstream_chords=2.0, transverse_chords=1.0,
                span_chords=0.5, memory_limit_gib=2, farfield=farfield)
    return section, flow, grid, Plan.create(section, flow, grid)


def rel_err(g_gpu, g_ref, fluid):
    f = np.abs(g_ref + W)
    return float((np.abs(g_gpu - g_ref) / f)[fluid].max())


def gpu_solver(section, flow, grid, plan, storage="fp64"):
    from aerokinetic_amd.solver import Solver
    return Solver(section, flow, grid, storage=storage, plan=plan)


@pytest.mark.parametrize("farfield", ["characteristic_transport_v1", "legacy"])
def test_initial_state_and_one_step(farfield):
    section, flow, grid, plan = small_case(farfield)
    print("case", farfield, plan.shape, flush=True)
    s = gpu_solver(section, flow, grid, plan)
    o = O.from_case(section, flow, grid, plan)
    fluid = np.broadcast_to(~s._solid_xy[:, :, None, None], (*plan.shape, 27))
    assert rel_err(s.populations(), o.pdf(), fluid) <= 1e-14
    rec = s.step(sample=True)
    ref = o.step(sample=True)
    assert rel_err(s.populations(), o.pdf(), fluid) <= 1e-12
    for k in ("cl", "cd", "cm"):
        assert rec[k] == pytest.approx(ref[k], abs=1e-11)
    np.testing.assert_allclose(rec["element_lattice_loads"], ref["element_lattice_loads"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("farfield", ["characteristic_transport_v1", "legacy"])
def test_thousand_steps_laminar(farfield):
    # Re 1e4 at 16 cells/chord over 1000 steps (t* = 3.1) is a laminar start-up
    section, flow, grid, plan = small_case(farfield, alpha=2.0, perturbation=0.0)
    print("case", farfield, plan.shape, flush=True)
    s = gpu_solver(section, flow, grid, plan)
    o = O.from_case(section, flow, grid, plan)
    recs = s.advance(1000, 100)
    refs = [o.step(sample=(k + 1) % 100 == 0) for k in range(1000)]
    refs = [r for r in refs if r is not None]
    fluid = np.broadcast_to(~s._solid_xy[:, :, None, None], (*plan.shape, 27))
    err = rel_err(s.populations(), o.pdf(), fluid)
    print("max relative population error after 1000 steps", err)
    assert err <= 1e-9
    for a, b in zip(recs, refs):
        for k in ("cl", "cd", "cm"):
            assert a[k] == pytest.approx(b[k], abs=1e-9)


@pytest.mark.parametrize("farfield", ["characteristic_transport_v1", "legacy"])
def test_body_near_the_edge_takes_the_separate_boundary_launches(farfield):
    # A wall link within 3 cells of the x = 0 edge: the walls no longer share a launch with
    # the edge phase (x-face copy or legacy links); the separate launches keep oracle parity.
    section = Section((naca0012(41),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=10000, alpha_deg=4.0, lattice_speed=0.05, perturbation=1e-3)
    grid = Grid(cells_per_chord=16, upstream_chords=0.125, downstream_chords=2.0, transverse_chords=1.0,
                span_chords=0.5, memory_limit_gib=2, farfield=farfield)
    plan = Plan.create(section, flow, grid)
    print("case near edge", farfield, plan.shape, flush=True)
    s = gpu_solver(section, flow, grid, plan)
    assert int(s.links["x"].min()) < 3
    o = O.from_case(section, flow, grid, plan)
    fluid = np.broadcast_to(~s._solid_xy[:, :, None, None], (*plan.shape, 27))
    recs = s.advance(20, 10)
    refs = [o.step(sample=(k + 1) % 10 == 0) for k in range(20)]
    refs = [r for r in refs if r is not None]
    assert rel_err(s.populations(), o.pdf(), fluid) <= 1e-12
    for a, b in zip(recs, refs):
        for k in ("cl", "cd", "cm"):
            assert a[k] == pytest.approx(b[k], abs=1e-11)


def _golden(name, last):
    project = Project.load(GOLDEN / name / "project.json")
    plan = project.plan()
    with open(GOLDEN / name / "history_head.csv", newline="") as fh:
        ref = {int(r["step"]): r for r in csv.DictReader(fh)}
    s = gpu_solver(project.section, project.flow, project.grid, plan)
    recs = s.advance(last, 8)
    worst = 0.0
    for rec in recs:
        r = ref[rec["step"]]
        for k in ("cl", "cd", "cm"):
            worst = max(worst, abs(rec[k] - float(r[k])))
        assert rec["local_mach_max"] == pytest.approx(float(r["local_mach_max"]), rel=1e-11)
        assert rec["density_error_max"] == pytest.approx(float(r["density_error_max"]), rel=1e-10)
    return worst


def test_gpu_matches_recorded_legacy_history():
    worst = _golden("n32_legacy_down3", 240)
    print("worst |dC| legacy", worst)
    assert worst < 5e-12


def test_gpu_matches_recorded_characteristic_history():
    worst = _golden("n32_characteristic_down3", 400)
    print("worst |dC| characteristic", worst)
    assert worst < 5e-12


def test_periodic_shear_wave_decay_and_parity():
    """Shear wave in a fully periodic box (tests/test_numerics.py:8-29 of the reference)."""
    from aerokinetic_amd.lattice import Lattice
    nu, amp, steps, n = 0.03, 1e-4, 120, 32
    omega = 1 / (0.5 + 3 * nu)
    lat = Lattice(n, 4, 4, storage="fp64", omega_shear=omega)
    lat.set_solid(np.zeros((n, 4), dtype=np.uint8))
    x = np.arange(n) + 0.5
    uy = amp * np.sin(2 * np.pi * x / n)
    g = np.zeros((n, 4, 4, 27))
    from aerokinetic_amd.d3q27 import CX, CY, CZ
    def phi(c, u):
        return np.where(c == 0, 2 / 3 - u * u, (1 / 3 + u * u + c * u) / 2)
    for i in range(27):
        g[..., i] = (phi(CX[i], 0.0) * phi(CY[i], uy) * phi(CZ[i], 0.0))[:, None, None] - W[i]
    lat.set_populations(g)
    lat.set_farfield("periodic", 0.0)
    o = O.Oracle(n, 4, 4, omega, 1.0)
    o.set_populations = o.set_pdf
    o.set_pdf(g)
    o.set_farfield("periodic")
    lat.run(steps, 0, 0)
    for _ in range(steps):
        o.step()
    gg, go = lat.populations(), o.pdf()
    assert np.max(np.abs(gg - go) / np.abs(go + W)) <= 1e-12
    rho = 1 + gg.sum(-1)
    vy = (gg * CY).sum(-1) / rho
    harmonic = abs(np.fft.fft(vy.mean(axis=(1, 2)))[1]) * 2 / n
    assert harmonic / amp == pytest.approx(np.exp(-nu * (2 * np.pi / n) ** 2 * steps), rel=0.012)
    np.testing.assert_allclose(rho, 1.0, atol=2e-12)


def test_macroscopic_export_matches_population_moments():
    from aerokinetic_amd.d3q27 import CX, CY, CZ
    section, flow, grid, plan = small_case("characteristic_transport_v1")
    s = gpu_solver(section, flow, grid, plan)
    for n in (3, 4):  # both AA parities
        s.advance(n, 0)
        g = s.populations()
        rho, u = s.lattice.macroscopic()
        r_ref = 1.0 + g.sum(axis=-1)
        j = np.stack([g @ CX, g @ CY, g @ CZ], axis=-1)
        np.testing.assert_allclose(rho, r_ref, rtol=0, atol=1e-14)
        np.testing.assert_allclose(u, j / r_ref[..., None], rtol=0, atol=1e-14)
