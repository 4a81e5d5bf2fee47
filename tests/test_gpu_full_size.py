"""Parity at the full sizes of the headline configurations (the sizes the benchmark runs).

* plain uniform NACA0012, 384x256x8, chord 32 at 4 deg, Re 1e4, Ma 0.1, characteristic
  farfield (aerokinetic_amd.cases.NACA0012_C32_384x256x8): per population against the FP64
  oracle (C/OpenMP) after 1 and 100 steps, loads every step;
* two-level, coarse 256x192x8 with the 160x96x16 fine patch, NACA0012 chord 64 fine cells at
  6 deg, Re 3e4 (cases.NACA0012_REFINED_256x192x8): every coarse cell and every owned fine
  fluid cell against the refined oracle for 3 coarse steps, loads every step;
* the 1536x1024x64 precision-study lattice in FP64, FP32 and FP16 storage (40 steps; guards,
  the FP64 mass check, loads of the reduced storages within 1 % of FP64).

Bars (FP64 storage) as the smaller parity cases: per population |g_gpu - g_oracle| / |f|
<= 1e-12 after one step (growing at most linearly per coarse step on the two-level case),
loads to 1e-11; after 100 uniform steps 1e-10.
"""
import numpy as np
import pytest

from aerokinetic_amd.d3q27 import W

pytestmark = pytest.mark.gpu


def _rel(a, b, mask):
    return float((np.abs(a - b) / np.abs(b + W))[mask].max())


def test_plain_384x256x8_matches_oracle():
    from aerokinetic_amd import cases
    from aerokinetic_amd.solver import Solver
    from oracle import oracle as O
    O.Oracle.set_threads(16)
    section, flow, grid, plan = cases.NACA0012_C32_384x256x8.build()
    print("case", cases.NACA0012_C32_384x256x8.name, plan.shape, flush=True)
    s = Solver(section, flow, grid, plan=plan)
    o = O.from_case(section, flow, grid, plan)
    fluid = np.broadcast_to(~s._solid_xy[:, :, None, None], (*plan.shape, 27))
    assert _rel(s.populations(), o.pdf(), fluid) <= 1e-14
    a, b = s.step(sample=True), o.step(sample=True)
    assert _rel(s.populations(), o.pdf(), fluid) <= 1e-12
    for k in ("cl", "cd", "cm"):
        assert a[k] == pytest.approx(b[k], abs=1e-11)
    recs = s.advance(99, 1)
    for r in recs:
        ref = o.step(sample=True)
        for k in ("cl", "cd", "cm"):
            assert r[k] == pytest.approx(ref[k], abs=1e-10), (r["step"], k)
    err = _rel(s.populations(), o.pdf(), fluid)
    print("100-step max relative population error", err, flush=True)
    assert err <= 1e-10


def test_two_level_256x192x8_matches_refined_oracle():
    from aerokinetic_amd import cases
    from aerokinetic_amd.refined import RefinedSolver
    from oracle.oracle import Oracle
    from oracle.refined_oracle import RefinedOracle
    Oracle.set_threads(16)
    project, settings, plan = cases.NACA0012_REFINED_256x192x8.build()
    print("case", cases.NACA0012_REFINED_256x192x8.name, plan.shape, plan.box, flush=True)
    s = RefinedSolver(project, settings, plan=plan)
    o = RefinedOracle(project, settings, plan)
    pad, nz2 = 2, 2 * plan.shape[2]
    fine_fluid = np.broadcast_to(~s._fine_solid_xy[:, :, None, None], (*s._fine_solid_xy.shape, nz2, 27))
    coarse_all = np.ones((*plan.shape, 27), bool)
    assert _rel(s.populations(0), o.coarse.pdf(), coarse_all) <= 1e-13
    for k in range(3):
        a, b = s.step(), o.step()
        for key in ("cl", "cd", "cm"):
            assert a[key] == pytest.approx(b[key], abs=1e-11), (k, key)
        assert _rel(s.populations(0), o.coarse.pdf(), coarse_all) <= 1e-12 * (k + 1)
        gf = s.populations(1)[pad:-pad, pad:-pad]
        of = o.fine.pdf()[pad:-pad, pad:-pad, pad:pad + nz2]
        assert _rel(gf, of, fine_fluid) <= 1e-12 * (k + 1)


def test_precision_study_lattice_three_storages():
    """The 1536x1024x64 lattice (100.7M cells; NACA0012 chord 128 at 8 deg, Re 1e5, Ma 0.1,
    characteristic farfield) at full size in FP64, FP32 and FP16 shifted storage for 40 steps
    with loads every step: the state stays finite and inside the low-Mach guards, the FP64
    density sum over the fluid cells (the global mass check) changes by less than 2e-6 of the
    mass in FP64 and FP32 storage (measured 4.7e-7: the characteristic farfield is open, and the
    impulsive start pushes a little mass through it) and by less than 3e-5 in FP16 storage
    (measured 1.0e-5: each cell's stored populations carry the FP16 rounding, a few units in
    the last place of the largest shifted population, ~7e-6 at Ma 0.1), and the mean Cl and Cd of steps 21-40 of
    FP32 and FP16 storage are within 1 % of FP64 storage."""
    import torch
    from aerokinetic_amd import cases
    from aerokinetic_amd.solver import Solver
    case = cases.NACA0012_C128_1536x1024x64
    section, flow, grid, plan = case.build()
    means = {}
    for storage in ("fp64", "fp32", "fp16"):
        print("case", case.name, storage, flush=True)
        s = Solver(section, flow, grid, storage=storage, plan=plan)
        nfluid = int(s._fluid.sum())
        m0 = s.diagnostics()["density_mean"] * nfluid
        recs = s.advance(40, 1)
        assert [r["step"] for r in recs] == list(range(1, 41))
        for r in recs:
            assert r["nonfinite"] == 0 and np.isfinite([r["cl"], r["cd"], r["cm"]]).all()
            assert r["local_mach_max"] < 0.3 and r["density_error_max"] < 0.1 and r["density_min"] > 0
        d = s.diagnostics()
        m1 = d["density_mean"] * nfluid
        print(storage, "mass", m0, m1, "relative change", (m1 - m0) / m0, "guards", d["local_mach_max"],
              d["density_error_max"], flush=True)
        assert abs(m1 - m0) / m0 < (3e-5 if storage == "fp16" else 2e-6)
        means[storage] = np.mean([[r["cl"], r["cd"]] for r in recs[20:]], axis=0)
        s.lattice.close()
        del s
        torch.cuda.empty_cache()
    for storage in ("fp32", "fp16"):
        rel = np.abs(means[storage] - means["fp64"]) / np.abs(means["fp64"])
        print(storage, "mean Cl, Cd relative to FP64", rel, flush=True)
        assert rel.max() < 0.01
