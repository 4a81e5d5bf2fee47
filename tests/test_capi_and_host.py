"""CPU-side checks: the C ABI library loads and exports every declared symbol; host
logic (case model, wall links) against the reference's rules and the oracle."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from aerokinetic_amd import _lib
from aerokinetic_amd.config import Flow, Grid, Plan
from aerokinetic_amd.geometry import Contour, Section, naca0012
from aerokinetic_amd.project import Project
from aerokinetic_amd.timing import duration_steps
from aerokinetic_amd.walls import extrude, labels, wall_links
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "aerokinetic.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*|void \*)\s*(ak_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.exported_symbols())
    dyn = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in dyn.stdout.splitlines() if " T ak_" in ln}
    assert exported == set(names)
    assert b"gfx950" in lib.ak_version()


def test_product_links_equal_independent_oracle_links():
    project = Project.load(ROOT / "tests" / "golden" / "n32_legacy_down3" / "project.json")
    plan = project.plan()
    contours = project.section.normalized(project.flow.alpha_deg)
    links, fallback = wall_links(contours, plan.shape[:2], plan.origin, plan.dx)
    o = O.Oracle(*plan.shape, plan.relaxation_rate)
    o.set_geometry(contours, plan.origin, plan.dx)
    ol = o.links()
    ext = extrude(links, plan.shape[2])
    a = np.lexsort((ext["z"], ext["d"], ext["y"], ext["x"], ext["element"]))
    b = np.lexsort((ol["z"], ol["d"], ol["y"], ol["x"], ol["element"]))
    for k in ("x", "y", "z", "d", "element"):
        np.testing.assert_array_equal(ext[k][a], ol[k][b])
    np.testing.assert_array_equal(ext["q"][a], ol["q"][b])
    assert fallback * plan.shape[2] == o.fallback_links()
    np.testing.assert_array_equal(labels(contours, plan.shape[:2], plan.origin, plan.dx)[1:-1, 1:-1] > 0, o.solid())


def test_link_fraction_against_exact_plane():
    # tests/test_geometry_and_units.py:6-10 of the reference
    box = Contour([[0.25, -1], [2, -1], [2, 1], [0.25, 1]])
    p = np.array([[0, 0], [0, 0.5], [0, -0.25]])
    offsets = np.array([[1, 0], [1, 1], [1, -1]])
    np.testing.assert_allclose(box.link_fraction(p, offsets), 0.25)


def test_links_land_on_the_polygon():
    # tests/test_numerics.py:40-56 of the reference: q puts the wall point on the contour
    section = Section((naca0012(31),))
    flow, grid = Flow(perturbation=0), Grid(cells_per_chord=16, upstream_chords=1, downstream_chords=2,
                                            transverse_chords=1, memory_limit_gib=1)
    plan = Plan.create(section, flow, grid)
    contours = section.normalized(flow.alpha_deg)
    links, _ = wall_links(contours, plan.shape[:2], plan.origin, plan.dx)
    from aerokinetic_amd.d3q27 import CX, CY
    d = links["d"]
    pts = np.column_stack(((links["x"] + 0.5) * plan.dx + plan.origin[0], (links["y"] + 0.5) * plan.dx + plan.origin[1]))
    q = links["q"]
    assert (q > 0).all()
    dirs = np.column_stack((CX[d], CY[d])) * plan.dx
    wall = pts + q[:, None] * dirs
    eps = 1e-6 * dirs
    assert not contours[0].contains(*(wall - eps).T).any()
    assert contours[0].contains(*(wall + eps).T).all()


def test_units_and_duration_rounding():
    a = Contour([[0, 0], [1, 0], [1, 0.1], [0, 0.1]])
    b = Contour([[1.2, 0], [2, -0.1], [2, -0.2], [1.2, -0.1]])
    section = Section((a, b))
    assert section.chord == pytest.approx(np.hypot(2, 0.15))
    flow = Flow(reynolds=30000, physical_chord_m=0.15)
    grid = Grid(cells_per_chord=64)
    plan = Plan.create(section, flow, grid)
    assert flow.lattice_speed * grid.cells_per_chord / plan.viscosity == pytest.approx(30000)
    assert plan.dt_seconds * flow.physical_speed / flow.physical_chord_m == pytest.approx(plan.dt_convective)
    assert duration_steps(0.0175, 0.04 / 16) == 7
    assert duration_steps(1.0, 0.1) == 10
    with pytest.raises(ValueError):
        Flow(reynolds=5000)
    with pytest.raises(ValueError, match="farfield"):
        Grid(farfield="misspelled")


def test_explicit_plan_puts_leading_edge_on_requested_cell():
    from aerokinetic_amd import cases
    section, flow, grid, plan = cases.NACA0012_C32_384x256x8.build()
    assert plan.shape == (384, 256, 8)
    le = section.leading_edge_normalized(flow.alpha_deg)
    np.testing.assert_allclose((le[0] - plan.origin[0]) / plan.dx, 128)
    np.testing.assert_allclose((le[1] - plan.origin[1]) / plan.dx, 128)
    assert flow.lattice_speed == pytest.approx(0.1 / np.sqrt(3))


def test_exterior_plan_keeps_the_physical_domain_and_rejects_unsupported_models():
    """tests/test_exterior.py:34-48 of the reference: the exterior farfield pads the storage,
    not the physical plan; GPU target, SGS and refined plans are rejected."""
    from dataclasses import replace
    from aerokinetic_amd.config import Flow, Grid, Plan
    from aerokinetic_amd.geometry import Section, naca4
    from aerokinetic_amd.project import Project
    from aerokinetic_amd.refined_plan import RefinedPlan, RefinementSettings
    mode = "exterior_regularized_fd_v1"
    p = Project(Section((naca4("0012", 31),)), Flow(reynolds=30000),
                Grid(cells_per_chord=16, upstream_chords=1, downstream_chords=2, transverse_chords=1))
    legacy = p.plan()
    explicit = replace(p, grid=replace(p.grid, farfield=mode))
    new = explicit.plan()
    for name in ("shape", "origin", "dx", "span", "viscosity", "dt_convective", "dt_seconds"):
        assert getattr(new, name) == getattr(legacy, name)
    assert new.estimated_memory_gib > legacy.estimated_memory_gib
    assert any("storage shape=(" in n for n in new.notes)
    with pytest.raises(ValueError, match="CPU"):
        replace(p.grid, farfield=mode, target="gpu")
    with pytest.raises(ValueError, match="sgs=none"):
        Plan.create(p.section, replace(p.flow, sgs="smagorinsky"), explicit.grid)
    with pytest.raises(ValueError, match="uniform"):
        RefinedPlan.create(explicit, RefinementSettings())


def test_profile_summary_takes_the_last_dispatches_in_start_order(tmp_path, capsys):
    """The kernel-trace summary used for the benchmark profiles: per kernel name, the mean of all
    dispatches and of the last N, in start order whatever the file order."""
    from aerokinetic_amd import profile_summary as P
    trace = tmp_path / "t_kernel_trace.csv"
    trace.write_text(
        "Kind,Agent_Id,Kernel_Name,Start_Timestamp,End_Timestamp\n"
        'K,1,"void ak::k_sweep<0, true, false, true>(double*, int)",9000,11000\n'
        'K,1,"void ak::k_sweep<0, true, false, true>(double*, int)",100,8100\n'
        'K,1,"void ak::k_sweep<0, true, false, true>(double*, int)",20000,23000\n'
        'K,1,"ak::k_fill<unsigned char>(unsigned char*, long long)",1,2001\n', encoding="utf-8")
    d = P.durations(str(trace))
    assert d['void ak::k_sweep<0, true, false, true>(double*, int)'] == [8.0, 2.0, 3.0]
    assert P.short('void ak::k_sweep<0, true, false, true>(double*, int)') == "ak::k_sweep<0, true, false, true>"
    P.main([str(trace), "--last", "2", "--match", "k_sweep"])
    out = capsys.readouterr().out
    assert "| `ak::k_sweep<0, true, false, true>` | 3 | 4.3 | 2.5 | 2.0–3.0 |" in out
    assert "k_fill" not in out


def test_profile_summary_counter_mode_per_wave_and_issue_shares(tmp_path, capsys):
    """The --pmc summary: counters per wave, the clock implied by SQ_BUSY_CYCLES over the shader
    engines, and the VALU issue share = VALU instructions per CU / cycles of the dispatch."""
    from aerokinetic_amd import profile_summary as P
    f = tmp_path / "c_counter_collection.csv"
    head = "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp\n"
    k = '"void ak::k_sweep<2, false, false, true>(ak::St)"'
    # 2 CUs, 1 engine, 1000 ns at 2 GHz = 2000 cycles; 4 waves x 500 VALU = 2000 -> 1000 per CU = 50 %
    rows = [f"7,{k},{c},{v},1000,2000" for c, v in
            (("SQ_WAVES", 4), ("SQ_INSTS_VALU", 2000), ("SQ_INSTS_SALU", 400), ("SQ_BUSY_CYCLES", 2000))]
    f.write_text(head + "\n".join(rows) + "\n", encoding="utf-8")
    r, = P.issue_rows(str(f), "k_sweep", cus=2, engines=1)
    assert r["dispatches"] == 1 and r["us"] == 1.0
    assert r["SQ_INSTS_VALU"] == 500.0 and r["SQ_INSTS_SALU"] == 100.0
    assert abs(r["ghz"] - 2.0) < 1e-12
    assert abs(r["valu_issue"] - 0.5) < 1e-12 and abs(r["salu_issue"] - 0.1) < 1e-12
    P.main([str(f), "--counters", "--cus", "2", "--engines", "1"])
    assert "| `ak::k_sweep<2, false, false, true>` | 1 | 1.0 | 100.0 | 500.0 | 2.00 | 50.0% | 10.0% |" \
        in capsys.readouterr().out
