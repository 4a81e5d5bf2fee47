"""Sweep tiles change which thread updates a cell, never what it computes: the autotuned
two-level case and a tuned batch give the same histories and populations, bit for bit, as
the library's default tiles, and tuning leaves a case's state as it was."""
import numpy as np
import pytest

from aerokinetic_amd import serving as S

pytestmark = pytest.mark.gpu


def _solver(req, storage, tile=None):
    from aerokinetic_amd.refined import RefinedSolver
    project, settings, plan = S.project_of(req)
    return RefinedSolver(project, settings, plan=plan, storage=storage, tile=tile)


def _reqs(n):
    return S.polar_requests(n, seed=11, chord=(20, 24), span=(8, 8), reynolds=(1e4, 3e4), alpha=(2.0, 8.0))


def _loads(recs):
    return np.array([[r["step"], r["cl"], r["cd"], r["cm"]] for r in recs])


@pytest.mark.parametrize("storage", ["fp64", "fp16"])
def test_autotuned_two_level_case_matches_default_tiles(storage):
    req = _reqs(1)[0]
    print("request", storage, req, flush=True)
    a = _solver(req, storage)
    b = _solver(req, storage, tile="auto")
    print("tiles", a.tile, b.tile, flush=True)
    assert b.steps == 0
    for (bx, by) in b.tile:
        assert (bx == 0 and by % 64 == 0) or (bx % 64 == 0 and bx * by <= 1024)
    for level in (0, 1):  # tuning restored the initial state
        np.testing.assert_array_equal(a.populations(level), b.populations(level))
    ra, rb = a.advance(12, 1), b.advance(12, 1)
    np.testing.assert_array_equal(_loads(ra), _loads(rb))
    for level in (0, 1):
        np.testing.assert_array_equal(a.populations(level), b.populations(level))
    np.testing.assert_array_equal(np.asarray(a.budget()), np.asarray(b.budget()))
    a.close()
    b.close()


def test_tuned_batch_matches_single_cases():
    """Three members tuned over the batched tiles (real, sampled steps), then stepped on: every
    member's history and final populations equal the single-case solver's."""
    from aerokinetic_amd.batch import BATCH_TILES, RefinedBatch
    storage = "fp32"
    reqs = _reqs(3)
    singles = [_solver(r, storage) for r in reqs]
    members = [_solver(r, storage) for r in reqs]
    batch = RefinedBatch(storage)
    for m in members:
        batch.add(m)
    tile, timings, rows = batch.tune(steps=2)
    print("batch tile", tile, timings, flush=True)
    assert tile in BATCH_TILES and set(timings) == set(BATCH_TILES)
    recs, failed = batch.advance(4)
    assert not failed
    tuned_steps = 4 * len(BATCH_TILES)
    for s, m in zip(singles, members):
        want = _loads(s.advance(tuned_steps + 4, 1))
        got = np.concatenate([rows[m][:, :4], _loads(recs[m])])
        np.testing.assert_array_equal(got, want)
        for level in (0, 1):
            np.testing.assert_array_equal(s.populations(level), m.populations(level))
    batch.close()


def test_persistent_batched_sweep_matches_one_block_per_tile():
    """The batched sweep as a persistent grid of 64 blocks looping over the members' tiles gives
    the same histories and populations as one block per tile."""
    from aerokinetic_amd.batch import RefinedBatch
    storage = "fp16"
    reqs = _reqs(3)
    plain = [_solver(r, storage) for r in reqs]
    looped = [_solver(r, storage) for r in reqs]
    a, b = RefinedBatch(storage), RefinedBatch(storage, sweep_grid=64)
    for m in plain:
        a.add(m)
    for m in looped:
        b.add(m)
    ra, fa = a.advance(6)
    rb, fb = b.advance(6)
    assert not fa and not fb
    for p, q in zip(plain, looped):
        np.testing.assert_array_equal(_loads(ra[p]), _loads(rb[q]))
        for level in (0, 1):
            np.testing.assert_array_equal(p.populations(level), q.populations(level))
    a.close()
    b.close()


def test_per_form_tiles_and_wave_caps_match_default_launch():
    """Each AA form may run its own tile and cap on resident waves (the cap only reserves LDS
    per block): a run with different launches for the two forms, switched mid-run, equals the
    library's default launch bit for bit, in every storage and with sampled steps."""
    from aerokinetic_amd.config import Flow, Grid, Plan
    from aerokinetic_amd.geometry import Section, naca0012
    from aerokinetic_amd.solver import Solver
    section = Section((naca0012(41),), 1.0, (0.25, 0.0))
    flow = Flow(reynolds=10000, alpha_deg=4.0, lattice_speed=0.05, perturbation=1e-3)
    grid = Grid(cells_per_chord=32, upstream_chords=1.0, downstream_chords=2.0, transverse_chords=1.0,
                span_chords=0.25, memory_limit_gib=2, farfield="characteristic_transport_v1")
    plan = Plan.create(section, flow, grid)
    for storage in ("fp64", "fp32", "fp16"):
        print("per-form launches", storage, plan.shape, flush=True)
        a = Solver(section, flow, grid, plan=plan, storage=storage)
        b = Solver(section, flow, grid, plan=plan, storage=storage)
        b.lattice.set_tile(256, 1, odd=True)
        b.lattice.set_waves(3, odd=True)
        b.lattice.set_tile(64, 1, odd=False)
        b.lattice.set_waves(5, odd=False)
        ra, rb = a.advance(7, 2), b.advance(7, 2)
        b.lattice.set_tile(0, 256, odd=False)  # a new even tile drops only the even form's cap
        b.lattice.set_waves(4, odd=False)
        ra += a.advance(6, 1)
        rb += b.advance(6, 1)
        keys = ("step", "cl", "cd", "cm", "local_mach_max", "density_error_max", "density_min", "nonfinite")
        np.testing.assert_array_equal(np.array([[r[k] for k in keys] for r in ra]),
                                      np.array([[r[k] for k in keys] for r in rb]))
        np.testing.assert_array_equal(a.populations(), b.populations())
        with pytest.raises(Exception, match="waves per SIMD"):
            b.lattice.set_waves(9, odd=True)
        a.lattice.close()
        b.lattice.close()
