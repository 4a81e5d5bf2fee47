"""Batched two-level cases: a member gives the same result alone and in any mix, a batch
of one equals the single-case solver, and the continuous-batching service completes a
stream of requests."""
import numpy as np
import pytest

from aerokinetic_amd import serving as S

pytestmark = pytest.mark.gpu


def _solver(req, storage, graphs=True):
    from aerokinetic_amd.refined import RefinedSolver
    project, settings, plan = S.project_of(req)
    return RefinedSolver(project, settings, plan=plan, storage=storage, graphs=graphs)


def _reqs():
    return S.polar_requests(3, seed=5, chord=(16, 20), span=(8, 8), reynolds=(1e4, 2e4), alpha=(2.0, 8.0))


def _loads(recs):
    return np.array([[r["step"], r["cl"], r["cd"], r["cm"]] for r in recs])


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_batch_of_one_matches_single_case_solver(storage):
    """The single case ends a step with one fused launch whose reflux and restriction write
    the next step's donor moments from registers (storage rounding applied there); the batch
    reads them back from memory in its own launches.  Both must agree bit for bit."""
    from aerokinetic_amd.batch import RefinedBatch
    req = _reqs()[0]
    print("request", storage, req, flush=True)
    single = _solver(req, storage)
    member = _solver(req, storage)
    batch = RefinedBatch(storage)
    batch.add(member)
    a = single.advance(9, 1)
    recs, failed = batch.advance(9)
    assert not failed
    b = recs[member]
    np.testing.assert_array_equal(_loads(a), _loads(b))
    for level in (0, 1):
        ga, gb = single.populations(level), member.populations(level)
        np.testing.assert_array_equal(ga, gb)
    batch.close()


@pytest.mark.parametrize("storage", ["fp32", "fp64", "fp16"])
def test_member_is_identical_alone_and_in_a_mix(storage):
    """Case A alone vs A sharing every batched launch with B (joined from the start) and C
    (joined after 4 steps, different size and age): A's loads and populations are
    bit-identical."""
    from aerokinetic_amd.batch import RefinedBatch
    ra, rb, rc = _reqs()
    print("requests", ra.id, rb.id, rc.id, storage, flush=True)
    alone_a = _solver(ra, storage)
    alone = RefinedBatch(storage)
    alone.add(alone_a)
    mix_a, mix_b, mix_c = _solver(ra, storage), _solver(rb, storage), _solver(rc, storage)
    mix_c.advance(3, 0)  # C is older: a different AA parity when it joins
    mix = RefinedBatch(storage)
    mix.add(mix_b)
    mix.add(mix_a)
    r1, _ = alone.advance(4)
    r2, _ = mix.advance(4)
    mix.add(mix_c)
    r3, _ = alone.advance(6)
    r4, f4 = mix.advance(6)
    assert not f4
    np.testing.assert_array_equal(_loads(r1[alone_a] + r3[alone_a]), _loads(r2[mix_a] + r4[mix_a]))
    assert [r["step"] for r in r4[mix_c]] == list(range(4, 10))
    mix.remove(mix_b)
    r5, _ = alone.advance(3)
    r6, _ = mix.advance(3)
    np.testing.assert_array_equal(_loads(r5[alone_a]), _loads(r6[mix_a]))
    for level in (0, 1):
        np.testing.assert_array_equal(alone_a.populations(level), mix_a.populations(level))
    # C after 3 own steps + 9 in the mix equals C stepped alone 12 times
    ref_c = _solver(rc, storage)
    ref_c.advance(12, 0)
    np.testing.assert_array_equal(ref_c.populations(1), mix_c.populations(1))
    alone.close()
    mix.close()


def test_batch_graph_replay_matches_direct_launches():
    from aerokinetic_amd.batch import RefinedBatch
    reqs = _reqs()[:2]
    out = []
    for graphs in (False, True):
        b = RefinedBatch("fp32", graphs=graphs)
        members = [_solver(r, "fp32") for r in reqs]
        for m in members:
            b.add(m)
        recs, _ = b.advance(5)
        out.append(([_loads(recs[m]) for m in members], [m.populations(1) for m in members]))
        b.close()
    for x, y in zip(out[0][0] + out[0][1], out[1][0] + out[1][1]):
        np.testing.assert_array_equal(x, y)


def test_service_admits_steps_and_retires_a_request_stream():
    # every request is available at t = 0 and the two builders prepare two of them side by
    # side, so the second joins while the first still runs (a case lives >= one window)
    reqs = S.polar_requests(4, seed=9, rate=None, chord=(16, 18), span=(8, 8), reynolds=(1e4, 2e4))
    print("arrivals", [round(r.arrival, 3) for r in reqs], flush=True)
    srv = S.Server(storage="fp32", chunk=40, window=400, tol=0.005, max_steps=800, builders=2)
    rep = srv.serve(reqs)
    s = rep.summary()
    assert s["cases"] == 4 and sum(s["reasons"].values()) == 4
    assert s["reasons"]["failed"] == 0
    assert s["peak_members"] >= 2
    for r in rep.results:
        assert r.latency > 0 and 400 <= r.steps <= 800 and np.isfinite(r.cl)


def test_polar_sweep_sized_members_identical_alone_and_in_a_mix():
    """Batched polar-sweep sizes (FP32 storage): three members drawn from the sweep's
    distribution -- NACA 4-digit camber 0-4 %, thickness 9-15 %, chord 32-96 coarse cells,
    span 8-32 cells, Re 1e4-1e5 -- one of them at the largest chord; each member's loads and
    populations after 6 coarse steps in the mix are bit-identical to the member stepped alone."""
    from aerokinetic_amd.batch import RefinedBatch
    reqs = S.polar_requests(3, seed=2026, chord=(32, 96), span=(8, 32))
    reqs[2] = S.CaseRequest(2, 0.0, 0.03, 0.4, 0.12, 6.0, 5e4, 96, 32)  # the largest case of the sweep
    print("requests", [(r.chord_cells, r.span_cells) for r in reqs],
          "cells", [S.case_cells(r) for r in reqs], flush=True)
    assert all(32 <= r.chord_cells <= 96 and 8 <= r.span_cells <= 32 for r in reqs)
    mixed = [_solver(r, "fp32") for r in reqs]
    mix = RefinedBatch("fp32")
    for m in mixed:
        mix.add(m)
    got, failed = mix.advance(6)
    assert not failed
    mix.close()
    for req, m in zip(reqs, mixed):
        alone_m = _solver(req, "fp32")
        alone = RefinedBatch("fp32")
        alone.add(alone_m)
        ref, _ = alone.advance(6)
        np.testing.assert_array_equal(_loads(ref[alone_m]), _loads(got[m]))
        for level in (0, 1):
            np.testing.assert_array_equal(alone_m.populations(level), m.populations(level))
        alone.close()
        alone_m.close()
        m.close()


def test_service_serves_a_flapped_request_of_millions_of_cells():
    """A serving-workload request (main element + 25 %-chord flap, FP16 storage, a lattice of
    more than 2M cells) through the service alongside a small one: both retire by the rule."""
    reqs = S.serving_requests(1, seed=11, rate=None, cells=(2.2e6, 2.6e6), flap=(20.0, 30.0))
    reqs += S.polar_requests(1, seed=3, rate=None, chord=(16, 16), span=(8, 8))
    reqs[1] = S.CaseRequest(1, 0.0, reqs[1].camber, 0.4, reqs[1].thickness, reqs[1].alpha_deg, reqs[1].reynolds, 16, 8)
    big = reqs[0]
    print("flap request", big, "cells", S.case_cells(big), flush=True)
    assert S.case_cells(big) >= 2_000_000 and big.flap_deflection_deg >= 20.0
    srv = S.Server(storage="fp16", chunk=100, window=400, tol=0.005, max_steps=600, builders=2)
    rep = srv.serve(reqs)
    s = rep.summary()
    assert s["cases"] == 2 and s["reasons"]["failed"] == 0
    by_id = {r.id: r for r in rep.results}
    assert by_id[0].cells >= 2_000_000 and 400 <= by_id[0].steps <= 600
    assert np.isfinite([by_id[0].cl, by_id[0].cd, by_id[0].cm]).all()


def test_service_retires_a_converged_case_at_its_window():
    """Cl-window retirement on the GPU: a symmetric section at zero incidence (Cl stays within
    1e-5 of zero, below tol x the 0.1 floor) retires as converged at its first full window,
    while a lifting case at Re 1e4, which sheds vortices, runs to the step cap."""
    reqs = [S.CaseRequest(0, 0.0, 0.0, 0.4, 0.12, 0.0, 1e4, 16, 8),
            S.CaseRequest(1, 0.0, 0.0, 0.4, 0.12, 2.0, 1e4, 16, 8)]
    srv = S.Server(storage="fp32", chunk=100, window=300, tol=0.005, max_steps=1200, builders=2)
    rep = srv.serve(reqs)
    by_id = {r.id: r for r in rep.results}
    print({i: (r.reason, r.steps, r.cl_window_variation) for i, r in by_id.items()}, flush=True)
    assert by_id[0].reason == "converged" and 300 <= by_id[0].steps <= 400
    assert by_id[0].cl_window_variation <= 0.005 and abs(by_id[0].cl) < 1e-3
    assert by_id[1].reason == "max_steps" and by_id[1].steps == 1200
