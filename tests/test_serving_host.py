"""Host logic of the batched service: seeded requests, case geometry, footprints, the
Cl-window retirement rule (no GPU)."""
import numpy as np
import pytest

from aerokinetic_amd import serving as S
from aerokinetic_amd.walls import labels


def test_polar_requests_are_seeded_and_in_range():
    a = S.polar_requests(40, seed=2026, rate=4.0)
    b = S.polar_requests(40, seed=2026, rate=4.0)
    assert a == b
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0) and arr[0] > 0
    # mean inter-arrival of a Poisson process with rate 4/s
    assert 0.15 < np.diff(np.r_[0.0, arr]).mean() < 0.4
    for r in a:
        assert 0 <= r.camber <= 0.04 and 0.09 <= r.thickness <= 0.15 and -4 <= r.alpha_deg <= 12
        assert 1e4 <= r.reynolds <= 1e5 and 32 <= r.chord_cells <= 96 and 8 <= r.span_cells <= 32
    assert all(r.arrival == 0.0 for r in S.polar_requests(5, seed=1))


def test_flap_cases_have_two_resolved_separate_elements():
    for r in S.polar_requests(6, seed=11, chord=(16, 24), span=(8, 8), flap=(0.0, 30.0)):
        project, settings, plan = S.project_of(r)
        contours = project.section.normalized(project.flow.alpha_deg)
        assert len(contours) == 2
        lab = labels(contours, plan.shape[:2], plan.origin, plan.coarse_level.dx / 2)
        assert (lab == 1).sum() > 0 and (lab == 2).sum() > 0
        # the fine box holds both elements with the guard band
        assert plan.minimum_wall_interface_distance_coarse_cells >= settings.guard_coarse_cells


def test_footprint_scales_with_storage_and_size():
    r = S.polar_requests(2, seed=3, chord=(16, 16), span=(8, 8))[0]
    plan = S.project_of(r)[2]
    f16, f32, f64 = (S.footprint_bytes(plan, s) for s in ("fp16", "fp32", "fp64"))
    assert f16 < f32 < f64
    nx, ny, nz = plan.shape
    assert f64 > (nx * ny * nz) * 27 * 8


def test_window_rule():
    steady = [0.5 + 1e-4 * np.sin(k) for k in range(400)]
    assert S._window_variation(steady, 200, 0.1) < 0.005
    drifting = list(np.linspace(0.2, 0.5, 400))
    assert S._window_variation(drifting, 200, 0.1) > 0.005
    # near-zero lift uses the floor instead of the mean
    assert S._window_variation([1e-4, -1e-4] * 100, 200, 0.1) == pytest.approx(2e-4 / 0.1)
    srv = S.Server(window=200, tol=0.005, max_steps=1000)
    assert srv._retire_reason({"cl": steady}) == "converged"
    assert srv._retire_reason({"cl": drifting}) is None
    assert srv._retire_reason({"cl": list(np.linspace(0, 1, 1000))}) == "max_steps"
