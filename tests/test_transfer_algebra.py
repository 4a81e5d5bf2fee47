"""Closed-form pins of the 2:1 transfer algebra of the oracle (restating the
reference's tests/test_refinement.py:24-209 and tests/test_interface.py:35-70)."""
import numpy as np
import pytest

from aerokinetic_amd.d3q27 import C, W
from aerokinetic_amd.refined_plan import AcousticLevel
from oracle import refined_oracle as R


def brute_moments(pdf):
    f = pdf + W
    rho = f.sum(-1)
    j = f @ C
    out = [rho] + [j[..., a] for a in range(3)]
    for a, b in R.SECOND:
        raw = (f * C[:, a] * C[:, b]).sum(-1)
        out.append(raw - rho * R.CS2 * (a == b) - j[..., a] * j[..., b] / rho)
    return np.stack(out, -1)


def random_state(shape, seed=0):
    rng = np.random.default_rng(seed)
    m = np.zeros((*shape, 10))
    m[..., 0] = 1 + 0.01 * rng.standard_normal(shape)
    m[..., 1:4] = 0.03 * rng.standard_normal((*shape, 3))
    m[..., 4:] = 1e-4 * rng.standard_normal((*shape, 6))
    return m


def test_moments_match_brute_force_and_roundtrip():
    m = random_state((5, 4, 3))
    pdf = R.populations(m)
    np.testing.assert_allclose(R.moments(pdf), m, atol=1e-15)
    np.testing.assert_allclose(brute_moments(pdf), m, atol=1e-14)
    for _ in range(100):
        pdf = R.populations(R.moments(pdf))
    np.testing.assert_allclose(R.moments(pdf), m, atol=1e-14)


def test_equilibrium_is_product_form():
    rho, u = 1.003, np.array([0.045, -0.025, 0.015])
    m = np.zeros(10)
    m[0], m[1:4] = rho, rho * u
    phi = lambda c, v: np.where(c == 0, 2 / 3 - v * v, (1 / 3 + v * v + c * v) / 2)
    expect = rho * phi(C[:, 0], u[0]) * phi(C[:, 1], u[1]) * phi(C[:, 2], u[2]) - W
    np.testing.assert_allclose(R.populations(m), expect, atol=3e-16)


def test_prolongation_conserves_parents_and_reproduces_quartics():
    rng = np.random.default_rng(1)
    v = rng.standard_normal((9, 8, 7, 2))
    child = R.prolong_cell_averages(v)
    np.testing.assert_allclose(R.restrict_cell_averages(child), v[2:-2, 2:-2, 2:-2], atol=1e-14)
    # exact cell averages of a degree-4 polynomial (one axis at a time)
    x = np.arange(-2, 7) + 0.5
    def avg(a, b, p):
        P = np.polyint(p)
        return (np.polyval(P, b) - np.polyval(P, a)) / (b - a)
    p = np.array([0.3, -0.2, 0.5, 1.0, -0.7])
    parents = avg(x - 0.5, x + 0.5, p)
    vals = np.broadcast_to(parents[:, None, None], (9, 5, 5)).copy()
    kids = R.prolong_cell_averages(vals)[:, 0, 0]
    xc = np.arange(0, 10) * 0.5 + 0.25 - 0.0
    lo = np.repeat(np.arange(0, 5), 2) + np.tile([0.0, 0.5], 5)
    np.testing.assert_allclose(kids, avg(lo, lo + 0.5, p), atol=1e-13)


def test_stress_rescale_is_acoustic_and_invertible():
    c = AcousticLevel(1.0, 1.0, 0.55, 0.83)
    f = c.refined()
    assert f.shear_viscosity == pytest.approx(c.shear_viscosity)
    assert f.bulk_viscosity == pytest.approx(c.bulk_viscosity)
    m = random_state((4,))
    back = R.rescale_stress(R.rescale_stress(m, c, f), f, c)
    np.testing.assert_allclose(back, m, atol=1e-16)
    up = R.rescale_stress(m, c, f)
    np.testing.assert_allclose(up[..., 4:].sum(-1) * 0, 0)
    dev = m[..., 4:7] - m[..., 4:7].mean(-1, keepdims=True)
    np.testing.assert_allclose(up[..., 4:7] - up[..., 4:7].mean(-1, keepdims=True),
                               dev * f.tau_shear * f.dt / (c.tau_shear * c.dt), atol=1e-17)


def test_face_flux_equals_direct_pull_streaming_balance():
    """Sum of face fluxes around a block = change of its content under pull streaming."""
    rng = np.random.default_rng(3)
    shape = (8, 7, 6)
    post = 1e-3 * rng.standard_normal((*shape, 27))
    lo, hi = (2, 2, 1), (5, 6, 4)
    streamed = np.zeros_like(post)
    for i, c in enumerate(C):
        streamed[..., i] = np.roll(post[..., i], shift=tuple(c), axis=(0, 1, 2))
    content = lambda p: np.concatenate([[p[..., :].sum()], (p.sum(axis=(0, 1, 2)) @ C)]) if p.ndim == 4 else None
    blk = tuple(slice(a, b) for a, b in zip(lo, hi))
    before = np.r_[post[blk].sum(), post[blk].sum(axis=(0, 1, 2)) @ C]
    after = np.r_[streamed[blk].sum(), streamed[blk].sum(axis=(0, 1, 2)) @ C]
    net = np.zeros(4)
    for axis in range(3):
        others = [a for a in range(3) if a != axis]
        ranges = [(lo[a], hi[a]) for a in others]
        low = R.FaceFlux(shape, axis, lo[axis] - 1, ranges, periodic=True)(post).sum(axis=(0, 1))
        high = R.FaceFlux(shape, axis, hi[axis] - 1, ranges, periodic=True)(post).sum(axis=(0, 1))
        net += low - high
    np.testing.assert_allclose(after - before, net, atol=1e-15)


def test_negative_populations_are_rejected():
    m = np.zeros(10)
    m[0] = 1.0
    m[4] = -1.0
    with pytest.raises(R.TransferError):
        R.populations(m)
