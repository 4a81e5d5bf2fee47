"""Population layout fallback: lattices whose z plane would exceed 2 GiB with interleaved
direction rows keep one plane per direction.  The kernels address both layouts through the
same strides, so a window of a large lattice (one plane per direction) must evolve
bit-identically to a small lattice (interleaved rows) holding the same initial field, away
from both boxes' boundaries (information travels one cell per step)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DX = 1.0 / 64.0  # exact: cell centres of both boxes are the same binary fractions


def _seed(U=0.05):
    return {"U": U, "amp": 2e-3, "phase_y": 0.3, "phase_z": 0.1, "ky": 2 * math.pi / 2.0,
            "kz": 2 * math.pi / (2 * DX), "lex": 31.2, "ley": 17.0}


def _lattice(nx, ny, nz, origin):
    from aerokinetic_amd.lattice import Lattice
    lat = Lattice(nx, ny, nz, storage="fp64", omega_shear=1.6, omega_bulk=1.0)
    lat.set_solid(np.zeros((nx, ny), dtype=np.uint8))
    lat.init_seeded(_seed(), origin, DX)
    lat.set_farfield("legacy", 0.05)
    return lat


def test_one_plane_per_direction_window_matches_interleaved_rows():
    big_shape, small_shape = (4096, 2432, 2), (256, 192, 2)
    x0, y0 = 1900, 1100  # small cell (x, y) is big cell (x + x0, y + y0); the seeded
    # envelope (centre lex + 0.5) sits in the middle of the small box
    # the big lattice's z plane with interleaved rows would be 27 * 4112 * 2434 * 8 bytes > 2 GiB
    assert 27 * 4112 * (big_shape[1] + 2) * 8 >= 1 << 31
    assert 27 * 272 * (small_shape[1] + 2) * 8 < 1 << 31
    print("case layout fallback", big_shape, small_shape, flush=True)
    big = _lattice(*big_shape, (0.0, 0.0))
    small = _lattice(*small_shape, (x0 * DX, y0 * DX))
    steps = 4
    big.run(steps, 0, 0)
    small.run(steps, 0, 0)
    m = steps + 2
    for z in range(2):
        rb, ub = big.plane(z)
        rs, us = small.plane(z)
        wb = (slice(x0 + m, x0 + small_shape[0] - m), slice(y0 + m, y0 + small_shape[1] - m))
        ws = (slice(m, small_shape[0] - m), slice(m, small_shape[1] - m))
        assert np.abs(us[ws][..., 1:]).max() > 1e-6  # a non-trivial field in the window
        np.testing.assert_array_equal(rb[wb], rs[ws])
        np.testing.assert_array_equal(ub[wb], us[ws])
    big.close()
    small.close()
