"""ak_diagnostics_noalloc (no allocation; partials in a file-scope device array) against
ak_diagnostics, through the C ABI: bit-identical results on two calls in a row, the second on
a smaller lattice, and on a two-level case's level (pre-collision state)."""
import numpy as np
import pytest

from aerokinetic_amd import _lib

pytestmark = pytest.mark.gpu


def _both(h):
    L = _lib.load()
    a, b = np.zeros(5), np.zeros(5)
    _lib.check(L.ak_diagnostics(h, _lib.dptr(a)), h, "ak_diagnostics")
    _lib.check(L.ak_diagnostics_noalloc(h, _lib.dptr(b)), h, "ak_diagnostics_noalloc")
    return a, b


def test_noalloc_diagnostics_equal_the_allocating_entry_point():
    from aerokinetic_amd import cases
    from aerokinetic_amd.refined import RefinedSolver
    from aerokinetic_amd.solver import Solver
    # first call: the plain 384x256x8 lattice (3 x 4096 cells per block of the stride)
    section, flow, grid, plan = cases.NACA0012_C32_384x256x8.build()
    big = Solver(section, flow, grid, plan=plan)
    big.advance(5, 0)
    print("lattice", plan.shape, flush=True)
    a, b = _both(big.lattice.h)
    print("allocating", a, "no-alloc", b, flush=True)
    np.testing.assert_array_equal(a, b)
    assert a[4] > 0.99 * int(big._fluid.sum()) and a[0] > 0
    big.lattice.close()
    # second call: a smaller lattice (the array keeps the first call's partials)
    small = cases.NACA0012_C32_384x256x8.with_shape((96, 64, 8), (32, 32))
    section, flow, grid, plan = small.build()
    s = Solver(section, flow, grid, plan=plan, storage="fp32")
    s.advance(3, 0)
    print("lattice", plan.shape, flush=True)
    a2, b2 = _both(s.lattice.h)
    np.testing.assert_array_equal(a2, b2)
    assert a2[4] < a[4] / 10
    s.lattice.close()
    # a level of a two-level case (pre-collision state)
    from aerokinetic_amd import serving as S
    project, settings, rplan = S.project_of(S.polar_requests(1, seed=5, chord=(16, 16), span=(8, 8))[0])
    r = RefinedSolver(project, settings, plan=rplan, storage="fp16")
    r.advance(2, 1)
    for level in r._levels:
        a3, b3 = _both(level)
        np.testing.assert_array_equal(a3, b3)
    r.close()
