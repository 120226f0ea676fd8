"""Device self-tests of kernel building blocks (matrix-core lane layouts)."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [(64, 1e-13), (32, 1e-5)])
def test_mfma_update_tile_matches_scalar(precision, tol):
    from jf_2026_05_25_amd import solver, _lib
    ctx = solver.Context(0)
    d = ctypes.c_double()
    _lib.check(ctx.h, ctx.L.jfem_selftest_update(ctx.h, precision, ctypes.byref(d)))
    ctx.close()
    assert d.value < tol, d.value
