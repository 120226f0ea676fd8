"""jfem_solve_dev_static (the solve that allocates nothing) against jfem_solve_dev:
the same results bit for bit, on two calls in a row with the second smaller,
in both factor precisions; the context's own scratch is never grown; a solve
whose scratch exceeds the static array is refused with BAD_ARG."""
import numpy as np
import pytest

from jf_2026_05_25_amd import meshes

pytestmark = pytest.mark.gpu


def _setup(ctx, m, precision):
    from jf_2026_05_25_amd import solver
    asm = solver.assemble_stiffness(m, ctx)
    fixed = solver.bc_partition(asm.K, m.spc)
    S = solver.Symbolic(asm.K, asm.mesh, fixed, ctx)
    F = solver.Factor(S, asm.K, precision, ctx)
    Bc = np.ascontiguousarray(m.loads.T)
    b = solver.DeviceBuffer(Bc.nbytes, ctx)
    b.upload(Bc)
    return asm, S, F, b, Bc


def _solve(ctx, F, K, b, Bc, nrhs, static):
    from jf_2026_05_25_amd import solver
    x = solver.DeviceBuffer(Bc.nbytes, ctx)
    rr, it, ok = F.solve_device(K, b, x, nrhs, tol=1e-10, max_refine=200, ctx=ctx, static_scratch=static)
    out = x.download(np.zeros_like(Bc))
    x.free()
    return out, rr, it, ok


@pytest.mark.parametrize("precision", [32, 64])
def test_static_solve_matches_solve_dev_on_two_calls(precision):
    from jf_2026_05_25_amd import solver
    ctx_static = solver.Context(0)          # a fresh context: its scratch must stay empty
    ctx_ref = solver.Context(0)
    try:
        # first a larger request, then a smaller one (fewer DOFs and right-hand sides):
        # the second must not see anything the first left in the shared array
        for m in (meshes.wing_box(48, 8, nrhs=6, tria_fraction=0.1), meshes.wing_box(12, 4, nrhs=2)):
            print(f"no-alloc solve: {m.name} ndof={m.ndof} nrhs={m.nrhs} fp{precision}", flush=True)
            asm, S, F, b, Bc = _setup(ctx_ref, m, precision)
            need = S.solve_scratch_bytes(m.nrhs, precision)
            assert 0 < need < 128 << 20
            before = ctx_static.device_bytes()
            xs, rrs, its, oks = _solve(ctx_static, F, asm.K, b, Bc, m.nrhs, True)
            assert ctx_static.device_bytes() == before
            xr, rrr, itr, okr = _solve(ctx_ref, F, asm.K, b, Bc, m.nrhs, False)
            assert oks and okr and np.max(rrs) <= 1e-10
            assert np.array_equal(xs.view(np.int64), xr.view(np.int64))
            assert np.array_equal(its, itr) and np.array_equal(rrs, rrr)
            b.free(); F.free(); S.free()
    finally:
        ctx_static.close()
        ctx_ref.close()


def test_static_solve_refuses_oversized_scratch():
    from jf_2026_05_25_amd import solver, _lib
    ctx = solver.Context(0)
    try:
        m = meshes.wing_box(128, 32, nrhs=16)
        asm, S, F, b, Bc = _setup(ctx, m, 32)
        assert S.solve_scratch_bytes(16, 32) > 128 << 20
        x = solver.DeviceBuffer(Bc.nbytes, ctx)
        with pytest.raises(_lib.JfemError) as e:
            F.solve_device(asm.K, b, x, 16, static_scratch=True)
        assert e.value.code == 5 and "static scratch" in str(e.value)
        x.free(); b.free(); F.free(); S.free()
    finally:
        ctx.close()
