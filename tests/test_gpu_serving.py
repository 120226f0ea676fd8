"""Serving loop on the GPU: every request's result is the same bits whether it is
solved alone or in a mix of concurrent requests; residuals meet the tolerance;
admission by free HBM serialises requests when memory is short; a failing request
does not stop the stream."""
import numpy as np
import pytest

from jf_2026_05_25_amd import meshes

pytestmark = pytest.mark.gpu


def _stream(n=10, seed=11):
    from jf_2026_05_25_amd import serving
    reqs = serving.poisson_stream(n, rate_per_s=1e4, seed=seed, dof_range=(2e3, 2e4), rhs_range=(1, 6))
    # same topology with new sections (shares pattern + symbolic) and same matrix with new loads (factor reuse)
    base = reqs[0].model
    m2 = meshes.Model(**{**base.__dict__})
    m2.quad_props = base.quad_props.copy(); m2.quad_props[:, meshes.SP_H] *= 1.3
    m2.loads = base.loads[:, ::-1].copy()
    m3 = meshes.Model(**{**base.__dict__})
    m3.loads = base.loads * 0.5
    reqs.append(serving.Request(len(reqs), m2, reqs[-1].arrival, 32))
    reqs.append(serving.Request(len(reqs), m3, reqs[-1].arrival, 32))
    return reqs


def test_mix_equals_alone_bitwise():
    from jf_2026_05_25_amd import serving
    reqs = _stream()
    srv = serving.SolveServer(slots=4)
    try:
        mixed = srv.serve(reqs, realtime=False)
        stats = serving.summarize(mixed)
        again = srv.serve([serving.Request(0, reqs[-1].model, 0.0, 32)], realtime=False)[0]
    finally:
        srv.close()
    assert stats["ok"] == len(reqs), [r.error for r in mixed if not r.ok]
    assert stats["max_relative_residual"] <= 1e-10
    assert again.cache_hit                                   # device-resident factor reused
    assert np.array_equal(again.u.view(np.int64), mixed[-1].u.view(np.int64))
    for r in reqs:
        solo = serving.SolveServer(slots=1)
        try:
            a = solo.serve([serving.Request(r.rid, r.model, 0.0, r.precision)], realtime=False)[0]
        finally:
            solo.close()
        b = mixed[r.rid]
        assert a.ok and b.ok
        assert np.array_equal(a.u.view(np.int64), b.u.view(np.int64)), r.rid
        assert np.array_equal(a.iters, b.iters)


def test_results_match_oracle(oracle):
    from jf_2026_05_25_amd import serving
    reqs = _stream(n=4, seed=5)[:4]
    srv = serving.SolveServer(slots=2)
    try:
        res = srv.serve(reqs, realtime=False)
    finally:
        srv.close()
    for q, r in zip(reqs, res):
        _, fixed, uo, rro, _ = oracle.sol101(q.model, refine=10)
        assert r.ok
        err = np.linalg.norm(r.u - uo) / np.linalg.norm(uo)
        assert err <= 1e-9, (q.rid, err)      # FP32 factor stopped at residual 1e-10
        assert np.max(r.relres) <= 1e-10


def test_admission_serialises_when_memory_is_short():
    from jf_2026_05_25_amd import serving
    reqs = _stream(n=6, seed=9)[:6]
    srv = serving.SolveServer(slots=4)
    try:
        free, _ = srv.ctxs[0].memory()
        srv.reserve = free                 # nothing fits beside a running request
        res = srv.serve(reqs, realtime=False)
        assert all(r.ok for r in res)
        assert srv.peak_active == 1
        assert srv.committed == 0 and srv.active == 0
    finally:
        srv.close()


def test_failing_request_does_not_stop_stream():
    from jf_2026_05_25_amd import serving
    good = _stream(n=3, seed=3)[:3]
    bad = meshes.flat_plate(4, 4)
    bad.xyz = bad.xyz.copy(); bad.xyz[6] = bad.xyz[7]          # collapsed edge -> degenerate CQUAD4
    reqs = good + [serving.Request(3, bad, 0.0, 32)]
    srv = serving.SolveServer(slots=2)
    try:
        res = srv.serve(reqs, realtime=False)
    finally:
        srv.close()
    assert [r.ok for r in res] == [True, True, True, False]
    assert "JfemError" in res[3].error
