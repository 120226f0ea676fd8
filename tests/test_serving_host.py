"""Host logic of the serving loop that needs no GPU: the request stream and the
summary statistics."""
import numpy as np

from jf_2026_05_25_amd import serving


def test_poisson_stream_shape_and_determinism():
    a = serving.poisson_stream(24, rate_per_s=50.0, seed=4, dof_range=(2e3, 2e4))
    b = serving.poisson_stream(24, rate_per_s=50.0, seed=4, dof_range=(2e3, 2e4))
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    for x, y in zip(a, b):
        assert x.arrival == y.arrival
        assert np.array_equal(x.model.loads, y.model.loads)
        assert np.array_equal(x.model.xyz, y.model.xyz)
    dofs = np.array([r.model.ndof for r in a])
    assert dofs.min() >= 1000 and dofs.max() <= 4e4
    nrhs = np.array([r.model.nrhs for r in a])
    assert nrhs.min() >= 1 and nrhs.max() <= 16
    frac = [len(r.model.trias) / max(1, len(r.model.trias) + len(r.model.quads)) for r in a]
    assert 0.03 < np.mean(frac) < 0.2


def test_keys_separate_topology_sections_and_constraints():
    a = serving.poisson_stream(1, 10.0, seed=2, dof_range=(2e3, 3e3))[0].model
    from jf_2026_05_25_amd import meshes
    b = meshes.Model(**{**a.__dict__})
    b.quad_props = a.quad_props.copy(); b.quad_props[0, meshes.SP_H] *= 2
    c = meshes.Model(**{**a.__dict__})
    c.spc = a.spc.copy(); c.spc[-1] = 1
    assert serving.topology_key(a) == serving.topology_key(b) == serving.topology_key(c)
    assert len({serving.matrix_key(a), serving.matrix_key(b), serving.matrix_key(c)}) == 3


def test_summarize():
    R = serving.Result
    res = [R(i, np.zeros(6), np.array([1e-12]), np.array([1]), 0.1 * i, 0.1 * i, 0.1 * i + 0.05 * (i + 1),
             6, 6, 2, True) for i in range(10)]
    res.append(R(10, None, None, None, 0.0, 0.0, 0.2, 6, 0, 2, False, "err"))
    s = serving.summarize(res)
    assert s["requests"] == 11 and s["ok"] == 10 and s["solves"] == 20
    lat = [0.05 * (i + 1) for i in range(10)]
    assert abs(s["p50_latency_ms"] - 1e3 * np.percentile(lat, 50)) < 1e-9
    assert abs(s["p99_latency_ms"] - 1e3 * np.percentile(lat, 99)) < 1e-9
    assert abs(s["makespan_s"] - (0.9 + 0.5)) < 1e-12
