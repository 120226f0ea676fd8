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


# ---------------------------------------------------------------- residency with a stub backend
class _Stub:
    """Device objects of a stub backend: they record their frees and flag a free
    that lands while a solve runs on the object or while a live object still
    references it (``deps``: what an object's device structures point to)."""
    registry = []
    lock = None

    def __init__(self, kind, nbytes=1000, deps=()):
        import threading
        self.kind, self.nbytes, self.deps = kind, nbytes, list(deps)
        self.frees = 0
        self.in_solve = 0
        self.bad_free = False
        with _Stub.lock:
            _Stub.registry.append(self)

    def device_bytes(self):
        return self.nbytes

    def free(self):
        with _Stub.lock:
            live_users = [o for o in _Stub.registry if o.frees == 0 and o is not self and
                          any(d is self for d in o.deps)]
        if self.in_solve or live_users:
            self.bad_free = True
        self.frees += 1


def _stub_backend(free_bytes=1 << 40):
    import threading
    import types
    _Stub.registry = []
    _Stub.lock = threading.Lock()

    class Ctx:
        def __init__(self, device=0):
            pass

        def memory(self):
            return free_bytes - sum(o.nbytes for o in _Stub.registry if o.frees == 0), 1 << 41

        def device_bytes(self):
            return 0

        def close(self):
            pass

    class Sym(_Stub):
        def __init__(self, K, mesh, fixed, ctx):
            super().__init__("S", 500, deps=[K.pattern])
            self.fixed = fixed
            self.info = {"nnz_L": 100}

        def factor_bytes(self, precision):
            return 4000

        def solve_scratch_bytes(self, nrhs, precision):
            return 100

    class Fac(_Stub):
        def __init__(self, S, K, precision):
            super().__init__("F", 4000, deps=[S])
            self.S, self.K = S, K

        def solve(self, K, B, tol, max_refine, ctx):
            import time
            self.in_solve += 1
            self.S.in_solve += 1
            K.in_solve += 1
            try:
                time.sleep(0.002)
                assert self.frees == 0 and self.S.frees == 0 and K.frees == 0, "solve on a freed object"
            finally:
                self.in_solve -= 1
                self.S.in_solve -= 1
                K.in_solve -= 1
            n = B.shape[1]
            return np.zeros_like(B), np.full(n, 1e-12), np.ones(n, np.int32), True

    def matrix(mesh, pattern, ctx):
        K = _Stub("K", 800, deps=[pattern])
        K.pattern = pattern
        return K

    B = types.SimpleNamespace(
        Context=Ctx,
        DeviceMesh=lambda model, ctx: _Stub("mesh", 200),
        Pattern=lambda mesh: _Stub("pattern", 300),
        DeviceMatrix=matrix,
        Symbolic=Sym,
        bc_partition=lambda K, spc: np.asarray(spc, np.uint8).copy(),
        factor_with_fallback=lambda S, K, precision, ctx: (Fac(S, K, precision), precision),
        small_pivot_dofs=lambda S, K, F, ctx=None: (np.zeros(0, np.int64), 64),
    )
    return B


def _variants(n_topo=3, n_props=4, repeats=3):
    from jf_2026_05_25_amd import meshes
    reqs = []
    for t in range(n_topo):
        base = meshes.flat_plate(2 + t, 2)
        for p in range(n_props):
            m = meshes.Model(**{**base.__dict__})
            m.quad_props = base.quad_props.copy()
            m.quad_props[:, meshes.SP_H] *= 1.0 + 0.1 * p
            for r in range(repeats):
                reqs.append(serving.Request(len(reqs), m, 0.0, 32 if (p + r) % 2 else 64))
    return reqs


def test_eviction_never_frees_an_object_in_use():
    B = _stub_backend()
    srv = serving.SolveServer(slots=4, backend=B, max_cached_factors=2, max_cached_topologies=1, cache_bytes=1 << 30)
    res = srv.serve(_variants(), realtime=False)
    assert all(r.ok for r in res), [r.error for r in res if not r.ok]
    assert srv.evictions > 0
    assert srv.committed == 0 and srv.active == 0 and srv.stage_holders == [0, 0]
    srv.close()
    objs = _Stub.registry
    assert not any(o.bad_free for o in objs)
    assert all(o.frees == 1 for o in objs), sorted({(o.kind, o.frees) for o in objs})


def test_cache_hits_keep_precisions_apart():
    B = _stub_backend()
    srv = serving.SolveServer(slots=1, backend=B, max_cached_factors=64, cache_bytes=1 << 30)
    reqs = _variants(n_topo=2, n_props=2, repeats=3)
    res = srv.serve(reqs, realtime=False)
    # per matrix: FP64, FP32, FP64 -> the third is the only hit
    assert [r.cache_hit for r in res] == [False, False, True] * 4
    assert {k[1] for k in srv.factors} == {32, 64} and len(srv.factors) == 8
    srv.close()
    assert all(o.frees == 1 for o in _Stub.registry)


def test_factor_key_separates_precision():
    from jf_2026_05_25_amd import meshes
    m = meshes.flat_plate(3, 3)
    assert serving.factor_key(m, 32) != serving.factor_key(m, 64)
    assert serving.factor_key(m, 32)[0] == serving.matrix_key(m)


def test_memory_pressure_evicts_idle_entries_and_serialises():
    # free memory covers about one request's factor beyond the reserve
    B = _stub_backend(free_bytes=12000)
    srv = serving.SolveServer(slots=4, backend=B, reserve_bytes=0, cache_bytes=1 << 30)
    res = srv.serve(_variants(n_topo=2, n_props=3, repeats=1), realtime=False)
    assert all(r.ok for r in res)
    assert srv.evictions > 0
    srv.close()
    assert all(o.frees == 1 for o in _Stub.registry)
    assert not any(o.bad_free for o in _Stub.registry)


def test_pre_admission_estimate_bounds_allocations():
    from jf_2026_05_25_amd import meshes
    m = meshes.flat_plate(6, 5, tria_fraction=0.3)
    ntrip = 576 * len(m.quads) + 324 * len(m.trias)
    est = serving.estimate_pre_bytes(m, True)
    # K values (nnz <= ntrip) + element buffer, pattern lists, mesh
    assert est >= 16 * ntrip + 24 * m.n_nodes
    assert serving.estimate_pre_bytes(m, False) < est
