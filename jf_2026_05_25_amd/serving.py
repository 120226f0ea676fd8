"""Continuous-batching serving loop for static solves on one GPU.

The reference answers one request at a time from a persistent worker
(utils/julia_tools/jfem_worker_jsonl.jl:61, warm_model_session.jl:21-61) and keeps a
factor cache per stiffness matrix (LinearSolveCacheEntry,
src/solver/boundary_conditions.jl:62-125).  Here requests of uneven size arrive
over time; each admitted request runs on its own HIP stream (one jfem context per
slot) so several requests' kernels overlap on the device:

* admission: a request waits for a free slot, then for free HBM covering its
  factor + matrix (exact sizes from its symbolic analysis) plus a reserve;
* residency: CSR pattern and symbolic factor are cached per topology (same grid
  points, connectivity and free set), numeric factors per stiffness matrix, all on
  the device; a request's right-hand sides share one factor;
* every result reports its relative residual, refinement iterations and timing.

Results do not depend on what else is running: every kernel is deterministic and
requests share no mutable device state, so a request solved alone and in the mix
gives identical bits.
"""
from __future__ import annotations

import hashlib
import heapq
import threading
import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from . import solver


@dataclass
class Request:
    rid: int
    model: object
    arrival: float = 0.0          # seconds after the stream starts
    precision: int = 32


@dataclass
class Result:
    rid: int
    u: np.ndarray | None
    relres: np.ndarray | None
    iters: np.ndarray | None
    t_arrival: float
    t_admit: float
    t_done: float
    ndof: int
    n_free: int
    nrhs: int
    ok: bool
    error: str = ""
    cache_hit: bool = False
    timings: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


def topology_key(model) -> str:
    h = hashlib.sha1()
    for a in (model.xyz, model.quads, model.trias, model.tets):
        h.update(np.ascontiguousarray(a).view(np.uint8))
    h.update(str(model.xyz.shape).encode())
    return h.hexdigest()


def matrix_key(model) -> str:
    h = hashlib.sha1(topology_key(model).encode())
    for a in (model.quad_props, model.tria_props, model.tet_props):
        h.update(np.ascontiguousarray(a).view(np.uint8))
    h.update(np.ascontiguousarray(model.spc, np.uint8).tobytes())
    h.update(repr(model.k6rot).encode())
    return h.hexdigest()


class SolveServer:
    """Serve static-solve requests on one GPU with ``slots`` concurrent streams."""

    def __init__(self, slots: int = 4, device: int = 0, precision: int = 32, tol: float = 1e-10,
                 reserve_bytes: int = 2 << 30, max_cached_factors: int = 8):
        self.ctxs = [solver.Context(device) for _ in range(slots)]
        self.precision = precision
        self.tol = tol
        self.reserve = reserve_bytes
        self.lock = threading.Lock()
        self.mem_cv = threading.Condition(self.lock)
        self.committed = 0            # bytes promised to admitted requests
        self.active = 0               # requests holding an admission
        self.peak_active = 0
        self.topo = {}                # topology key -> (DeviceMesh, Pattern, {fixed hash: Symbolic})
        self.factors = OrderedDict()  # matrix key -> (DeviceMesh, DeviceMatrix, Factor, fixed)
        self.max_cached_factors = max_cached_factors
        self.topo_locks = {}

    def close(self):
        for f in self.factors.values():
            f[2].free()
        self.factors.clear()
        for c in self.ctxs:
            c.close()

    # ---------------------------------------------------------------- one request
    def _topology(self, ctx, model):
        key = topology_key(model)
        with self.lock:
            lk = self.topo_locks.setdefault(key, threading.Lock())
        with lk:
            ent = self.topo.get(key)
            if ent is None:
                mesh = solver.DeviceMesh(model, ctx)
                pattern = solver.Pattern(mesh)
                ent = (mesh, pattern, {})
                with self.lock:
                    self.topo[key] = ent
            return ent

    def _admit_bytes(self, nbytes: int):
        with self.mem_cv:
            while True:
                free, total = self.ctxs[0].memory()
                if free - self.committed - self.reserve >= nbytes or self.active == 0:
                    self.committed += nbytes
                    self.active += 1
                    self.peak_active = max(self.peak_active, self.active)
                    return
                self.mem_cv.wait(timeout=0.05)

    def _release_bytes(self, nbytes: int):
        with self.mem_cv:
            self.committed -= nbytes
            self.active -= 1
            self.mem_cv.notify_all()

    def process(self, ctx, req: Request, t0: float) -> Result:
        t_admit = time.perf_counter() - t0
        m = req.model
        tim = {}
        try:
            ta = time.perf_counter()
            mkey = matrix_key(m)
            with self.lock:
                cached = self.factors.get(mkey)
                if cached is not None:
                    self.factors.move_to_end(mkey)
            if cached is not None:
                mesh_p, K, F, fixed = cached
                u, rr, it, ok = F.solve(K, m.loads, tol=self.tol, max_refine=100, ctx=ctx)
                tim["solve_s"] = time.perf_counter() - ta
                return Result(req.rid, u, rr, it, req.arrival, t_admit, time.perf_counter() - t0, m.ndof,
                              int((fixed == 0).sum()), m.nrhs, bool(ok), "" if ok else "not converged", True, tim)
            mesh, pattern, syms = self._topology(ctx, m)
            # per-request mesh with this request's sections (same topology)
            mesh_r = solver.DeviceMesh(m, ctx)
            K = solver.DeviceMatrix(mesh_r, pattern, ctx)
            fixed = solver.bc_partition(K, m.spc)
            fkey = hashlib.sha1(fixed.tobytes()).hexdigest()
            tb = time.perf_counter()
            with self.lock:
                S = syms.get(fkey)
            if S is None:
                S = solver.Symbolic(K, mesh, fixed, ctx)
                with self.lock:
                    syms.setdefault(fkey, S)
                    S = syms[fkey]
            tim["symbolic_s"] = time.perf_counter() - tb
            es = 4 if req.precision == 32 else 8
            need = (S.info["nnz_L"] + S.info["update_entries"]) * es
            self._admit_bytes(need)
            try:
                tc = time.perf_counter()
                F, used = solver.factor_with_fallback(S, K, req.precision, ctx)
                small = F.small_pivots(K)
                if len(small):
                    F.free()
                    fixed = fixed.copy()
                    fixed[small] = 1
                    S = solver.Symbolic(K, mesh, fixed, ctx)
                    F, used = solver.factor_with_fallback(S, K, req.precision, ctx)
                tim["factor_precision"] = used
                tim["factor_s"] = time.perf_counter() - tc
                ts = time.perf_counter()
                u, rr, it, ok = F.solve(K, m.loads, tol=self.tol, max_refine=100, ctx=ctx)
                tim["solve_s"] = time.perf_counter() - ts
                with self.lock:
                    self.factors[mkey] = (mesh_r, K, F, fixed)
                    while len(self.factors) > self.max_cached_factors:
                        _, old = self.factors.popitem(last=False)
                        old[2].free()
            finally:
                self._release_bytes(need)
            return Result(req.rid, u, rr, it, req.arrival, t_admit, time.perf_counter() - t0, m.ndof,
                          int((fixed == 0).sum()), m.nrhs, bool(ok), "" if ok else "not converged", False, tim)
        except Exception as e:  # a failing request does not stop the server
            return Result(req.rid, None, None, None, req.arrival, t_admit, time.perf_counter() - t0, m.ndof, 0,
                          m.nrhs, False, repr(e), False, tim)

    # ---------------------------------------------------------------- the loop
    def serve(self, requests: list[Request], realtime: bool = True) -> list[Result]:
        """Run the stream: requests become visible at their arrival times; each
        free slot takes the earliest visible request (continuous batching)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        heap = []
        results = []
        cv = threading.Condition()
        t0 = time.perf_counter()
        state = {"next": 0, "done": 0}

        def release_arrivals(now):
            while state["next"] < len(pending) and (not realtime or pending[state["next"]].arrival <= now):
                r = pending[state["next"]]
                heapq.heappush(heap, (r.arrival, r.rid, r))
                state["next"] += 1

        def worker(ctx):
            while True:
                with cv:
                    while True:
                        release_arrivals(time.perf_counter() - t0)
                        if heap:
                            _, _, req = heapq.heappop(heap)
                            break
                        if state["next"] >= len(pending):
                            return
                        nxt = pending[state["next"]].arrival
                        cv.wait(timeout=max(0.0, min(0.01, nxt - (time.perf_counter() - t0))))
                res = self.process(ctx, req, t0)
                with cv:
                    results.append(res)
                    state["done"] += 1
                    cv.notify_all()

        threads = [threading.Thread(target=worker, args=(c,), daemon=True) for c in self.ctxs]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        return sorted(results, key=lambda r: r.rid)


def summarize(results: list[Result]) -> dict:
    ok = [r for r in results if r.ok]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    span = max(r.t_done for r in results) - min(r.t_arrival for r in results) if results else 0.0
    nsolves = sum(r.nrhs for r in ok)
    return {"requests": len(results), "ok": len(ok), "solves": nsolves,
            "solves_per_s": nsolves / span if span > 0 else 0.0,
            "requests_per_s": len(ok) / span if span > 0 else 0.0,
            "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
            "max_relative_residual": float(max(np.max(r.relres) for r in ok)) if ok else None,
            "dof_solves_per_s": sum(r.n_free * r.nrhs for r in ok) / span if span > 0 else 0.0,
            "makespan_s": span}


def poisson_stream(n: int, rate_per_s: float, seed: int = 20260525, dof_range=(2e3, 2e5),
                   rhs_range=(1, 16), precision: int = 32):
    """The serving workload: n requests, Poisson arrivals, flat-plate or box shells
    with DOFs log-uniform in dof_range, 10 % CTRIA3, random thickness, 1-16 RHS."""
    from . import meshes
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate_per_s)
        dofs = float(np.exp(rng.uniform(np.log(dof_range[0]), np.log(dof_range[1]))))
        nrhs = int(rng.integers(rhs_range[0], rhs_range[1] + 1))
        m = meshes.shell_request(dofs, rng, nrhs, tria_fraction=0.1, idx=i)
        reqs.append(Request(i, m, t, precision))
    return reqs
