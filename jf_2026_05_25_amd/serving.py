"""Continuous-batching serving loop for static solves on one GPU.

The reference answers one request at a time from a persistent worker
(utils/julia_tools/jfem_worker_jsonl.jl:61, warm_model_session.jl:21-61) and keeps a
factor cache per stiffness matrix (LinearSolveCacheEntry,
src/solver/boundary_conditions.jl:62-125).  Here requests of uneven size arrive
over time; each admitted request runs on its own HIP stream (one jfem context per
slot) so several requests' kernels overlap on the device:

* admission, in two stages, each waiting for free HBM (less what earlier
  admissions promised, less a reserve): before anything is allocated, the
  request's mesh, CSR matrix and (for a new topology) pattern; after its
  symbolic analysis, its numeric factor and the growth of its slot's solve
  scratch.  Under memory pressure unused cached entries are evicted first; a
  stage lets a request through when no other request holds it, so the stream
  never stalls on an estimate;
* residency: CSR pattern and symbolic factors are cached per topology (same grid
  points and connectivity; symbolics per free set), numeric factors per
  (stiffness matrix, constraints, precision), all on the device, each cache LRU
  with a count and a byte bound.  Entries are reference-counted: a request holds
  what it uses, a cached factor holds its topology, and an evicted entry is freed
  by its last holder, never while a solve still runs on it;
* every result reports its relative residual, refinement iterations and timing.

Results do not depend on what else is running: every kernel is deterministic,
requests share no mutable device state, and the dataflow solves take their tasks
by ticket so that concurrent launches cannot starve each other; a request solved
alone and in the mix gives identical bits.
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
    """Stiffness matrix and constraints (sections, SPC, K6ROT on one topology)."""
    h = hashlib.sha1(topology_key(model).encode())
    for a in (model.quad_props, model.tria_props, model.tet_props):
        h.update(np.ascontiguousarray(a).view(np.uint8))
    h.update(np.ascontiguousarray(model.spc, np.uint8).tobytes())
    h.update(repr(model.k6rot).encode())
    return h.hexdigest()


def factor_key(model, precision: int) -> tuple:
    """A cached factor serves only requests of the same matrix AND precision (an
    FP64 request must not be answered from an FP32 factor)."""
    return matrix_key(model), int(precision)


class Resident:
    """Device objects held by the caches: ``users`` counts the holders (running
    requests, cached factors holding their topology); an entry evicted while held
    is freed by its last holder."""

    def __init__(self, key, objs: dict, nbytes: int, parent: "Resident | None" = None):
        self.key = key
        self.objs = objs
        self.nbytes = int(nbytes)
        self.users = 0
        self.evicted = False
        self.freed = False
        self.parent = parent


def estimate_pre_bytes(model, new_topology: bool) -> int:
    """Upper bound of what a request allocates before its symbolic analysis: its
    mesh, its CSR values and element-matrix buffer, and for a new topology the
    pattern (nnz <= number of element-matrix entries)."""
    ntrip = 576 * len(model.quads) + 324 * len(model.trias) + 900 * len(model.tets)
    mesh = 24 * model.n_nodes + 4 * (4 * len(model.quads) + 3 * len(model.trias) + 10 * len(model.tets)) + \
        8 * (8 * len(model.quads) + 8 * len(model.trias) + 2 * len(model.tets))
    K = 16 * ntrip
    pattern = (8 * (model.ndof + 1) + 12 * ntrip + 8 + 4 * ntrip + 288 * model.n_nodes) if new_topology else 0
    return mesh + K + pattern


class SolveServer:
    """Serve static-solve requests on one GPU with ``slots`` concurrent streams.

    ``backend`` supplies Context, DeviceMesh, Pattern, DeviceMatrix, Symbolic,
    bc_partition, factor_with_fallback and small_pivot_dofs (default: the device
    path in :mod:`solver`)."""

    def __init__(self, slots: int = 4, device: int = 0, precision: int = 32, tol: float = 1e-10,
                 reserve_bytes: int = 2 << 30, max_cached_factors: int = 8, max_cached_topologies: int = 16,
                 cache_bytes: int | None = None, backend=solver):
        self.B = backend
        self.ctxs = [backend.Context(device) for _ in range(slots)]
        self.precision = precision
        self.tol = tol
        self.reserve = reserve_bytes
        self.lock = threading.Lock()
        self.mem_cv = threading.Condition(self.lock)
        self.committed = 0            # bytes promised to admitted requests, not yet released
        self.stage_holders = [0, 0]   # requests holding a stage-1 / stage-2 admission
        self.active = 0               # requests holding any admission
        self.peak_active = 0
        self.topo = OrderedDict()     # topology key -> Resident(mesh, pattern, syms{fixed hash: Symbolic})
        self.factors = OrderedDict()  # factor key -> Resident(mesh, K, F, S, fixed)
        self.max_cached_factors = max_cached_factors
        self.max_cached_topologies = max_cached_topologies
        if cache_bytes is None:
            _, total = self.ctxs[0].memory()
            cache_bytes = total // 2
        self.cache_bytes = cache_bytes
        self.topo_locks = {}
        self.evictions = 0

    def clear_caches(self):
        """Evict every cached factor and topology (entries in use are freed by
        their last holder)."""
        with self.lock:
            for d in (self.factors, self.topo):
                for k in list(d):
                    self._evict(d, k)

    def close(self):
        with self.lock:
            for d in (self.factors, self.topo):
                for ent in list(d.values()):
                    ent.evicted = True
                    if ent.users == 0:
                        self._free(ent)
                d.clear()
        for c in self.ctxs:
            c.close()

    # ------------------------------------------------------------ residency (under self.lock)
    def _free(self, ent: Resident):
        if ent.freed:
            return
        ent.freed = True
        o = ent.objs
        if "syms" in o:              # topology: symbolics, then pattern and mesh
            for S in o["syms"].values():
                S.free()
            o["pattern"].free()
            o["mesh"].free()
        else:                        # factor: factor, its own symbolic, matrix, mesh
            for name in ("F", "S_own", "K", "mesh"):
                if o.get(name) is not None:
                    o[name].free()
        if ent.parent is not None:
            self._release(ent.parent)

    def _hold(self, ent: Resident):
        ent.users += 1

    def _release(self, ent: Resident):
        ent.users -= 1
        if ent.users == 0 and ent.evicted:
            self._free(ent)

    def _evict(self, d: OrderedDict, key):
        ent = d.pop(key)
        ent.evicted = True
        self.evictions += 1
        if ent.users == 0:
            self._free(ent)

    def _cached_bytes(self) -> int:
        return sum(e.nbytes for e in self.factors.values()) + sum(e.nbytes for e in self.topo.values())

    def _trim(self, need_free: int | None = None):
        """Evict least-recently-used entries beyond the count and byte bounds
        (and, with ``need_free``, until that many bytes are free); entries in use
        are skipped by the memory-pressure pass."""
        while len(self.factors) > self.max_cached_factors:
            self._evict(self.factors, next(iter(self.factors)))
        while len(self.topo) > self.max_cached_topologies:
            idle = [k for k, e in self.topo.items() if e.users == 0]
            if not idle:
                break
            self._evict(self.topo, idle[0])
        while self._cached_bytes() > self.cache_bytes:
            if not self._evict_one_idle():
                break
        if need_free is not None:
            while self._available() < need_free:
                if not self._evict_one_idle():
                    break

    def _evict_one_idle(self) -> bool:
        for d in (self.factors, self.topo):
            for k, e in d.items():
                if e.users == 0:
                    self._evict(d, k)
                    return True
        return False

    def _available(self) -> int:
        free, _ = self.ctxs[0].memory()
        return free - self.committed - self.reserve

    # ------------------------------------------------------------ admission
    def _admit(self, stage: int, nbytes: int):
        with self.mem_cv:
            while True:
                if self._available() < nbytes:
                    self._trim(need_free=nbytes)
                if self._available() >= nbytes or self.stage_holders[stage] == 0:
                    self.committed += nbytes
                    self.stage_holders[stage] += 1
                    if stage == 0:
                        self.active += 1
                        self.peak_active = max(self.peak_active, self.active)
                    return
                self.mem_cv.wait(timeout=0.05)

    def _leave(self, stage: int, nbytes: int):
        with self.mem_cv:
            self.committed -= nbytes
            self.stage_holders[stage] -= 1
            if stage == 0:
                self.active -= 1
            self.mem_cv.notify_all()

    # ------------------------------------------------------------ one request
    def _topology(self, ctx, model) -> Resident:
        """The topology's cached mesh / pattern / symbolics, held for the caller."""
        key = topology_key(model)
        with self.lock:
            lk = self.topo_locks.setdefault(key, threading.Lock())
        with lk:
            with self.lock:
                ent = self.topo.get(key)
                if ent is not None:
                    self.topo.move_to_end(key)
                    self._hold(ent)
                    return ent
            mesh = self.B.DeviceMesh(model, ctx)
            pattern = self.B.Pattern(mesh)
            ent = Resident(key, {"mesh": mesh, "pattern": pattern, "syms": {}},
                           mesh.device_bytes() + pattern.device_bytes())
            with self.lock:
                self.topo[key] = ent
                self._hold(ent)
                self._trim()
            return ent

    def _symbolic(self, topo: Resident, K, fixed, ctx):
        fkey = hashlib.sha1(np.ascontiguousarray(fixed).tobytes()).hexdigest()
        syms = topo.objs["syms"]
        with self.lock:
            S = syms.get(fkey)
        if S is None:
            S_new = self.B.Symbolic(K, topo.objs["mesh"], fixed, ctx)
            with self.lock:
                S = syms.setdefault(fkey, S_new)
                if S is S_new:
                    topo.nbytes += S.device_bytes()
            if S is not S_new:
                S_new.free()
        return S

    def process(self, ctx, req: Request, t0: float) -> Result:
        t_admit = time.perf_counter() - t0
        m = req.model
        tim = {}
        try:
            ta = time.perf_counter()
            fkey = factor_key(m, req.precision)
            with self.lock:
                cached = self.factors.get(fkey)
                if cached is not None:
                    self.factors.move_to_end(fkey)
                    self._hold(cached)
            if cached is not None:
                try:
                    o = cached.objs
                    u, rr, it, ok = o["F"].solve(o["K"], m.loads, tol=self.tol, max_refine=100, ctx=ctx)
                finally:
                    with self.lock:
                        self._release(cached)
                tim["solve_s"] = time.perf_counter() - ta
                return Result(req.rid, u, rr, it, req.arrival, t_admit, time.perf_counter() - t0, m.ndof,
                              int((o["fixed"] == 0).sum()), m.nrhs, bool(ok), "" if ok else "not converged", True,
                              tim)
            with self.lock:
                new_topo = topology_key(m) not in self.topo
            pre = estimate_pre_bytes(m, new_topo)
            self._admit(0, pre)
            try:
                return self._process_new(ctx, req, t0, t_admit, fkey, tim)
            finally:
                self._leave(0, pre)
        except Exception as e:  # a failing request does not stop the server
            return Result(req.rid, None, None, None, req.arrival, t_admit, time.perf_counter() - t0, m.ndof, 0,
                          m.nrhs, False, repr(e), False, tim)

    def _process_new(self, ctx, req, t0, t_admit, fkey, tim) -> Result:
        m = req.model
        t0_ = time.perf_counter()
        topo = self._topology(ctx, m)
        tim["topology_s"] = time.perf_counter() - t0_
        keep = False
        mesh_r = K = F = S_own = None
        try:
            # per-request mesh with this request's sections (same topology)
            ta = time.perf_counter()
            mesh_r = self.B.DeviceMesh(m, ctx)
            K = self.B.DeviceMatrix(mesh_r, topo.objs["pattern"], ctx)
            tim["assemble_s"] = time.perf_counter() - ta
            ta = time.perf_counter()
            fixed = self.B.bc_partition(K, m.spc)
            tim["bc_s"] = time.perf_counter() - ta
            tb = time.perf_counter()
            S = self._symbolic(topo, K, fixed, ctx)
            tim["symbolic_s"] = time.perf_counter() - tb
            need = S.factor_bytes(req.precision) + max(0, S.solve_scratch_bytes(m.nrhs, req.precision) -
                                                      ctx.device_bytes())
            self._admit(1, need)
            try:
                tc = time.perf_counter()
                F, used = self.B.factor_with_fallback(S, K, req.precision, ctx)
                small, _ = self.B.small_pivot_dofs(S, K, F, ctx=ctx)
                if len(small):   # post-factor pivot check (src/solver/boundary_conditions.jl:1684-1710)
                    F.free()
                    F = None
                    fixed = fixed.copy()
                    fixed[small] = 1
                    S_own = S = self.B.Symbolic(K, topo.objs["mesh"], fixed, ctx)
                    F, used = self.B.factor_with_fallback(S, K, req.precision, ctx)
                tim["factor_precision"] = used
                tim["factor_s"] = time.perf_counter() - tc
                ts = time.perf_counter()
                u, rr, it, ok = F.solve(K, m.loads, tol=self.tol, max_refine=100, ctx=ctx)
                tim["solve_s"] = time.perf_counter() - ts
            finally:
                self._leave(1, need)
            nbytes = mesh_r.device_bytes() + K.device_bytes() + F.device_bytes() + \
                (S_own.device_bytes() if S_own is not None else 0)
            ent = Resident(fkey, {"mesh": mesh_r, "K": K, "F": F, "S_own": S_own, "fixed": fixed}, nbytes,
                           parent=topo)
            with self.lock:
                old = self.factors.pop(fkey, None)
                if old is not None:          # another slot cached the same matrix meanwhile
                    old.evicted = True
                    if old.users == 0:
                        self._free(old)
                self.factors[fkey] = ent     # holds the topology (released when freed)
                self._trim()
            keep = True
            return Result(req.rid, u, rr, it, req.arrival, t_admit, time.perf_counter() - t0, m.ndof,
                          int((fixed == 0).sum()), m.nrhs, bool(ok), "" if ok else "not converged", False, tim)
        finally:
            if not keep:
                for o in (F, S_own, K, mesh_r):
                    if o is not None:
                        o.free()
                with self.lock:
                    self._release(topo)

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


def stage_times(results: list[Result]) -> dict:
    """Mean seconds per request of each timed stage (requests that ran it)."""
    keys = ("topology_s", "assemble_s", "bc_s", "symbolic_s", "factor_s", "solve_s")
    out = {}
    for k in keys:
        v = [r.timings[k] for r in results if k in r.timings]
        if v:
            out[k] = float(np.mean(v))
    return out


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
