"""Host side of the drop-in: the reference's assemble / factor / solve calls,
re-expressed over the device C ABI (include/jfem_hip.h).

Mirrors, for this path, the reference's
  * FEM.stiffness_quad4_matrices / stiffness_tria3_matrices / stiffness_tetra10
    (src/FEMKernels.jl:2839, 9811; src/tetra10.jl:68),
  * Solver.assemble_stiffness (src/solver/assembly.jl:2630),
  * Solver.apply_bc_and_solve (src/solver/boundary_conditions.jl:1324) with its
    linear-solve cache (:62-125) and post-factor pivot check (:1662-1710).
Everything runs on the GPU through libjfem_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, cm, ptr

FP64 = 64
FP32 = 32
KIND = {"CQUAD4": 0, "CTRIA3": 1, "CTETRA10": 2}


class Context:
    """One HIP stream on one device (one per host thread / serving slot)."""

    def __init__(self, device: int = 0):
        self.L = _lib.load()
        h = C.c_void_p()
        code = self.L.jfem_ctx_create(device, C.byref(h))
        if code != 0:
            raise _lib.JfemError(code, "jfem_ctx_create: no usable HIP device")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.L.jfem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.h, self.L.jfem_ctx_synchronize(self.h))

    def memory(self):
        f = C.c_int64(); t = C.c_int64()
        check(self.h, self.L.jfem_device_memory(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def device_bytes(self) -> int:
        """Solve scratch and BC buffers this context holds on the device."""
        return int(self.L.jfem_ctx_device_bytes(self.h))

    def enable_timing(self, on=True):
        check(self.h, self.L.jfem_ctx_enable_timing(self.h, int(on)))

    def reset_timing(self):
        check(self.h, self.L.jfem_ctx_reset_kernel_time(self.h))

    def kernel_time(self, family: int):
        ms = C.c_double(); n = C.c_int64()
        check(self.h, self.L.jfem_ctx_kernel_time(self.h, family, C.byref(ms), C.byref(n)))
        return ms.value, n.value


_default = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(0)
        return _default


# --------------------------------------------------------------------------- elements
def stiffness_quad4_matrices(coords, Cm, Cb, Cs, h, E_ref=None, *, k6rot=100.0, drill_scale=1.0,
                             rigid_shear=False, membrane_modes=15, ctx: Context | None = None):
    """Local-frame CQUAD4 stiffness (src/FEMKernels.jl:2839), batched.

    ``coords``: (4,2) or (n,4,2); ``Cm``/``Cb``: (3,3) or (n,3,3); ``Cs``: (2,2) or
    (n,2,2); ``h``, ``k6rot``, ``drill_scale``: scalar or (n,).  Returns (24,24) or
    (n,24,24).  Raises JfemError(DEGENERATE) like the reference's ArgumentError."""
    ctx = ctx or default_context()
    coords = np.asarray(coords, float)
    single = coords.ndim == 2
    coords = coords.reshape(-1, 4, 2)
    n = coords.shape[0]

    def per(a, shape):
        a = np.asarray(a, float)
        a = np.broadcast_to(a, (n,) + shape) if a.shape == shape else a.reshape((n,) + shape)
        return np.ascontiguousarray(np.transpose(a, (0,) + tuple(range(len(shape), 0, -1))).reshape(n, -1))

    lc = np.ascontiguousarray(coords.reshape(n, 8))
    Cm_ = per(Cm, (3, 3)); Cb_ = per(Cb, (3, 3)); Cs_ = per(Cs, (2, 2))
    h_ = np.ascontiguousarray(np.broadcast_to(np.asarray(h, float), (n,)))
    k6 = np.ascontiguousarray(np.broadcast_to(np.asarray(k6rot, float), (n,)))
    ds = np.ascontiguousarray(np.broadcast_to(np.asarray(drill_scale, float), (n,)))
    flags = (np.broadcast_to(np.asarray(rigid_shear, np.int32), (n,)).astype(np.int32) |
             ((np.broadcast_to(np.asarray(membrane_modes, np.int32), (n,)) & 15) << 4)).astype(np.int32)
    flags = np.ascontiguousarray(flags)
    Ke = np.zeros(n * 576)
    bad = C.c_int64(-1)
    check(ctx.h, ctx.L.jfem_q4_local_batch(ctx.h, n, ptr(lc), ptr(Cm_), ptr(Cb_), ptr(Cs_), ptr(h_), ptr(k6),
                                           ptr(ds), ptr(flags), ptr(Ke), C.byref(bad)))
    Ke = Ke.reshape(n, 24, 24).transpose(0, 2, 1)   # column-major blocks -> (n, row, col)
    return Ke[0] if single else Ke


def _per_element(a, n, shape):
    """(n,) + shape array (or one shape broadcast to n) as n column-major blocks."""
    a = np.asarray(a, float)
    a = np.broadcast_to(a, (n,) + shape) if a.shape == shape else a.reshape((n,) + shape)
    return np.ascontiguousarray(np.transpose(a, (0,) + tuple(range(len(shape), 0, -1))).reshape(n, -1))


def stiffness_tria3_matrices(coords, Dm, Db, Ds, h=1.0, G_ref=1.0, *, bend_ratio=1.0, k6rot=100.0,
                             ctx: Context | None = None):
    """Element-frame CTRIA3 stiffness (src/FEMKernels.jl:9811), batched.

    ``coords``: (3,2) or (n,3,2); ``Dm``/``Db``: (3,3) or (n,3,3); ``Ds``: (2,2) or
    (n,2,2); ``h``, ``G_ref``, ``bend_ratio``, ``k6rot``: scalar or (n,).  Returns
    (18,18) or (n,18,18)."""
    ctx = ctx or default_context()
    coords = np.asarray(coords, float)
    single = coords.ndim == 2
    coords = coords.reshape(-1, 3, 2)
    n = coords.shape[0]
    lc = np.ascontiguousarray(coords.reshape(n, 6))
    Dm_, Db_, Ds_ = _per_element(Dm, n, (3, 3)), _per_element(Db, n, (3, 3)), _per_element(Ds, n, (2, 2))
    vec = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, float), (n,))) for v in (h, G_ref, bend_ratio, k6rot)]
    Ke = np.zeros(n * 324)
    bad = C.c_int64(-1)
    check(ctx.h, ctx.L.jfem_t3_local_batch(ctx.h, n, ptr(lc), ptr(Dm_), ptr(Db_), ptr(Ds_), *[ptr(v) for v in vec],
                                           ptr(Ke), C.byref(bad)))
    Ke = Ke.reshape(n, 18, 18).transpose(0, 2, 1)
    return Ke[0] if single else Ke


def stiffness_tetra10(coords, E, nu, ctx: Context | None = None):
    """CTETRA10 stiffness (src/tetra10.jl:68), batched: ``coords`` (10,3) or
    (n,10,3), ``E``/``nu`` scalar or (n,).  Returns (30,30) or (n,30,30).  Raises
    JfemError(INVERTED) on a non-positive Jacobian like the reference's
    DomainError, JfemError(BAD_ARG) on bad E / nu."""
    ctx = ctx or default_context()
    X = np.asarray(coords, float)
    single = X.ndim == 2
    X = X.reshape(-1, 10, 3)
    n = X.shape[0]
    Xc = _per_element(X, n, (10, 3))
    E_ = np.ascontiguousarray(np.broadcast_to(np.asarray(E, float), (n,)))
    nu_ = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, float), (n,)))
    Ke = np.zeros(n * 900)
    bad = C.c_int64(-1)
    check(ctx.h, ctx.L.jfem_tet10_batch(ctx.h, n, ptr(Xc), ptr(E_), ptr(nu_), ptr(Ke), C.byref(bad)))
    Ke = Ke.reshape(n, 30, 30).transpose(0, 2, 1)
    return Ke[0] if single else Ke


class DeviceMesh:
    """A model's grid points, connectivity and sections resident on the device."""

    def __init__(self, model, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.model = model
        L = self.ctx.L
        x = np.ascontiguousarray(model.xyz[:, 0]); y = np.ascontiguousarray(model.xyz[:, 1])
        z = np.ascontiguousarray(model.xyz[:, 2])
        self._keep = [cm(model.quads, np.int32), cm(model.quad_props, float), cm(model.trias, np.int32),
                      cm(model.tria_props, float), cm(model.tets, np.int32), cm(model.tet_props, float)]
        qc, qp, tc, tp, ec, ep = self._keep
        h = C.c_void_p()
        check(self.ctx.h, L.jfem_mesh_create(self.ctx.h, model.n_nodes, ptr(x), ptr(y), ptr(z),
                                             len(model.quads), ptr(qc), ptr(qp), len(model.trias), ptr(tc), ptr(tp),
                                             len(model.tets), ptr(ec), ptr(ep), float(model.k6rot), C.byref(h)))
        self.h = h

    def update_props(self, model):
        qp, tp, ep = cm(model.quad_props, float), cm(model.tria_props, float), cm(model.tet_props, float)
        check(self.ctx.h, self.ctx.L.jfem_mesh_update_props(self.ctx.h, self.h, ptr(qp), ptr(tp), ptr(ep)))
        self.model = model

    def device_bytes(self) -> int:
        return int(self.ctx.L.jfem_mesh_device_bytes(self.h))

    def element_matrices(self, kind: str):
        """Global-frame element matrices exactly as assembled, (n, nd, nd)."""
        m = self.model
        n, nd = {"CQUAD4": (len(m.quads), 24), "CTRIA3": (len(m.trias), 18), "CTETRA10": (len(m.tets), 30)}[kind]
        out = np.zeros(n * nd * nd)
        bad = C.c_int64(-1)
        check(self.ctx.h, self.ctx.L.jfem_element_matrices(self.ctx.h, self.h, KIND[kind], ptr(out), C.byref(bad)))
        return out.reshape(n, nd, nd).transpose(0, 2, 1)

    def shell_recover(self, u, mean_shear: bool = True):
        """Shell force / stress / strain recovery from one displacement vector u
        (ndof, global DOF order) on the device: raw CQUAD4 rows (n_quad4, 44) and
        CTRIA3 rows (n_tria3, 20), layout of jfem_shell_recover; mean_shear selects
        the branch of solution sequences 101/105/106 (recover_shell_stresses!,
        src/solver/stress_recovery.jl:505)."""
        m = self.model
        u = np.ascontiguousarray(np.asarray(u, dtype=float).reshape(-1))
        if u.size != m.ndof:
            raise ValueError(f"u has {u.size} entries, the model {m.ndof} DOFs")
        q = np.zeros((len(m.quads), 44))
        t = np.zeros((len(m.trias), 20))
        bad = C.c_int64(-1)
        check(self.ctx.h, self.ctx.L.jfem_shell_recover(self.ctx.h, self.h, ptr(u), int(bool(mean_shear)),
                                                        ptr(q), ptr(t), C.byref(bad)))
        return q, t

    def free(self):
        if self.h:
            self.ctx.L.jfem_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Pattern:
    """CSR pattern of K plus the element-slot gather lists (topology only)."""

    def __init__(self, mesh: DeviceMesh):
        self.ctx = mesh.ctx
        h = C.c_void_p(); nnz = C.c_int64()
        check(self.ctx.h, self.ctx.L.jfem_pattern_create(self.ctx.h, mesh.h, C.byref(h), C.byref(nnz)))
        self.h = h
        self.nnz = nnz.value
        self.ndof = mesh.model.ndof

    def device_bytes(self) -> int:
        return int(self.ctx.L.jfem_pattern_device_bytes(self.h))

    def free(self):
        if self.h:
            self.ctx.L.jfem_pattern_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMatrix:
    """Assembled global K (CSR values, FP64) resident on the device."""

    def __init__(self, mesh: DeviceMesh, pattern: Pattern, ctx: Context | None = None):
        self.ctx = ctx or mesh.ctx
        self.pattern = pattern
        h = C.c_void_p(); bad = C.c_int64(-1)
        code = self.ctx.L.jfem_assemble(self.ctx.h, mesh.h, pattern.h, C.byref(h), C.byref(bad))
        self.bad_element = bad.value
        check(self.ctx.h, code)
        self.h = h

    def device_bytes(self) -> int:
        return int(self.ctx.L.jfem_matrix_device_bytes(self.h))

    def reassemble(self, mesh: DeviceMesh):
        bad = C.c_int64(-1)
        check(self.ctx.h, self.ctx.L.jfem_reassemble(self.ctx.h, mesh.h, self.h, C.byref(bad)))

    def to_scipy(self):
        import scipy.sparse as sp
        n = self.pattern.ndof
        rp = np.zeros(n + 1, np.int64); col = np.zeros(self.pattern.nnz, np.int32); val = np.zeros(self.pattern.nnz)
        check(self.ctx.h, self.ctx.L.jfem_matrix_download(self.ctx.h, self.h, ptr(rp), ptr(col), ptr(val)))
        return sp.csr_matrix((val, col, rp), shape=(n, n))

    def free(self):
        if self.h:
            self.ctx.L.jfem_matrix_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class AssembledStiffness:
    """assemble_stiffness result (src/solver/assembly.jl:5949), device-resident K."""
    K: DeviceMatrix
    mesh: DeviceMesh
    pattern: Pattern
    ndof: int
    node_coords: np.ndarray
    timings: dict = field(default_factory=dict)


def assemble_stiffness(model, ctx: Context | None = None, mesh: DeviceMesh | None = None,
                       pattern: Pattern | None = None) -> AssembledStiffness:
    """Element generation + deterministic assembly of the global K on the GPU."""
    ctx = ctx or default_context()
    t0 = time.perf_counter()
    mesh = mesh or DeviceMesh(model, ctx)
    t1 = time.perf_counter()
    pattern = pattern or Pattern(mesh)
    t2 = time.perf_counter()
    K = DeviceMatrix(mesh, pattern, ctx)
    t3 = time.perf_counter()
    return AssembledStiffness(K, mesh, pattern, model.ndof, model.xyz,
                              {"upload_s": t1 - t0, "pattern_s": t2 - t1, "assemble_s": t3 - t2})


def bc_partition(K: DeviceMatrix, spc: np.ndarray, eps_zero: float = 1e-8):
    """Fixed-DOF mask: SPC + AUTOSPC + grid-point singularity (src/solver/boundary_conditions.jl:1480-1607)."""
    ctx = K.ctx
    spc = np.ascontiguousarray(np.asarray(spc, np.uint8))
    fixed = np.zeros(K.pattern.ndof, np.uint8)
    nf = C.c_int64()
    check(ctx.h, ctx.L.jfem_bc_partition(ctx.h, K.h, ptr(spc), eps_zero, ptr(fixed), C.byref(nf)))
    return fixed


# --------------------------------------------------------------------- load vectors
def single_precision(a):
    """FORCE / MOMENT magnitudes and directions as the reference's reader keeps
    them: rounded to single precision when finite and in range
    (load_single_precision, src/parsing/extract_loads.jl:15-22)."""
    a = np.asarray(a, dtype=float)
    ok = np.isfinite(a) & (np.abs(a) <= np.finfo(np.float32).max)
    with np.errstate(over="ignore"):
        return np.where(ok, a.astype(np.float32).astype(float), a)


class LoadSet:
    """One load SID on the device: FORCE / MOMENT cards on grid points, uniform
    PLOAD4 cards on shell elements and GRAV cards on shell mass, in card order
    (the FORCE, MOMENT, shell PLOAD4 and shell GRAV branches of _resolve_loads!,
    src/solver/loads.jl:517-587, 680-772, 1338-1340).

    pload4 = (elem, p[, dir3]): shell element e is CQUAD4 e (e < n_quad) or CTRIA3
    e - n_quad; a dir3 row of norm <= 1e-30 (or no dir3) takes the element normal.
    force / moment = (node, mag, dir3), values as given (see single_precision)."""

    def __init__(self, mesh: DeviceMesh, pload4=None, force=None, moment=None, ctx: Context | None = None,
                 grav=None, shell_areal=None):
        self.ctx = ctx or mesh.ctx
        self.mesh = mesh
        L = self.ctx.L

        def cards(c, nd):
            if c is None or len(c[0]) == 0:
                return 0, None, None, None
            idx = np.ascontiguousarray(np.asarray(c[0], np.int32))
            val = np.ascontiguousarray(np.asarray(c[1], float).reshape(len(idx)))
            d = None
            if len(c) > 2 and c[2] is not None:
                d = np.ascontiguousarray(np.asarray(c[2], float).reshape(len(idx), 3))
            elif nd:
                raise ValueError("FORCE / MOMENT cards need directions")
            return len(idx), idx, val, d

        npl, pe, pp, pd = cards(pload4, False)
        nf, fn, fm, fd = cards(force, True)
        nm, mn, mm, md = cards(moment, True)
        self._keep = [pe, pp, pd, fn, fm, fd, mn, mm, md]
        h = C.c_void_p()
        check(self.ctx.h, L.jfem_loadset_create(self.ctx.h, mesh.h, npl, ptr(pe), ptr(pp), ptr(pd),
                                                nf, ptr(fn), ptr(fm), ptr(fd), nm, ptr(mn), ptr(mm), ptr(md),
                                                C.byref(h)))
        self.h = h
        self.ndof = 6 * mesh.model.n_nodes
        if grav is not None and len(grav[0]):
            self.set_gravity(grav, shell_areal)

    def set_gravity(self, grav, shell_areal):
        """GRAV cards on the shells: grav = (A (n,), N (n, 3)) in the basic frame,
        shell_areal = rho T + NSM per shell (CQUAD4s then CTRIA3s, 0 without mass):
        the lumped shell branch of the GRAV loads (src/solver/loads.jl:680-772)."""
        m = self.mesh.model
        a = np.ascontiguousarray(np.asarray(grav[0], float).reshape(-1))
        n3 = np.ascontiguousarray(np.asarray(grav[1], float).reshape(len(a), 3))
        ar = np.ascontiguousarray(np.asarray(shell_areal, float).reshape(len(m.quads) + len(m.trias)))
        self._gkeep = [a, n3, ar]
        check(self.ctx.h, self.ctx.L.jfem_loadset_set_gravity(self.ctx.h, self.h, len(a), ptr(a), ptr(n3), ptr(ar)))

    def vector(self, scale: float = 1.0) -> np.ndarray:
        """The load vector (all 6 n DOFs) with the SID's scale factor."""
        F = np.zeros(self.ndof)
        check(self.ctx.h, self.ctx.L.jfem_load_vector(self.ctx.h, self.h, float(scale), ptr(F)))
        return F

    def vector_dev(self, F_dev, scale: float = 1.0):
        """Write the load vector into device memory F_dev (6 n doubles), on the
        context's stream."""
        dst = F_dev if isinstance(F_dev, C.c_void_p) else C.c_void_p(F_dev)
        check(self.ctx.h, self.ctx.L.jfem_load_vector_dev(self.ctx.h, self.h, float(scale), dst))

    def free(self):
        if self.h:
            self.ctx.L.jfem_loadset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# --------------------------------------------------------------------- factor / solve
class Symbolic:
    def __init__(self, K: DeviceMatrix, mesh: DeviceMesh, fixed: np.ndarray, ctx: Context | None = None):
        self.ctx = ctx or K.ctx
        self.fixed = np.ascontiguousarray(np.asarray(fixed, np.uint8))
        h = C.c_void_p()
        info = _lib.SymbolicInfo()
        check(self.ctx.h, self.ctx.L.jfem_symbolic_create(self.ctx.h, K.pattern.h, mesh.h, ptr(self.fixed),
                                                          C.byref(h), C.byref(info)))
        self.h = h
        self.info = {k: getattr(info, k) for k, _ in _lib.SymbolicInfo._fields_}
        self.ndof = K.pattern.ndof

    def device_bytes(self) -> int:
        return int(self.ctx.L.jfem_symbolic_device_bytes(self.h))

    def order(self) -> np.ndarray:
        """Elimination order: the global DOF eliminated at each position."""
        out = np.zeros(max(self.info["n_free"], 1), np.int64)
        check(self.ctx.h, self.ctx.L.jfem_symbolic_order(self.h, ptr(out)))
        return out[:self.info["n_free"]]

    def factor_bytes(self, precision: int) -> int:
        """Device bytes of a numeric factor of this symbolic (L, update matrices,
        inverse diagonal blocks)."""
        es = 8 if precision == FP64 else 4
        return (self.info["nnz_L"] + self.info["update_entries"] + self.info["inv_entries"]) * es

    def solve_scratch_bytes(self, nrhs: int, precision: int) -> int:
        """Context scratch one solve with ``nrhs`` right-hand sides needs."""
        return int(self.ctx.L.jfem_solve_scratch_bytes(self.h, self.ndof, int(nrhs), int(precision)))

    def free(self):
        if self.h:
            self.ctx.L.jfem_symbolic_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Factor:
    def __init__(self, S: Symbolic, K: DeviceMatrix, precision: int = FP64, ctx: Context | None = None):
        self.ctx = ctx or K.ctx
        self.symbolic = S
        self.precision = precision
        h = C.c_void_p(); bad = C.c_int64(-1)
        code = self.ctx.L.jfem_factor_create(self.ctx.h, S.h, K.h, int(precision), C.byref(h), C.byref(bad))
        self.bad_column = bad.value
        check(self.ctx.h, code)
        self.h = h

    def refactor(self, K: DeviceMatrix):
        bad = C.c_int64(-1)
        code = self.ctx.L.jfem_refactor(self.ctx.h, self.h, K.h, C.byref(bad))
        self.bad_column = bad.value
        check(self.ctx.h, code)

    def small_pivots(self, K: DeviceMatrix, threshold: float = 1e-7):
        cap = self.symbolic.info["n_free"]
        out = np.zeros(max(cap, 1), np.int64); cnt = C.c_int64()
        check(self.ctx.h, self.ctx.L.jfem_factor_small_pivots(self.ctx.h, self.h, K.h, threshold, ptr(out), cap,
                                                              C.byref(cnt)))
        return out[:cnt.value].copy()

    def pivot_ratios(self, K: DeviceMatrix) -> np.ndarray:
        """L_kk^2 / |K_gg| per elimination position (see Symbolic.order)."""
        out = np.zeros(max(self.symbolic.info["n_free"], 1))
        check(self.ctx.h, self.ctx.L.jfem_factor_pivot_ratios(self.ctx.h, self.h, K.h, ptr(out)))
        return out[:self.symbolic.info["n_free"]]

    def device_bytes(self):
        return int(self.ctx.L.jfem_factor_device_bytes(self.h))

    def solve(self, K: DeviceMatrix, B: np.ndarray, tol: float = 1e-10, max_refine: int = 100,
              ctx: Context | None = None, details: bool = False):
        """Solve K_ff X = B_ff with refinement on ``ctx``'s stream (default: the
        factor's own context).  Solve scratch lives in the context, so concurrent
        solves with one factor must use different contexts.

        Returns ``(X, relres, iters, ok)``; with ``details`` also a dict holding
        ``x_lo`` (low parts of the double-double solution) and ``relres_fp64``
        (residual of the returned double X itself)."""
        ctx = ctx or self.ctx
        B = np.asarray(B, float)
        one = B.ndim == 1
        if one:
            B = B[:, None]
        nrhs = B.shape[1]
        Bc = np.ascontiguousarray(B.T)        # column-major ndof x nrhs
        X = np.zeros_like(Bc)
        Xl = np.zeros_like(Bc) if details else None
        iters = np.zeros(nrhs, np.int32); rr = np.zeros(nrhs); rr64 = np.zeros(nrhs)
        code = ctx.L.jfem_solve(ctx.h, self.h, K.h, nrhs, ptr(Bc), ptr(X), tol, max_refine, ptr(iters), ptr(rr),
                                ptr(Xl) if details else None, ptr(rr64))
        if code not in (0, 8):
            check(ctx.h, code)
        X = X.T.copy()
        out = ((X[:, 0] if one else X), rr, iters, code == 0)
        if details:
            Xl = Xl.T.copy()
            return out + ({"x_lo": Xl[:, 0] if one else Xl, "relres_fp64": rr64},)
        return out

    def solve_device(self, K: DeviceMatrix, b_dev: "DeviceBuffer", x_dev: "DeviceBuffer", nrhs: int,
                     tol: float = 1e-10, max_refine: int = 100, ctx: Context | None = None,
                     static_scratch: bool = False):
        """As ``solve`` with right-hand sides and solution in device memory
        (ndof x nrhs column-major FP64, ``DeviceBuffer``): nothing crosses PCIe but
        the per-RHS residuals.  ``static_scratch`` selects jfem_solve_dev_static,
        which allocates nothing (its scratch is the library's static device array;
        calls that overlap on one device race).  Returns ``(relres, iters, ok)``."""
        ctx = ctx or self.ctx
        iters = np.zeros(nrhs, np.int32); rr = np.zeros(nrhs)
        fn = ctx.L.jfem_solve_dev_static if static_scratch else ctx.L.jfem_solve_dev
        code = fn(ctx.h, self.h, K.h, nrhs, b_dev.ptr, x_dev.ptr, tol, max_refine, ptr(iters), ptr(rr), None, None)
        if code not in (0, 8):
            check(ctx.h, code)
        return rr, iters, code == 0

    def free(self):
        if self.h:
            self.ctx.L.jfem_factor_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBatch:
    """Design variants of one topology on the device: the variants' section
    records (each model must share the mesh's grid points and connectivity), their
    element matrices and K values, and after ``factor`` the numeric factors of all
    of them over one symbolic.  Element generation + assembly and the numeric
    factorization each run as one launch sequence over every variant.

    ``matrix(v)`` / ``factor_view(v)`` give variant v's DeviceMatrix / Factor,
    owned by the batch (valid until ``free``)."""

    def __init__(self, mesh: DeviceMesh, pattern: Pattern, models, ctx: Context | None = None):
        self.ctx = ctx or mesh.ctx
        self.mesh, self.pattern = mesh, pattern
        self.V = len(models)
        qp, tp, ep = self._props(models)
        h = C.c_void_p()
        check(self.ctx.h, self.ctx.L.jfem_batch_create(self.ctx.h, mesh.h, pattern.h, self.V, ptr(qp), ptr(tp),
                                                       ptr(ep), C.byref(h)))
        self.h = h
        self._mviews = []
        for v in range(self.V):
            vh = C.c_void_p()
            check(self.ctx.h, self.ctx.L.jfem_batch_matrix(self.h, v, C.byref(vh)))
            K = object.__new__(DeviceMatrix)
            K.ctx, K.pattern, K.h, K.bad_element = self.ctx, pattern, vh, -1
            self._mviews.append(K)
        self._fviews = []
        self.symbolic = None

    @staticmethod
    def _props(models):
        qp = np.concatenate([cm(m.quad_props, float) for m in models])
        tp = np.concatenate([cm(m.tria_props, float) for m in models])
        ep = np.concatenate([cm(m.tet_props, float) for m in models])
        return qp, tp, ep

    def update(self, models):
        """New section records for the same number of variants."""
        assert len(models) == self.V
        qp, tp, ep = self._props(models)
        check(self.ctx.h, self.ctx.L.jfem_batch_update_props(self.ctx.h, self.h, ptr(qp), ptr(tp), ptr(ep)))

    def assemble(self):
        bv = C.c_int64(-1); be = C.c_int64(-1)
        code = self.ctx.L.jfem_batch_assemble(self.ctx.h, self.h, C.byref(bv), C.byref(be))
        self.bad_variant, self.bad_element = bv.value, be.value
        check(self.ctx.h, code)

    def matrix(self, v: int) -> DeviceMatrix:
        return self._mviews[v]

    def factor(self, S: Symbolic, precision: int):
        bv = C.c_int64(-1); bc = C.c_int64(-1)
        code = self.ctx.L.jfem_batch_factor(self.ctx.h, self.h, S.h, int(precision), C.byref(bv), C.byref(bc))
        self.bad_variant, self.bad_column = bv.value, bc.value
        check(self.ctx.h, code)
        if self.symbolic is not S or not self._fviews or self._fviews[0].precision != precision:
            self._fviews = []
            for v in range(self.V):
                vh = C.c_void_p()
                check(self.ctx.h, self.ctx.L.jfem_batch_factor_view(self.h, v, C.byref(vh)))
                F = object.__new__(Factor)
                F.ctx, F.symbolic, F.precision, F.h, F.bad_column = self.ctx, S, precision, vh, -1
                self._fviews.append(F)
            self.symbolic = S

    def factor_view(self, v: int) -> "Factor":
        return self._fviews[v]

    def free(self):
        if self.h:
            for o in self._mviews + self._fviews:
                o.h = None                    # views die with the batch
            self.ctx.L.jfem_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def factor_with_fallback(S: Symbolic, K: DeviceMatrix, precision: int, ctx: Context | None = None):
    """Numeric factor in ``precision``; an FP32 factor that meets a non-positive
    pivot (cond(K_ff)·2^-24 too large) is redone in FP64, the way the reference
    falls back from a failed Cholesky (src/solver/boundary_conditions.jl:1660-1710).
    Returns ``(factor, precision_used)``; an FP64 breakdown raises."""
    try:
        return Factor(S, K, precision, ctx), precision
    except _lib.JfemError as e:
        if e.code != _lib.JFEM_ERR_NOT_SPD or precision == FP64:
            raise
    return Factor(S, K, FP64, ctx), FP64


def small_pivot_dofs(S: Symbolic, K: DeviceMatrix, Fct: Factor, threshold: float = 1e-7, guard: float = 1e-4,
                     ctx: Context | None = None):
    """Post-factor pivot check: global DOFs whose pivot ratio L_jj^2 / |K_jj| is
    below ``threshold`` (src/solver/boundary_conditions.jl:1662-1684; the ratio is
    taken under this factor's own elimination order, as the reference takes it
    under CHOLMOD's).  The reference decides on an FP64 factor.  An FP32 factor's
    ratios carry errors of order 2^-24 relative to |K_jj| times the pivot growth,
    so for one the ratios below ``guard`` are undecided and the whole set is
    decided on an FP64 factor of the same symbolic; with none below ``guard`` the
    FP32 factor decides (nothing is small).  Returns ``(dofs, decided_in)``."""
    if Fct.precision == FP64:
        return Fct.small_pivots(K, threshold), FP64
    if len(Fct.small_pivots(K, guard)) == 0:
        return np.zeros(0, np.int64), FP32
    F64 = Factor(S, K, FP64, ctx or Fct.ctx)
    try:
        return F64.small_pivots(K, threshold), FP64
    finally:
        F64.free()


class DeviceBuffer:
    """Device memory owned by the caller of the ``_dev`` entry points."""

    def __init__(self, nbytes: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(self.ctx.h, self.ctx.L.jfem_dev_alloc(self.ctx.h, max(self.nbytes, 8), C.byref(p)))
        self.ptr = p

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(self.ctx.h, self.ctx.L.jfem_memcpy_h2d(self.ctx.h, self.ptr, ptr(a), a.nbytes))

    def download(self, out: np.ndarray):
        assert out.flags["C_CONTIGUOUS"] and out.nbytes <= self.nbytes
        check(self.ctx.h, self.ctx.L.jfem_memcpy_d2h(self.ctx.h, ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx.L.jfem_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class LinearSolveCache:
    """Factor reuse across right-hand sides sharing one K and one free set
    (LinearSolveCacheEntry, src/solver/boundary_conditions.jl:62-125)."""

    def __init__(self):
        self.entries = {}

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, value):
        self.entries[key] = value


def apply_bc_and_solve(asm: AssembledStiffness, model, F_applied=None, *, precision: int = FP64,
                       tol: float = 1e-10, max_refine: int = 100, pivot_threshold: float = 1e-7,
                       linear_cache: LinearSolveCache | None = None, symbolic: Symbolic | None = None,
                       autospc: bool | None = None):
    """BC partition, factor and solve of K_ff u = F_ff on the GPU.

    Returns ``(u_global, fixed_dofs, spc_dofs, diagnostics)`` like the reference
    (src/solver/boundary_conditions.jl:1324, 1974).  ``precision`` = 64 (plain FP64
    factor) or 32 (FP32 factor + FP64 iterative refinement to ``tol``).
    ``autospc`` (default: ``model.meta["autospc"]``, else on, as PARAM,AUTOSPC
    defaults to YES: boundary_conditions.jl:3-18) gates the post-factor pivot
    check the way the reference gates it (:1684).  A ``symbolic`` passed in is
    used when its free set matches and is never freed here."""
    ctx = asm.K.ctx
    F = model.loads if F_applied is None else np.asarray(F_applied, float)
    if autospc is None:
        autospc = bool(getattr(model, "meta", {}).get("autospc", True))
    diag = {"bc_partition": {}, "linear_solver": {}}
    t0 = time.perf_counter()
    key = (id(asm.K), precision)
    entry = linear_cache.get(key) if linear_cache is not None else None
    if entry is None:
        fixed = bc_partition(asm.K, model.spc)
        own_S = not (symbolic is not None and np.array_equal(symbolic.fixed, fixed))
        S = Symbolic(asm.K, asm.mesh, fixed, ctx) if own_S else symbolic
        t1 = time.perf_counter()
        Fct, used = factor_with_fallback(S, asm.K, precision, ctx)
        n_post, decided = 0, used
        if autospc:
            small, decided = small_pivot_dofs(S, asm.K, Fct, pivot_threshold, ctx=ctx)
            if len(small):   # post-factor pivot check: extra SPC and refactor (:1684-1710)
                n_post = len(small)
                fixed = fixed.copy()
                fixed[small] = 1
                Fct.free()
                if own_S:
                    S.free()
                S = Symbolic(asm.K, asm.mesh, fixed, ctx)
                Fct, used = factor_with_fallback(S, asm.K, precision, ctx)
        t2 = time.perf_counter()
        entry = (fixed, S, Fct, t1 - t0, t2 - t1, n_post, used)
        diag["bc_partition"]["post_factorization_check_precision"] = decided
        if linear_cache is not None:
            linear_cache.put(key, entry)
        diag["linear_solver"]["cache_hit"] = False
    else:
        diag["linear_solver"]["cache_hit"] = True
    fixed, S, Fct, t_sym, t_num, n_post, used = entry
    t3 = time.perf_counter()
    u, rr, iters, ok, extra = Fct.solve(asm.K, F, tol=tol, max_refine=max_refine, details=True)
    t4 = time.perf_counter()
    spc = np.asarray(model.spc, np.uint8)
    diag["bc_partition"].update({"free_dofs": int((fixed == 0).sum()), "fixed_dofs": int(fixed.sum()),
                                 "explicit_spc_dofs": int(spc.sum()),
                                 "autospc_dofs": int(((fixed == 1) & (spc == 0)).sum()),
                                 "post_factorization_singular_dofs": n_post})
    diag["linear_solver"].update({"backend": f"jfem_hip_supernodal_fp{used}", "factor_precision": used,
                                  "fp32_breakdown_fallback": used != precision, "relative_residual": rr.tolist(),
                                  "relative_residual_fp64_output": extra["relres_fp64"].tolist(),
                                  "refinement_iterations": iters.tolist(), "converged": bool(ok),
                                  "symbolic": S.info, "time_symbolic_s": t_sym, "time_factor_s": t_num,
                                  "time_solve_s": t4 - t3})
    return u, np.where(fixed)[0], np.where(spc)[0], diag
