"""SOL 105 linear buckling on the device: the differential stiffness of a static
preload, shifted factorizations on the symbolic analysis of K, and EIGB root
extraction by subspace inverse iteration.

Reference flow (_solve_sol105, src/JFEMSolver.jl:3156-3600): the static preload
subcase is solved with the elastic K; assemble_geometric_stiffness
(src/solver/assembly.jl:6106) builds Kg from the preload displacements; the
buckling subcase's EIGB request (_resolve_eigb_request, src/solver/helpers.jl:
56-97) selects the positive roots lambda of K_ff phi = lambda (-Kg_ff) phi in
(L1, L2], at most NDP of them (3 NEP when NDP is blank), and
_solve_eigb_buckling (:280-420) returns them ascending with each mode scaled to
unit largest free component (NORM = MAX).

Here Kg is generated and assembled on the device on K's pattern
(jfem_assemble_kg), K + sigma Kg is factored on K's symbolic analysis
(jfem_matrix_shift + jfem_factor_create), the number of roots in a finite
window comes from the inertia of the scaled shifted matrices at its ends
(jfem_matrix_shift_scaled + jfem_factor_inertia: an unpivoted signed
factorization on the same symbolic analysis), and the roots come from subspace
inverse iteration: X <- (K + sigma Kg)^{-1} (Kg X) with the device factor and
refined solves, Rayleigh-Ritz on the projected pencil (X'KX, -X'KgX) whose
O(ndof) products run on the device (jfem_matrix_spmv, jfem_block_gram,
jfem_block_combine); only the p x p projected eigenproblem is solved on the
host.  The iteration stops when the wanted modes' residuals
||K x + lambda Kg x|| / ||K x|| (formed on the device from K Y and Kg Y) are below
``res_tol`` and the wanted Ritz values have settled.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sl

from . import solver
from ._lib import JfemError, check, ptr

MAX_BLOCK = 32          # columns of the device block kernels
SOLVE_CHUNK = 16        # right-hand sides per refined solve


@dataclass
class EigbRequest:
    """An EIGB entry (SID METHOD L1 L2 NEP NDP NDN; NORM = MAX) resolved as
    _resolve_eigb_request does (src/solver/helpers.jl:56-97)."""
    sid: int = 1
    method: str = "INV"
    lower: float = -math.inf
    upper: float = math.inf
    nep: int | None = None
    ndp: int | None = None
    ndn: int | None = None

    def __post_init__(self):
        if self.method not in ("INV", "SINV"):
            raise ValueError(f"EIGB SID={self.sid} requires METHOD=INV or SINV; received {self.method!r}")
        for k in ("nep", "ndp", "ndn"):
            v = getattr(self, k)
            if v is not None and not (isinstance(v, (int, np.integer)) and v > 0):
                raise ValueError(f"EIGB SID={self.sid} {k.upper()} must be a positive integer; received {v!r}")
        if self.nep is not None:
            if self.ndp is None:
                self.ndp = 3 * self.nep
            if self.ndn is None:
                self.ndn = 3 * self.nep
        if self.method == "INV" and self.nep is None:
            raise ValueError(f"EIGB SID={self.sid} METHOD=INV requires NEP")
        if not self.lower < self.upper:
            raise ValueError(f"EIGB SID={self.sid} requires finite specified L1 < L2")
        if self.upper > 0 and self.ndp is None and not math.isfinite(self.upper):
            raise ValueError(f"EIGB SID={self.sid} requires L2 or a positive-root count NDP")

    @property
    def positive_window(self):
        """(lower, upper, quota) of the positive branch, or None when empty."""
        lo, hi = max(0.0, self.lower), self.upper
        return (lo, hi, self.ndp) if hi > lo else None


class GeometricStiffness(solver.DeviceMatrix):
    """Kg of the shells (CQUAD4, CTRIA3) for a preload u (device CSR values on K's pattern)."""

    def __init__(self, asm: solver.AssembledStiffness, u, ctx: solver.Context | None = None):
        self.ctx = ctx or asm.K.ctx
        self.pattern = asm.pattern
        u = np.ascontiguousarray(np.asarray(u, float).reshape(-1))
        if u.size != asm.ndof:
            raise ValueError("u must hold one value per DOF")
        h = C.c_void_p(); bad = C.c_int64(-1)
        code = self.ctx.L.jfem_assemble_kg(self.ctx.h, asm.mesh.h, asm.pattern.h, ptr(u), C.byref(h), C.byref(bad))
        self.bad_element = bad.value
        check(self.ctx.h, code)
        self.h = h


class ShiftedMatrix(solver.DeviceMatrix):
    """K + sigma Kg on K's pattern (jfem_matrix_shift), or with ``scaled`` its
    diagonal scaling D^-1/2 (K + sigma Kg) D^-1/2, D = diag(K)
    (jfem_matrix_shift_scaled); ``shift`` overwrites it."""

    def __init__(self, K: solver.DeviceMatrix, sigma: float, Kg: solver.DeviceMatrix, scaled: bool = False):
        self.ctx = K.ctx
        self.pattern = K.pattern
        self.h = None
        self.scaled = scaled
        self.shift(K, sigma, Kg)

    def shift(self, K, sigma, Kg):
        h = C.c_void_p(self.h.value if self.h is not None else None)
        fn = self.ctx.L.jfem_matrix_shift_scaled if self.scaled else self.ctx.L.jfem_matrix_shift
        check(self.ctx.h, fn(self.ctx.h, K.h, float(sigma), Kg.h, C.byref(h)))
        self.h = h
        self.sigma = float(sigma)


def inertia(S: solver.Symbolic, A: solver.DeviceMatrix):
    """(negative pivots, smallest, largest pivot magnitude) of the unpivoted signed
    factorization of A_ff on the symbolic analysis S (jfem_factor_inertia)."""
    ctx = A.ctx
    neg = C.c_int64(0); mn = C.c_double(0); mx = C.c_double(0); bad = C.c_int64(-1)
    check(ctx.h, ctx.L.jfem_factor_inertia(ctx.h, S.h, A.h, C.byref(neg), C.byref(mn), C.byref(mx), C.byref(bad)))
    return int(neg.value), float(mn.value), float(mx.value)


def interval_count(K: solver.DeviceMatrix, Kg: solver.DeviceMatrix, S: solver.Symbolic, lower: float,
                   upper: float) -> int:
    """Roots of K_ff phi = lambda (-Kg_ff) phi in (lower, upper], 0 <= lower <
    upper < inf, by inertia: the negative pivots of D^-1/2 (K + sigma Kg) D^-1/2 at
    the padded ends, at two paddings that must agree (_eigb_interval_count,
    src/solver/helpers.jl:105-143).  Raises ValueError where the count is not
    established (pivots within 64 eps of zero relative to the largest, unequal
    counts at the two paddings)."""
    if not (0.0 <= lower < upper < math.inf):
        raise ValueError("the interval count requires finite 0 <= lower < upper")
    scale = max(abs(lower), abs(upper), np.finfo(float).tiny)
    pad = 128 * np.finfo(float).eps * scale
    eps = np.finfo(float).eps
    A = None

    def at(sigma):
        nonlocal A
        if A is None:
            A = ShiftedMatrix(K, sigma, Kg, scaled=True)
        else:
            A.shift(K, sigma, Kg)
        neg, mn, mx = inertia(S, A)
        if mn <= 64 * eps * mx:
            raise ValueError("interval end within rounding of a root: unpivoted LDL' pivots are unsafe; "
                             "move L1/L2 away from a root")
        return neg

    try:
        counts = []
        for mult in (1.0, 8.0):
            a = max(0.0, lower - mult * pad)
            b = upper + mult * pad
            na = 0 if a == 0.0 else at(a)
            nb = at(b)
            if nb < na:
                raise ValueError("positive-branch inertia count is inconsistent")
            counts.append(nb - na)
        if counts[0] != counts[1]:
            raise ValueError("interval count is unstable near an endpoint; move L1/L2 away from a root")
        return counts[0]
    finally:
        if A is not None:
            A.free()


def element_kg(mesh: solver.DeviceMesh, u) -> tuple[np.ndarray, np.ndarray]:
    """Per-element global differential stiffness for displacements u:
    (n_quad4, 24, 24) of the CQUAD4s and (n_tria3, 18, 18) of the CTRIA3s."""
    ctx = mesh.ctx
    nq, nt = mesh.model.quads.shape[0], mesh.model.trias.shape[0]
    out = np.zeros(max(576 * nq + 324 * nt, 1))
    u = np.ascontiguousarray(np.asarray(u, float).reshape(-1))
    bad = C.c_int64(-1)
    code = ctx.L.jfem_shell_kg_batch(ctx.h, mesh.h, ptr(u), ptr(out), C.byref(bad))
    if code:
        check(ctx.h, code)
    kq = out[:576 * nq].reshape(nq, 24, 24).transpose(0, 2, 1).copy()
    kt = out[576 * nq:576 * nq + 324 * nt].reshape(nt, 18, 18).transpose(0, 2, 1).copy()
    return kq, kt


class _Blocks:
    """Device blocks of ndof x p doubles (column-major) and the block products."""

    def __init__(self, ctx: solver.Context, n: int, p: int, fixed: np.ndarray):
        self.ctx, self.n, self.p = ctx, n, p
        self.fixed = np.ascontiguousarray(np.asarray(fixed, np.uint8))
        self.bufs = {k: solver.DeviceBuffer(8 * n * p, ctx) for k in ("X", "W", "Y", "KY", "GY")}

    def col(self, name, c):
        return C.c_void_p(self.bufs[name].ptr.value + 8 * self.n * c)

    def spmv(self, A, src, dst, ncol):
        check(self.ctx.h, self.ctx.L.jfem_matrix_spmv(self.ctx.h, A.h, ptr(self.fixed), ncol, self.bufs[src].ptr,
                                                      self.bufs[dst].ptr))

    def gram(self, a, b, p, q):
        out = np.zeros(p * q)
        check(self.ctx.h, self.ctx.L.jfem_block_gram(self.ctx.h, self.n, p, self.bufs[a].ptr, q, self.bufs[b].ptr,
                                                     ptr(out)))
        return out.reshape(q, p).T

    def combine(self, src, Q, dst, accumulate=False):
        p, q = Q.shape
        Qc = np.ascontiguousarray(Q.T).ravel()       # column-major p x q
        check(self.ctx.h, self.ctx.L.jfem_block_combine(self.ctx.h, self.n, p, self.bufs[src].ptr, q, ptr(Qc),
                                                        self.bufs[dst].ptr, int(bool(accumulate))))

    def solve(self, F: solver.Factor, A, src, dst, ncol, tol, max_refine):
        iters = 0
        for c0 in range(0, ncol, SOLVE_CHUNK):
            nr = min(SOLVE_CHUNK, ncol - c0)
            it = np.zeros(nr, np.int32); rr = np.zeros(nr)
            code = self.ctx.L.jfem_solve_dev(self.ctx.h, F.h, A.h, nr, self.col(src, c0), self.col(dst, c0), tol,
                                             max_refine, ptr(it), ptr(rr), None, None)
            if code not in (0, 8):
                check(self.ctx.h, code)
            iters += int(it.sum())
        return iters

    def download(self, name, ncol):
        out = np.zeros((ncol, self.n))
        self.bufs[name].download(out)
        return out.T.copy()

    def free(self):
        for b in self.bufs.values():
            b.free()


@dataclass
class BucklingResult:
    eigenvalues: np.ndarray                  # ascending positive roots in the window
    modes: np.ndarray                        # ndof x m, NORM = MAX
    diagnostics: dict = field(default_factory=dict)


def _ritz(Kr, Br):
    """Projected pencil Br q = mu Kr q (Kr SPD): eigenvalues lambda = 1/mu of the
    positive mu, ascending, then the rest; Kr-orthonormal vectors."""
    Kr = 0.5 * (Kr + Kr.T)
    Br = 0.5 * (Br + Br.T)
    mu, Q = sl.eigh(Br, Kr)
    order = np.argsort(-mu)                  # largest mu: smallest positive lambda
    mu, Q = mu[order], Q[:, order]
    with np.errstate(divide="ignore"):
        lam = np.where(mu > 0, 1.0 / mu, np.inf)
    return lam, mu, Q


def buckling_roots(K: solver.DeviceMatrix, Kg: solver.DeviceMatrix, fixed: np.ndarray, factor: solver.Factor,
                   window: tuple, *, factor_matrix: solver.DeviceMatrix | None = None, block: int | None = None,
                   tol: float = 1e-10, res_tol: float = 1e-9, max_iter: int = 500, solve_tol: float = 1e-13,
                   seed: int = 20260525) -> BucklingResult:
    """Positive roots of K_ff phi = lambda (-Kg_ff) phi in (lower, upper], at most
    ``quota`` (window = (lower, upper, quota)).  ``factor`` factors
    ``factor_matrix`` (default K), i.e. K + sigma Kg for a shift sigma below the
    first root, on the free set ``fixed == 0``; refined solves run against
    ``factor_matrix``.  Converged when every wanted mode's residual
    ||K x + lambda Kg x|| / ||K x|| is at most ``res_tol`` and the wanted Ritz
    values, with the first one past the window or the quota, change by at most
    ``tol`` relative between iterations."""
    t0 = time.perf_counter()
    ctx = K.ctx
    A = factor_matrix or K
    lower, upper, quota = window
    n = K.pattern.ndof
    free = np.asarray(fixed) == 0
    n_free = int(free.sum())
    want = min(quota if quota is not None else n_free, n_free)
    p = block or min(MAX_BLOCK, n_free, max(want + 8, 2 * want))
    if p < 1:
        return BucklingResult(np.zeros(0), np.zeros((n, 0)), {"free_dofs": n_free})
    rng = np.random.default_rng(seed)
    X0 = rng.standard_normal((n, p))
    X0[~free] = 0.0
    B = _Blocks(ctx, n, p, fixed)
    res = np.zeros(0)
    try:
        B.bufs["X"].upload(np.ascontiguousarray(X0.T))
        lam_prev = None
        solves = 0
        converged = False
        for it in range(1, max_iter + 1):
            B.spmv(Kg, "X", "W", p)                              # W = Kg X
            solves += B.solve(factor, A, "W", "Y", p, solve_tol, 100)   # Y = A^{-1} Kg X
            B.spmv(K, "Y", "KY", p)
            B.spmv(Kg, "Y", "GY", p)
            Kr = B.gram("Y", "KY", p, p)
            Br = -B.gram("Y", "GY", p, p)
            lam, mu, Q = _ritz(Kr, Br)
            B.combine("Y", Q, "X")                               # X'KX = I
            m = min(int(np.sum((lam > lower) & (lam <= upper))), want)
            # residuals of the wanted Ritz pairs: K x = (K Y) q, Kg x = (Kg Y) q
            if m:
                Qm = Q[:, :m]
                B.combine("KY", Qm, "W")
                kk = np.diag(B.gram("W", "W", m, m)).copy()
                B.combine("GY", Qm * lam[:m], "W", accumulate=True)
                rr = np.diag(B.gram("W", "W", m, m))
                res = np.sqrt(np.maximum(rr, 0.0) / np.maximum(kk, 1e-300))
            else:
                res = np.zeros(0)
            # the wanted values and the first Ritz value past the window or the quota
            k_chk = min(m + 1, p - 1) if m < p else m
            if lam_prev is not None and k_chk > 0:
                rel = np.abs(lam[:k_chk] - lam_prev[:k_chk]) / np.maximum(np.abs(lam[:k_chk]), 1e-300)
                finite = np.isfinite(rel)
                if np.all(rel[finite] <= tol) and np.all(res <= res_tol):
                    converged = True
                    break
            lam_prev = lam
        X = B.download("X", p)[:, :m]
    finally:
        B.free()
    vals = lam[:m].copy()
    # NORM = MAX over the analysis (free) set, largest component positive
    modes = X.copy()
    for k in range(m):
        col = modes[:, k]
        i = int(np.argmax(np.abs(np.where(free, col, 0.0))))
        a = col[i]
        if a != 0.0:
            modes[:, k] = col / a
    diag = {"free_dofs": n_free, "block": p, "iterations": it, "converged": bool(converged),
            "triangular_solves": solves, "mode_residual": res.tolist(),
            "window": {"lower": lower, "upper": upper, "quota": quota},
            "in_window_count": int(np.sum((lam > lower) & (lam <= upper))),
            "ritz_values": [float(v) for v in lam],
            "complete": bool(np.any(lam[m:] > upper)) or m == want,
            "time_s": time.perf_counter() - t0}
    if not converged:
        raise JfemError(8, f"buckling roots did not converge in {max_iter} iterations (residuals {res.tolist()})")
    return BucklingResult(vals, modes, diag)


def solve_buckling(model, request: EigbRequest, F_static=None, *, ctx: solver.Context | None = None,
                   shift: float = 0.0, asm: solver.AssembledStiffness | None = None, **kw):
    """SOL 105 on the device for one static preload and one EIGB request:
    assemble K, solve the preload (FP64 factor, refined), assemble Kg on K's
    pattern, factor K + shift Kg on K's symbolic analysis (the static factor
    itself when shift == 0) and extract the roots.  Returns
    (BucklingResult, u_static, diagnostics)."""
    ctx = ctx or solver.default_context()
    t0 = time.perf_counter()
    asm = asm or solver.assemble_stiffness(model, ctx)
    cache = solver.LinearSolveCache()
    u, fixed, spc, sdiag = solver.apply_bc_and_solve(asm, model, F_static, precision=64, linear_cache=cache)
    u = u.reshape(model.ndof, -1)[:, 0]
    fixed_e, S, Fct = next(iter(cache.entries.values()))[:3]
    t1 = time.perf_counter()
    Kg = GeometricStiffness(asm, u, ctx)
    t2 = time.perf_counter()
    win = request.positive_window
    if win is None:
        raise NotImplementedError("EIGB windows without a positive branch (L2 <= 0) are not on this path")
    lower, upper, quota = win
    Ash, Fsh = None, Fct
    available = None
    try:
        if math.isfinite(upper):
            # the roots in the window by inertia; the target is NDP of them at most
            available = interval_count(asm.K, Kg, S, lower, upper)
            quota = available if quota is None else min(quota, available)
        t_count = time.perf_counter()
        if quota == 0:
            res = BucklingResult(np.zeros(0), np.zeros((model.ndof, 0)),
                                 {"free_dofs": int((fixed_e == 0).sum()), "iterations": 0})
        else:
            if shift != 0.0:
                Ash = ShiftedMatrix(asm.K, shift, Kg)
                Fsh = solver.Factor(S, Ash, solver.FP64, ctx)
            res = buckling_roots(asm.K, Kg, fixed_e, Fsh, (lower, upper, quota), factor_matrix=Ash, **kw)
            if available is not None and len(res.eigenvalues) != quota:
                raise JfemError(8, f"EIGB SID={request.sid}: {len(res.eigenvalues)} of {quota} required roots "
                                   f"recovered (interval inertia {available})")
    finally:
        if Fsh is not Fct:
            Fsh.free()
        if Ash is not None:
            Ash.free()
        Kg.free()
    t3 = time.perf_counter()
    res.diagnostics["interval_count"] = available
    res.diagnostics["target_count"] = quota
    diag = {"static": sdiag, "phase_timings": {"static_reference_solve": t1 - t0, "kg_assembly": t2 - t1,
                                               "interval_count": t_count - t2, "buckling_eigensolve": t3 - t_count},
            "eigenvalue_extraction_method": request.method, "shift": shift}
    return res, u, diag
