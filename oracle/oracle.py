"""ctypes front end of the CPU oracle (libjfem_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
ip64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libjfem_oracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libjfem_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_q4_local.argtypes = [dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.or_q4_macneal_parts.argtypes = [dp, dp, dp, C.c_int, dp, dp, dp, dp, dp]
        L.or_t3_local.argtypes = [dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.or_tet10.argtypes = [dp, C.c_double, C.c_double, dp]
        L.or_t3_shear_parts.argtypes = [dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_t3_guard.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_q4_global.argtypes = [dp, dp, C.c_double, dp]
        L.or_t3_global.argtypes = [dp, dp, C.c_double, dp]
        L.or_q4_recover.argtypes = [dp, dp, dp, C.c_int, dp]
        L.or_t3_recover.argtypes = [dp, dp, dp, dp]
        L.or_q4_kg.argtypes = [dp, dp, dp, dp]
        L.or_q4_kg_local.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.or_t3_kg.argtypes = [dp, dp, dp, dp]
        L.or_t3_kg_local.argtypes = [dp, dp, C.c_double, dp]
        L.or_assemble.argtypes = [C.c_int64, dp, dp, dp, C.c_int64, C.c_void_p, C.c_void_p,
                                  C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                  C.c_double, C.POINTER(C.c_int64), C.POINTER(C.POINTER(C.c_int64)),
                                  C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_double)),
                                  C.POINTER(C.c_int64)]
        L.or_free.argtypes = [C.c_void_p]
        L.or_bc_partition.argtypes = [C.c_int64, ip64, ip32, dp, u8p, C.c_double]
        L.or_bc_partition.restype = C.c_int64
        L.or_solve.argtypes = [C.c_int64, ip64, ip32, dp, u8p, C.c_int, C.c_void_p, C.c_int, dp, dp, dp,
                               C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int]
        L.or_pivot_ratios.argtypes = [C.c_int64, ip64, ip32, dp, u8p, ip64, dp, C.POINTER(C.c_int64)]
        L.or_eig3.argtypes = [dp, dp, dp]
        vpp = C.c_void_p
        L.or_load_vector.argtypes = [C.c_int64, dp, dp, dp, C.c_int64, vpp, C.c_int64, vpp, C.c_int64, vpp, vpp, vpp,
                                     C.c_int64, vpp, vpp, vpp, C.c_int64, vpp, vpp, vpp, C.c_double,
                                     vpp, C.c_int64, vpp, vpp, dp]
        L.or_jl_norm3.argtypes = [dp]
        L.or_jl_norm3.restype = C.c_double
        L.or_jl_normalize3.argtypes = [dp]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"oracle status {code} {what}")
        self.code = code


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _colmajor(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).ravel()


def q4_local(lc, Cm, Cb, Cs, h=1.0, E_ref=1.0, k6rot=100.0, drill_scale=1.0, rigid=False):
    Ke = np.zeros(576)
    st = lib().or_q4_local(_f(np.asarray(lc, float).reshape(4, 2).ravel()), _colmajor(Cm), _colmajor(Cb),
                           _colmajor(Cs), h, E_ref, k6rot, drill_scale, int(rigid), Ke)
    if st:
        raise OracleError(st)
    return Ke.reshape(24, 24, order="F")


def q4_macneal_parts(lc, Cb, Cs, rigid=False):
    D = np.zeros(48); Zb = np.zeros(16); Zs = np.zeros(16); Z = np.zeros(16); Kp = np.zeros(144)
    st = lib().or_q4_macneal_parts(_f(np.asarray(lc, float).reshape(4, 2).ravel()), _colmajor(Cb),
                                   _colmajor(Cs), int(rigid), D, Zb, Zs, Z, Kp)
    if st:
        raise OracleError(st)
    return {"D": D.reshape(4, 12, order="F"), "Zb": Zb.reshape(4, 4, order="F"),
            "Zs": Zs.reshape(4, 4, order="F"), "Z": Z.reshape(4, 4, order="F"),
            "Kp": Kp.reshape(12, 12, order="F")}


def t3_shear_parts(lc, Db, Ds):
    """CTRIA3 MITC3 shear pieces: tying rows D3 (3x9), tying rigidity W (3x3),
    residual bending flexibility Zb3 before and after the guard, and the rigidity
    that enters Ks = D3' W D3.  ``applies`` is False when Zb3 is skipped."""
    D3 = np.zeros(27); W = np.zeros(9); Zb = np.zeros(9); Zg = np.zeros(9); Wf = np.zeros(9)
    st = lib().or_t3_shear_parts(_f(np.asarray(lc, float).reshape(3, 2).ravel()), _colmajor(Db), _colmajor(Ds),
                                 D3, W, Zb, Zg, Wf)
    if st < 0:
        raise OracleError(-st)
    return {"D3": D3.reshape(3, 9), "W": W.reshape(3, 3, order="F"), "Zb": Zb.reshape(3, 3, order="F"),
            "Zb_guarded": Zg.reshape(3, 3, order="F"), "W_total": Wf.reshape(3, 3, order="F"), "applies": st == 1}


def t3_guard(Zb, La, Lb, c, s, Dn):
    """tria3_guard_bending_flexibility! on a copy of Zb: returns (Zb', changed)."""
    Z = _colmajor(Zb)
    st = lib().or_t3_guard(Z, La, Lb, c, s, Dn)
    if st < 0:
        raise OracleError(5)
    return Z.reshape(3, 3, order="F"), bool(st)


def t3_local(lc, Dm, Db, Ds, h=1.0, G_ref=1.0, k6rot=100.0):
    Ke = np.zeros(324)
    st = lib().or_t3_local(_f(np.asarray(lc, float).reshape(3, 2).ravel()), _colmajor(Dm), _colmajor(Db),
                           _colmajor(Ds), h, G_ref, k6rot, Ke)
    if st:
        raise OracleError(st)
    return Ke.reshape(18, 18, order="F")


def tet10(X, E, nu):
    Ke = np.zeros(900)
    st = lib().or_tet10(_colmajor(np.asarray(X, float).reshape(10, 3)), E, nu, Ke)
    if st:
        raise OracleError(st)
    return Ke.reshape(30, 30, order="F")


def q4_global(p, prop, k6rot=100.0):
    Ke = np.zeros(576)
    st = lib().or_q4_global(_f(np.asarray(p, float).reshape(4, 3).ravel()), _f(prop), k6rot, Ke)
    if st:
        raise OracleError(st)
    return Ke.reshape(24, 24, order="F")


def t3_global(p, prop, k6rot=100.0):
    Ke = np.zeros(324)
    st = lib().or_t3_global(_f(np.asarray(p, float).reshape(3, 3).ravel()), _f(prop), k6rot, Ke)
    if st:
        raise OracleError(st)
    return Ke.reshape(18, 18, order="F")


Q4R_N, T3R_N = 44, 20


def q4_recover(p, prop, ug, mean_shear=True):
    """Raw CQUAD4 recovery vector (44 doubles, layout of or_q4_recover) for the
    4 x 3 corner coordinates p, section prop and global displacements ug (24);
    mean_shear selects the SOL 101/105/106 branch."""
    out = np.zeros(Q4R_N)
    st = lib().or_q4_recover(_f(np.asarray(p, float).reshape(4, 3).ravel()), _f(prop),
                             _f(np.asarray(ug, float).ravel()), int(bool(mean_shear)), out)
    if st:
        raise OracleError(st)
    return out


def t3_recover(p, prop, ug):
    """Raw CTRIA3 recovery vector (20 doubles, layout of or_t3_recover)."""
    out = np.zeros(T3R_N)
    st = lib().or_t3_recover(_f(np.asarray(p, float).reshape(3, 3).ravel()), _f(prop),
                             _f(np.asarray(ug, float).ravel()), out)
    if st:
        raise OracleError(st)
    return out


def shell_recover(model, u, mean_shear=True):
    """Per-element raw recovery of a model for one displacement vector u (ndof):
    (CQUAD4 rows (n_quad4, 44), CTRIA3 rows (n_tria3, 20))."""
    u = np.asarray(u, float).reshape(-1)
    q = np.zeros((len(model.quads), Q4R_N))
    for e, nodes in enumerate(np.asarray(model.quads)):
        ug = np.concatenate([u[6 * n:6 * n + 6] for n in nodes])
        q[e] = q4_recover(model.xyz[nodes], model.quad_props[e], ug, mean_shear)
    t = np.zeros((len(model.trias), T3R_N))
    for e, nodes in enumerate(np.asarray(model.trias)):
        ug = np.concatenate([u[6 * n:6 * n + 6] for n in nodes])
        t[e] = t3_recover(model.xyz[nodes], model.tria_props[e], ug)
    return q, t


def q4_kg(p, prop, ug):
    """Global 24 x 24 differential stiffness of one flat isotropic CQUAD4 (or_q4_kg)
    for corner coordinates p (4 x 3), section prop and global displacements ug."""
    out = np.zeros(576)
    st = lib().or_q4_kg(_f(np.asarray(p, float).reshape(4, 3).ravel()), _f(prop),
                        _f(np.asarray(ug, float).ravel()), out)
    if st:
        raise OracleError(st)
    return out.reshape(24, 24, order="F")


def q4_kg_local(lc, ue, E, nu, h, assumed_transverse=True):
    """The element-frame kernel (or_q4_kg_local) on local coordinates lc (4 x 2)
    and element-frame displacements ue (24)."""
    out = np.zeros(576)
    st = lib().or_q4_kg_local(_f(np.asarray(lc, float).reshape(4, 2).ravel()), _f(np.asarray(ue, float).ravel()),
                              float(E), float(nu), float(h), int(bool(assumed_transverse)), out)
    if st:
        raise OracleError(st)
    return out.reshape(24, 24, order="F")


def t3_kg(p, prop, ug):
    """Global 18 x 18 differential stiffness of one CTRIA3 (or_t3_kg)."""
    out = np.zeros(324)
    st = lib().or_t3_kg(_f(np.asarray(p, float).reshape(3, 3).ravel()), _f(prop),
                        _f(np.asarray(ug, float).ravel()), out)
    if st:
        raise OracleError(st)
    return out.reshape(18, 18, order="F")


def t3_kg_local(lc, sig, h):
    """The CTRIA3 element-frame kernel (or_t3_kg_local) for membrane stresses sig."""
    out = np.zeros(324)
    st = lib().or_t3_kg_local(_f(np.asarray(lc, float).reshape(3, 2).ravel()), _f(np.asarray(sig, float)),
                              float(h), out)
    if st:
        raise OracleError(st)
    return out.reshape(18, 18, order="F")


def kg_assemble(model, u):
    """Global differential stiffness (scipy CSR, ndof x ndof) of a shell model
    (CQUAD4, CTRIA3) for the displacement vector u: element matrices summed in
    triplet order (CQUAD4s, then CTRIA3s)."""
    u = np.asarray(u, float).reshape(-1)
    rows, cols, vals = [], [], []
    for kind, conn, props, fn, nn in (("q", model.quads, model.quad_props, q4_kg, 4),
                                      ("t", model.trias, model.tria_props, t3_kg, 3)):
        for e, nodes in enumerate(np.asarray(conn)):
            ug = np.concatenate([u[6 * n:6 * n + 6] for n in nodes])
            kg = fn(model.xyz[nodes], props[e], ug)
            dofs = np.concatenate([np.arange(6 * n, 6 * n + 6) for n in nodes])
            rows.append(np.repeat(dofs, 6 * nn)); cols.append(np.tile(dofs, 6 * nn)); vals.append(kg.ravel())
    if len(model.tets):
        raise OracleError(5)
    n = model.ndof
    if not vals:
        return sp.csr_matrix((n, n))
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsr()


def buckling_dense(K, Kg, fixed):
    """All buckling factors of the pencil K_ff phi = lambda (-Kg_ff) phi on the free
    DOFs (dense, for small models): (sorted finite eigenvalues, modes (ndof x m))
    with each mode scaled to unit largest component, K_ff positive definite."""
    import scipy.linalg as sl
    free = np.where(np.asarray(fixed) == 0)[0]
    Kf = K[free][:, free].toarray()
    Gf = -Kg[free][:, free].toarray()
    Gf = 0.5 * (Gf + Gf.T)
    # K_ff SPD: eigen-pairs of (G, K): G phi = mu K phi, lambda = 1 / mu
    mu, V = sl.eigh(Gf, Kf)
    keep = np.abs(mu) > 1e-14 * np.abs(mu).max()
    lam = 1.0 / mu[keep]
    V = V[:, keep]
    order = np.argsort(lam)
    lam, V = lam[order], V[:, order]
    modes = np.zeros((K.shape[0], len(lam)))
    modes[free] = V
    modes /= np.abs(modes).max(axis=0)
    return lam, modes


def _cm_i32(a):
    a = np.asarray(a, dtype=np.int32)
    return np.ascontiguousarray(a.T).ravel() if a.size else np.zeros(1, np.int32)


def _cm_f64(a):
    a = np.asarray(a, dtype=np.float64)
    return np.ascontiguousarray(a.T).ravel() if a.size else np.zeros(1)


def assemble(model):
    """Global K (scipy CSR, ndof x ndof) with sparse(I,J,V) duplicate summation."""
    L = lib()
    x = _f(model.xyz[:, 0]); y = _f(model.xyz[:, 1]); z = _f(model.xyz[:, 2])
    qc, qp = _cm_i32(model.quads), _cm_f64(model.quad_props)
    tc, tp = _cm_i32(model.trias), _cm_f64(model.tria_props)
    ec, ep = _cm_i32(model.tets), _cm_f64(model.tet_props)
    nnz = C.c_int64(); rp = C.POINTER(C.c_int64)(); ci = C.POINTER(C.c_int32)(); cv = C.POINTER(C.c_double)()
    bad = C.c_int64(-1)
    st = L.or_assemble(model.n_nodes, x, y, z, len(model.quads), qc.ctypes.data, qp.ctypes.data,
                       len(model.trias), tc.ctypes.data, tp.ctypes.data,
                       len(model.tets), ec.ctypes.data, ep.ctypes.data, model.k6rot,
                       C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(cv), C.byref(bad))
    if st:
        raise OracleError(st, f"element {bad.value}")
    n = model.ndof
    indptr = np.ctypeslib.as_array(rp, shape=(n + 1,)).copy()
    indices = np.ctypeslib.as_array(ci, shape=(max(nnz.value, 1),))[:nnz.value].copy()
    data = np.ctypeslib.as_array(cv, shape=(max(nnz.value, 1),))[:nnz.value].copy()
    L.or_free(rp); L.or_free(ci); L.or_free(cv)
    return sp.csr_matrix((data, indices, indptr), shape=(n, n))


def bc_partition(K, spc, eps_zero=1e-8):
    """Fixed-DOF mask after SPC + AUTOSPC + grid-point singularity processing."""
    K = sp.csr_matrix(K)
    K.sort_indices()
    fixed = np.ascontiguousarray(np.asarray(spc, dtype=np.uint8).copy())
    lib().or_bc_partition(K.shape[0], K.indptr.astype(np.int64), K.indices.astype(np.int32),
                          _f(K.data), fixed, eps_zero)
    return fixed


def solve(K, fixed, B, order="nd", xyz=None, refine=0):
    """Reference sparse Cholesky solve of K_ff u = F_ff for each column of B.
    ``refine`` = 0 is the plain solve; ``refine`` > 0 allows that many steps of
    iterative refinement with long-double residual and iterate, returning the
    solution correctly rounded to double up to a few units (for
    cond(K_ff)·2^-53 << 1) and the residual of that double vector.
    Returns (u (ndof, nrhs), relres (nrhs,), info dict)."""
    K = sp.csr_matrix(K)
    K.sort_indices()
    B = np.asarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B[:, None]
    n = K.shape[0]
    nrhs = B.shape[1]
    Bc = np.ascontiguousarray(B.T).ravel()
    U = np.zeros(n * nrhs)
    rr = np.zeros(nrhs)
    nzl = C.c_int64(); fl = C.c_double(); ns = C.c_int64()
    xyzc = _f(np.asarray(xyz, float).ravel()) if xyz is not None else None
    st = lib().or_solve(n, K.indptr.astype(np.int64), K.indices.astype(np.int32), _f(K.data),
                        np.ascontiguousarray(np.asarray(fixed, np.uint8)),
                        1 if (order == "nd" and xyz is not None) else 0,
                        xyzc.ctypes.data if xyzc is not None else None, nrhs, Bc, U, rr,
                        C.byref(nzl), C.byref(fl), C.byref(ns), int(refine))
    if st:
        raise OracleError(st)
    return U.reshape(nrhs, n).T, rr, {"nnz_L": nzl.value, "flops_colcount2": fl.value,
                                      "small_pivots": ns.value}


def pivot_ratios(K, fixed, order):
    """Post-factor pivot ratios L_kk^2/|K_gg| of K_ff's Cholesky factor under the
    elimination order ``order`` (global DOFs, a permutation of the free ones);
    position k's ratio belongs to DOF order[k].  Raises OracleError on a
    non-positive pivot."""
    K = sp.csr_matrix(K)
    K.sort_indices()
    order = np.ascontiguousarray(np.asarray(order, np.int64))
    ratios = np.zeros(max(len(order), 1))
    ns = C.c_int64()
    st = lib().or_pivot_ratios(K.shape[0], K.indptr.astype(np.int64), K.indices.astype(np.int32), _f(K.data),
                               np.ascontiguousarray(np.asarray(fixed, np.uint8)), order, ratios, C.byref(ns))
    if st:
        raise OracleError(st)
    return ratios[:len(order)]


def eig3(A):
    w = np.zeros(3); V = np.zeros(9)
    lib().or_eig3(_colmajor(A), w, V)
    return w, V.reshape(3, 3, order="F")


def sol101(model, order="nd", refine=0):
    """assemble -> BC partition -> factor -> solve, all on the CPU oracle.
    ``refine`` > 0 adds long-double iterative refinement (see ``solve``)."""
    K = assemble(model)
    fixed = bc_partition(K, model.spc)
    u, rr, info = solve(K, fixed, model.loads, order=order, xyz=model.xyz, refine=refine)
    info["free_dofs"] = int((fixed == 0).sum())
    return K, fixed, u, rr, info


def jl_norm3(v):
    """Julia's norm of a 3-vector, as restated in the oracle."""
    return lib().or_jl_norm3(np.ascontiguousarray(v, dtype=float))


def jl_normalize3(v):
    a = np.array(v, dtype=float)
    lib().or_jl_normalize3(a)
    return a


def load_vector(model, pload4=None, force=None, moment=None, scale=1.0, grav=None, shell_areal=None):
    """Load vector of one load set (FORCE / MOMENT / shell PLOAD4 / GRAV, card
    order), all 6 n DOFs: src/solver/loads.jl:510-587, 680-772, 1338-1340.  Cards
    as in jf_2026_05_25_amd.solver.LoadSet; grav: (A (n,), N (n, 3)) with the
    per-shell areal mass shell_areal (rho T + NSM, CQUAD4s then CTRIA3s)."""
    keep = []

    def arr(a, dt, shape=None):
        if a is None:
            return None
        a = np.ascontiguousarray(np.asarray(a, dt))
        if shape is not None:
            a = a.reshape(shape)
        keep.append(a)
        return a.ctypes.data if a.size else None

    def cards(c):
        if c is None or len(c[0]) == 0:
            return 0, None, None, None
        n = len(c[0])
        d = c[2] if len(c) > 2 else None
        return n, arr(c[0], np.int32), arr(c[1], float, (n,)), arr(d, float, (n, 3))

    npl, pe, pp, pd = cards(pload4)
    nf, fn, fm, fd = cards(force)
    nm, mn, mm, md = cards(moment)
    xyz = np.asarray(model.xyz, float)
    x, y, z = (np.ascontiguousarray(xyz[:, k]) for k in range(3))
    qc = _cm_i32(model.quads)
    tc = _cm_i32(model.trias)
    keep += [qc, tc]
    F = np.zeros(6 * len(xyz))
    ng, ga, gn, sa = 0, None, None, None
    if grav is not None and len(grav[0]):
        ng = len(grav[0])
        ga, gn = arr(grav[0], float, (ng,)), arr(grav[1], float, (ng, 3))
        sa = arr(shell_areal, float, (len(model.quads) + len(model.trias),))
    st = lib().or_load_vector(len(xyz), x, y, z, len(model.quads), qc.ctypes.data if qc.size else None,
                              len(model.trias), tc.ctypes.data if tc.size else None, npl, pe, pp, pd,
                              nf, fn, fm, fd, nm, mn, mm, md, float(scale), sa, ng, ga, gn, F)
    if st != 0:
        raise OracleError(st, "load_vector")
    return F
