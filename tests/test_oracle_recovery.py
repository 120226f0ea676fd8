"""Pins of the oracle's shell force / stress / strain recovery.

The oracle restates recover_shell_stresses! (src/solver/stress_recovery.jl:505)
for isotropic PSHELL CQUAD4 / CTRIA3 in both of its branches: the one of
solution sequences 101/105/106 (MacNeal area-mean shear, corner moments of the
enriched section fields) and the one of other sequences (MITC4 centre shear
blended with the equilibrium shear, corner moments of the compatible
curvature).  It is checked against

* the reference's own recorded rows: the MYSTRAN deck (SOL 1, so the second
  branch) holds forces, bilinear corner forces, fibre stresses and strains of
  its CQUAD4 for a membrane and a pressure subcase;
* closed forms, on distorted elements in arbitrary 3D orientation (the element
  frame is not reproduced here: results are compared through tensor
  invariants, trace and determinant, which any in-plane frame preserves):
  rigid motion recovers nothing; a constant membrane strain gives N = A eps at
  the centre and at every corner with zero M and Q; a constant curvature gives
  the printed moment -D kappa everywhere with zero N and Q;
* the definition of the area-mean shear: Q times the area is the work of the
  element's own stiffness (oracle.q4_global) against a uniform rotation about
  the element's in-plane axes, which a wrong row, sign or area would break;
* the MITC3 resultant of a constant transverse shear, (5/6) G h gamma.
"""
import numpy as np
import pytest

from jf_2026_05_25_amd import meshes, recovery

E, NU, H = 7.0e10, 0.3, 0.004
PROP = meshes.pshell_props(1, H, E, NU, 0.0, 1.0, 5.0 / 6.0)[0]


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


QUAD_PLANE = np.array([[0.0, 0.0], [1.3, 0.1], [1.1, 1.2], [-0.2, 0.9]]) * 0.35   # distorted, counter-clockwise
TRIA_PLANE = np.array([[0.0, 0.0], [0.9, 0.15], [0.25, 0.7]]) * 0.35


def _element(plane, seed):
    """Plane coordinates (n, 2) placed in 3D by a random rotation and offset."""
    R = _rotation(seed)
    off = np.random.default_rng(seed + 1).normal(size=3)
    X = np.c_[plane, np.zeros(len(plane))] @ R.T + off
    return X, R, off


def _nodal_u(plane, R, grad=None, kappa=None, shear=None):
    """Global nodal displacements (6 per node) of plane fields: membrane
    u_plane = grad @ (x, y); constant curvature kappa = (k1, k2, k3) through the
    rotations theta_y = k1 x + k3/2 y, theta_x = -k2 y - k3/2 x and the matching
    w = -(k1 x^2 + k3 x y + k2 y^2)/2 (no transverse shear); a constant
    transverse shear (gx, gy) through w = gx x + gy y with zero rotations."""
    u = np.zeros((len(plane), 6))
    for k, (x, y) in enumerate(plane):
        t = np.zeros(3); r = np.zeros(3)
        if grad is not None:
            t[:2] += grad @ np.array([x, y])
        if kappa is not None:
            k1, k2, k3 = kappa
            r[1] += k1 * x + 0.5 * k3 * y
            r[0] += -k2 * y - 0.5 * k3 * x
            t[2] += -(k1 * x * x + k3 * x * y + k2 * y * y) / 2
        if shear is not None:
            t[2] += shear[0] * x + shear[1] * y
        u[k, :3] = R @ t
        u[k, 3:] = R @ r
    return u.ravel()


def _inv(v):
    """(trace, determinant) of the symmetric 2 x 2 tensor of an engineering
    triple (xx, yy, 2 xy) or a resultant triple (xx, yy, xy)."""
    return v[0] + v[1], v[0] * v[1] - v[2] * v[2]


def _tensor_triple(eng):
    return np.array([eng[0], eng[1], eng[2] / 2])


PLANE_D = E / (1 - NU * NU) * np.array([[1, NU, 0], [NU, 1, 0], [0, 0, (1 - NU) / 2]])


def _quad(oracle, X, u, mean_shear):
    r = oracle.q4_recover(X, PROP, u, mean_shear)
    return r


WARPED = np.array([[0.0, 0.0, 0.0], [0.4, 0.02, 0.03], [0.45, 0.38, -0.02], [-0.03, 0.35, 0.04]])


@pytest.mark.parametrize("mean_shear", [True, False])
def test_rigid_motion_recovers_nothing(oracle, mean_shear):
    """Also for a warped CQUAD4, whose corner heights would turn a rigid rotation
    into membrane strain without the finite-warp displacement map."""
    rng = np.random.default_rng(3)
    for seed in range(4):
        X, R, off = _element(QUAD_PLANE, seed)
        Xt, Rt, offt = _element(TRIA_PLANE, seed)
        Xw = WARPED @ R.T + off
        for pts, fn in ((X, lambda u: oracle.q4_recover(X, PROP, u, mean_shear)),
                        (Xw, lambda u: oracle.q4_recover(Xw, PROP, u, mean_shear)),
                        (Xt, lambda u: oracle.t3_recover(Xt, PROP, u))):
            a, w = rng.normal(size=3), rng.normal(size=3) * 0.01
            u = np.concatenate([np.r_[a + np.cross(w, p), w] for p in pts])
            r = fn(u)
            assert np.max(np.abs(r)) < 1e-12 * E * H, np.max(np.abs(r))


@pytest.mark.parametrize("mean_shear", [True, False])
def test_constant_membrane_strain(oracle, mean_shear):
    """N = h D eps at the centre and at each corner (the Wilson modes vanish for
    the patch state), M = Q = 0; both fibres carry D eps and the strain rows eps."""
    g = np.array([[2e-4, -3e-5], [7e-5, -1e-4]])
    eps = np.array([g[0, 0], g[1, 1], g[0, 1] + g[1, 0]])
    Nx = H * PLANE_D @ eps
    for seed in range(3):
        X, R, _ = _element(QUAD_PLANE, seed)
        r = oracle.q4_recover(X, PROP, _nodal_u(QUAD_PLANE, R, grad=g), mean_shear)
        scale = np.max(np.abs(Nx))
        for n in [r[0:3]] + [r[20 + 3 * k:23 + 3 * k] for k in range(4)]:
            assert np.allclose(_inv(n), _inv(Nx), rtol=1e-10, atol=1e-10 * scale ** 2), (n, Nx)
        assert np.max(np.abs(r[3:8])) < 1e-10 * scale
        assert np.max(np.abs(r[32:44])) < 1e-10 * scale
        for s in (r[8:11], r[11:14]):
            assert np.allclose(_inv(s), _inv(PLANE_D @ eps), rtol=1e-10)
        for e in (r[14:17], r[17:20]):
            assert np.allclose(_inv(_tensor_triple(e)), _inv(_tensor_triple(eps)), rtol=1e-10)
        Xt, Rt, _ = _element(TRIA_PLANE, seed)
        rt = oracle.t3_recover(Xt, PROP, _nodal_u(TRIA_PLANE, Rt, grad=g))
        assert np.allclose(_inv(rt[0:3]), _inv(Nx), rtol=1e-10)
        assert np.max(np.abs(rt[3:8])) < 1e-10 * scale


@pytest.mark.parametrize("mean_shear", [True, False])
def test_constant_curvature(oracle, mean_shear):
    """The printed moment -(h^3/12) D kappa at the centre and at every corner, no
    membrane force, no transverse shear; fibre stresses -/+ (h/2) D kappa."""
    kap = np.array([0.03, -0.011, 0.017])
    Mx = -(H ** 3 / 12) * PLANE_D @ kap
    scale = np.max(np.abs(Mx))
    for seed in range(3):
        X, R, _ = _element(QUAD_PLANE, seed)
        r = oracle.q4_recover(X, PROP, _nodal_u(QUAD_PLANE, R, kappa=kap), mean_shear)
        for m in [r[3:6]] + [r[32 + 3 * k:35 + 3 * k] for k in range(4)]:
            assert np.allclose(_inv(m), _inv(Mx), rtol=1e-9, atol=1e-9 * scale ** 2), (m, Mx)
        assert np.max(np.abs(r[0:3])) < 1e-9 * scale / H
        assert np.max(np.abs(r[20:32])) < 1e-9 * scale / H
        assert np.max(np.abs(r[6:8])) < 1e-9 * scale / H, r[6:8]
        for s, z in ((r[8:11], -H / 2), (r[11:14], H / 2)):
            assert np.allclose(_inv(s), _inv(z * PLANE_D @ kap), rtol=1e-9)
        Xt, Rt, _ = _element(TRIA_PLANE, seed)
        rt = oracle.t3_recover(Xt, PROP, _nodal_u(TRIA_PLANE, Rt, kappa=kap))
        assert np.allclose(_inv(rt[3:6]), _inv(Mx), rtol=1e-9)
        assert np.max(np.abs(rt[6:8])) < 1e-9 * scale / H


def _rectangle(seed, a=0.3, b=0.2):
    """A 2a x 2b rectangle centred at the origin of its plane, turned in the plane
    by psi and tilted into 3D.  Its element axes are its sides (its diagonals are
    symmetric about them): returns the corners, the element axes e1, e2, the
    normal, and the corners' element coordinates."""
    rng = np.random.default_rng(100 + seed)
    psi = rng.uniform(-np.pi, np.pi)
    R = _rotation(seed)
    c, s_ = np.cos(psi), np.sin(psi)
    e1, e2, n = R @ np.r_[c, s_, 0.0], R @ np.r_[-s_, c, 0.0], R @ np.r_[0.0, 0.0, 1.0]
    loc = np.array([[-a, -b], [a, -b], [a, b], [-a, b]])
    off = rng.normal(size=3)
    X = np.array([off + x * e1 + y * e2 for x, y in loc])
    return X, e1, e2, n, loc


def _frame_u(loc, e1, e2, n, uvw, rot):
    """Global nodal displacements from element-frame fields uvw(x, y) -> (u, v, w)
    and rot(x, y) -> (theta_x, theta_y)."""
    out = []
    for x, y in loc:
        u, v, w = uvw(x, y)
        tx, ty = rot(x, y)
        out.append(np.r_[u * e1 + v * e2 + w * n, tx * e1 + ty * e2])
    return np.concatenate(out)


@pytest.mark.parametrize("mean_shear", [True, False])
def test_in_plane_bending_of_a_rectangle(oracle, mean_shear):
    """u = b x y - a y^2/2, v = a x y - b x^2/2: the condensed Wilson modes remove
    the parasitic shear and the Poisson coupling of the bilinear field, leaving
    uniaxial bending in each direction: N = E h (b y, a x, 0) at each corner (in
    connectivity order) and zero at the centre."""
    a, b = 2e-3, -3e-3
    for seed in range(3):
        X, e1, e2, n, loc = _rectangle(seed)
        u = _frame_u(loc, e1, e2, n, lambda x, y: (b * x * y - a * y * y / 2, a * x * y - b * x * x / 2, 0.0),
                     lambda x, y: (0.0, 0.0))
        r = oracle.q4_recover(X, PROP, u, mean_shear)
        scale = E * H * 0.3 * 3e-3
        assert np.max(np.abs(r[0:3])) < 1e-12 * scale
        for k, (x, y) in enumerate(loc):
            exp = E * H * np.array([b * y, a * x, 0.0])
            assert np.allclose(r[20 + 3 * k:23 + 3 * k], exp, rtol=0, atol=1e-12 * scale), (k, r[20 + 3 * k:23 + 3 * k], exp)


@pytest.mark.parametrize("mean_shear", [True, False])
def test_plate_bending_of_a_rectangle(oracle, mean_shear):
    """Rotations theta_y = b x y - a y^2/2, theta_x = -(a x y - b x^2/2): curvatures
    (b y, a x) with the bilinear field's parasitic twist.  Sequences 101/105/106:
    the bending enrichment of the section fields removes the twist and the Poisson
    coupling like the Wilson modes above, corner moments -(h^3/12) E (b y, a x);
    other sequences: the compatible curvature, corner moments -(h^3/12) D (b y, a x, 0)
    (mx, my; the twist is the centre value in both)."""
    a, b = 2e-3, -3e-3
    D = PLANE_D
    for seed in range(3):
        X, e1, e2, n, loc = _rectangle(seed)
        u = _frame_u(loc, e1, e2, n, lambda x, y: (0.0, 0.0, 0.0),
                     lambda x, y: (-(a * x * y - b * x * x / 2), b * x * y - a * y * y / 2))
        r = oracle.q4_recover(X, PROP, u, mean_shear)
        scale = H ** 3 / 12 * E * 0.3 * 3e-3
        assert np.max(np.abs(r[3:6])) < 1e-12 * scale
        for k, (x, y) in enumerate(loc):
            k_ = np.array([b * y, a * x, 0.0])
            exp = -(H ** 3 / 12) * (E * k_ if mean_shear else D @ k_)
            got = r[32 + 3 * k:35 + 3 * k]
            assert np.allclose(got[:2], exp[:2], rtol=0, atol=1e-12 * scale), (k, got, exp)
            assert got[2] == r[5]


def test_legacy_shear_blend_on_a_rectangle(oracle):
    """Other sequences, a state whose pieces are all exact on a rectangle: the
    plate bending state above plus w = g0 x.  Its MITC4 tying shears are exact
    (w quadratic and the rotations linear along each edge): (g0, -b a_x^2) with
    a_x the half-width, so the MITC4 centre shear is phi2 (5/6) G h (g0, -b a_x^2),
    phi2 = min(1, 10 h^2 / area).  The equilibrium shear of the compatible corner
    moments is -(h^3/12) D12 (a, b).  Each component keeps the smaller magnitude:
    here the MITC4 x value (g0 small) and the equilibrium y value."""
    a, b, g0, ax, ay = 2e-3, -3e-3, 1e-9, 0.3, 0.2
    G = E / (2 * (1 + NU))
    phi2 = min(1.0, 10 * H * H / (4 * ax * ay))
    mitc = phi2 * 5 / 6 * G * H * np.array([g0, -b * ax * ax])
    eq = -(H ** 3 / 12) * PLANE_D[0, 1] * np.array([a, b])
    assert abs(mitc[0]) < abs(eq[0]) and abs(mitc[1]) > abs(eq[1])
    for seed in range(3):
        X, e1, e2, n, loc = _rectangle(seed, ax, ay)
        u = _frame_u(loc, e1, e2, n, lambda x, y: (0.0, 0.0, g0 * x - b * x * x * y / 2 + a * x * y * y / 2),
                     lambda x, y: (-(a * x * y - b * x * x / 2), b * x * y - a * y * y / 2))
        q = oracle.q4_recover(X, PROP, u, False)[6:8]
        assert np.isclose(q[0], mitc[0], rtol=1e-9, atol=0) and np.isclose(q[1], eq[1], rtol=1e-9, atol=0), (q, mitc, eq)


def test_tria3_mitc3_tying(oracle):
    """MITC3 assumed shear (a + c s, b - c r) in the covariant components of the
    natural coordinates, fitted to the exact tangential shear at the three edge
    midpoints (edges 1-2, 1-3 along r, s; edge 2-3 along s - r), evaluated at the
    centroid: here solved as that 3 x 3 interpolation problem, for a rotation
    field with curl (w = 0, theta_y = alpha y, theta_x = -beta x), whose shear is
    not a gradient.  The element x axis is edge 1-2, y completes the plane."""
    alpha, beta = 2e-3, -5e-4
    G = E / (2 * (1 + NU))
    for seed in range(3):
        X, R, off = _element(TRIA_PLANE, seed)
        e1 = X[1] - X[0]; e1 /= np.linalg.norm(e1)
        nrm = np.cross(e1, X[2] - X[0]); nrm /= np.linalg.norm(nrm)
        e2 = np.cross(nrm, e1)
        loc = np.array([[(p - X.mean(0)) @ e1, (p - X.mean(0)) @ e2] for p in X])
        u = _frame_u(loc, e1, e2, nrm, lambda x, y: (0.0, 0.0, 0.0), lambda x, y: (-beta * x, alpha * y))
        # exact shear field gamma = grad w + (theta_y, -theta_x) = (alpha y, beta x)
        gam = lambda p: np.array([alpha * p[1], beta * p[0]])
        J = np.array([loc[1] - loc[0], loc[2] - loc[0]])          # rows: dx/dr, dx/ds
        mA, mB, mC = (loc[0] + loc[1]) / 2, (loc[0] + loc[2]) / 2, (loc[1] + loc[2]) / 2
        # tangential (covariant) shears: along r at A, along s at B, along s - r at C
        tA, tB, tC = J[0] @ gam(mA), J[1] @ gam(mB), (J[1] - J[0]) @ gam(mC)
        # (a + c s) at A (r, s) = (1/2, 0); (b - c r) at B (0, 1/2); (b - c r) - (a + c s) at C (1/2, 1/2)
        Amat = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [-1.0, 1.0, -1.0]])
        a_, b_, c_ = np.linalg.solve(Amat, np.array([tA, tB, tC]))
        gcov = np.array([a_ + c_ / 3, b_ - c_ / 3])
        expQ = 5 / 6 * G * H * np.linalg.solve(J, gcov)
        q = oracle.t3_recover(X, PROP, u)[6:8]
        assert np.allclose(q, expQ, rtol=1e-10, atol=1e-12 * np.max(np.abs(expQ))), (q, expQ)


def test_mean_shear_is_work_conjugate_to_uniform_rotation(oracle):
    """Q_x A = w_y' K u and Q_y A = -w_x' K u, with K the element's own stiffness
    (no drilling) and w_x, w_y uniform rotations about the element's in-plane
    axes, which produce neither membrane strain nor curvature: the work is that
    of the shear operator alone.  The element is a rectangle (turned in its plane
    by psi, then tilted into 3D), whose axes are its sides: its diagonals are
    symmetric about them.  Its area is 4 |J_c|."""
    rng = np.random.default_rng(11)
    a, b = 0.37, 0.21
    for seed in range(4):
        psi = rng.uniform(-np.pi, np.pi)
        c, s_ = np.cos(psi), np.sin(psi)
        rot = np.array([[c, -s_], [s_, c]])
        plane = np.array([[0, 0], [a, 0], [a, b], [0, b]]) @ rot.T
        X, R, _ = _element(plane, seed)
        e1, e2 = R @ np.r_[c, s_, 0.0], R @ np.r_[-s_, c, 0.0]
        K = oracle.q4_global(X, PROP, k6rot=0.0)
        u = rng.normal(size=24) * 1e-4
        q = oracle.q4_recover(X, PROP, u, True)[6:8]
        wy = np.concatenate([np.r_[0, 0, 0, e2] for _ in range(4)])
        wx = np.concatenate([np.r_[0, 0, 0, e1] for _ in range(4)])
        qx, qy = wy @ K @ u / (a * b), -(wx @ K @ u) / (a * b)
        scale = max(abs(qx), abs(qy))
        assert abs(q[0] - qx) < 1e-9 * scale and abs(q[1] - qy) < 1e-9 * scale, (q, qx, qy)


def test_tria3_constant_transverse_shear(oracle):
    """MITC3: w = gx x + gy y with zero rotations has constant shear strain
    (gx, gy); the centroid resultant is (5/6) G h gamma (its magnitude here, the
    element frame being arbitrary in the plane)."""
    G = E / (2 * (1 + NU))
    gam = np.array([3e-4, -1e-4])
    for seed in range(3):
        X, R, _ = _element(TRIA_PLANE, seed)
        r = oracle.t3_recover(X, PROP, _nodal_u(TRIA_PLANE, R, shear=gam))
        assert np.isclose(np.hypot(*r[6:8]), 5 / 6 * G * H * np.hypot(*gam), rtol=1e-12)
        assert np.max(np.abs(r[0:6])) < 1e-9 * G * H * np.hypot(*gam)


def test_legacy_branch_keeps_the_smaller_shear(oracle):
    """Other solution sequences: for a constant transverse shear the compatible
    corner moments are constant, so their equilibrium shear is zero and the
    blend keeps it (the smaller magnitude) in place of the MITC4 value."""
    X, R, _ = _element(QUAD_PLANE, 5)
    r = oracle.q4_recover(X, PROP, _nodal_u(QUAD_PLANE, R, shear=np.array([2e-4, 1e-4])), False)
    assert np.max(np.abs(r[6:8])) < 1e-6 * (E / 2.6) * H * 2e-4


def _rel_close(a, b, rtol, floor):
    return abs(a - b) <= rtol * max(abs(b), floor)


def test_mystran_rows_match_reference(oracle):
    """The recorded rows of the MYSTRAN case (SOL 1: the legacy branch) recovered
    from the recorded displacements of the same run: forces, corner forces,
    stresses (with von Mises and principal values) and strains, for the membrane
    and the pressure subcase, to 1e-9 of each row's largest entry.  (Recovery is
    compared from the reference's own displacements: the pressure subcase's
    displacements are reproduced only to the historical tolerance of this deck.)"""
    from _helpers import golden_cases, golden_model
    c = [g for g in golden_cases() if g["case"] == "MYSTRAN_sol101_simple"][0]
    assert c["sol"] == 1
    checked = 0
    for s in c["subcases"]:
        ref = s["shell_recovery"]
        m = golden_model(c, s, membrane_modes=6)
        q, t = oracle.shell_recover(m, m.meta["expected"], mean_shear=False)
        grids = [[int(g) for g in e["nodes"]] for e in c["cquad4"]]
        rows = recovery.shell_rows(q, [e["eid"] for e in c["cquad4"]], m.quad_props[:, 0], "quad4", grids)
        for key in ("forces", "forces_bilin", "stresses", "strains"):
            got, exp = rows[key], ref[key]
            assert len(got) == len(exp)
            for a, b in zip(got, exp):
                flat_a, flat_b = _flatten(a), _flatten(b)
                assert flat_a.keys() == flat_b.keys(), (key, flat_a.keys() ^ flat_b.keys())
                scale = max(abs(v) for v in flat_b.values() if isinstance(v, float))
                for k in flat_b:
                    if isinstance(flat_b[k], float):
                        assert _rel_close(flat_a[k], flat_b[k], 1e-9, scale), (key, k, flat_a[k], flat_b[k])
                    else:
                        assert flat_a[k] == flat_b[k], (key, k)
                checked += 1
    assert checked == 2 * (1 + 5 + 1 + 1)


def _flatten(d, pre=""):
    out = {}
    for k, v in d.items():
        if isinstance(v, dict):
            out.update(_flatten(v, pre + k + "."))
        elif isinstance(v, (int, float)) and not isinstance(v, bool) and k not in ("eid", "grid_id"):
            out[pre + k] = float(v)
        else:
            out[pre + k] = v
    return out
