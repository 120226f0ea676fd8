"""Pins of the oracle's shell transverse-shear pieces that the element-level
closed forms (CST membrane, constant curvature, rigid motions) leave open.

CTRIA3 (stiffness_tria3_matrices_generic, src/FEMKernels.jl:9375-9809):
  * residual bending flexibility Zb3: for an isotropic bending section the
    directional law (:9659-9700, per-edge channels divided by their own
    directional stiffness) must reduce exactly to the derived isotropic law
    written with different descriptors (:9584-9590, sum/difference form);
    and it may depend on Db only through the directional stiffnesses
    (the perturbation that leaves every D_dir unchanged has no effect, :9640-9648);
  * tying rigidity Wsh (:9541-9550): the MITC3 three-midpoint integral
    (:9552-9560) with the third mode (u s, -v r), u = 3 La / P, v = 3 Lb / P,
    and the r s cross-moment (2 - c)/6 (A/3) in place of 1/4 (A/3)
    (:9524-9536); at the equilateral triangle u = v = 1 and c = 1/2, so Wsh is
    the plain MITC3 sum;
  * the guard tria3_guard_bending_flexibility! (:9281-9312): the reference's own
    property tests, utils/test/tria_orthogonal_residual_energy_test.jl:24-50 and
    utils/test/tria_bending_flexibility_guard_test.jl:7-20;
  * the element's material and geometric scaling laws
    (utils/test/tria_bending_flexibility_guard_test.jl:21-62).
CQUAD4 (quad4_natural_shear_flexibility!, src/FEMKernels.jl:4223-4352):
  * Zb scales as 1/Cb and Zs as 1/Cs, Z_total(s^3 Cb, s Cs) = Zb/s^3 + Zs/s;
  * a cyclic relabelling of the corners with the matching 90-degree rotation of
    the coordinates and of Cb, Cs preserves the shear energies; rigid shear
    gives Z_total = Zb; zero bending rigidity gives a vanishing plate block
    (utils/test/natural_shear_work_transport_test.jl:132-172).
"""
import numpy as np
import pytest

from test_oracle_elements import CORNERS, macneal_geometries, rotation_fields, tying_rows


# ---------------------------------------------------------------- CTRIA3 helpers
def tri(angle_deg, La, Lb, rot=0.0, shift=(0.0, 0.0)):
    """Triangle with edge 1-2 of length La along x (then rotated by ``rot``), edge
    1-3 of length Lb at ``angle_deg`` from it."""
    a = np.radians(angle_deg)
    X = np.array([[0.0, 0.0], [La, 0.0], [Lb * np.cos(a), Lb * np.sin(a)]])
    R = np.array([[np.cos(rot), -np.sin(rot)], [np.sin(rot), np.cos(rot)]])
    return X @ R.T + np.asarray(shift)


def iso_bending(D, nu):
    return D * np.array([[1.0, nu, 0.0], [nu, 1.0, 0.0], [0.0, 0.0, (1.0 - nu) / 2.0]])


def descriptors(X):
    e12, e13, e23 = X[1] - X[0], X[2] - X[0], X[2] - X[1]
    La, Lb, Lc = np.linalg.norm(e12), np.linalg.norm(e13), np.linalg.norm(e23)
    c = e12 @ e13 / (La * Lb)
    return La, Lb, La + Lb + Lc, c, np.sqrt(1.0 - c * c)


SHAPES = [(ang, La, Lb) for ang in (30.0, 55.0, 90.0, 120.0, 150.0)
          for (La, Lb) in ((1.0, 1.0), (1.0, 3.0), (3.0, 1.0), (0.7, 1.9))]


@pytest.mark.parametrize("ang,La,Lb", SHAPES)
@pytest.mark.parametrize("nu", [0.0, 0.3, 0.45])
def test_t3_zb3_isotropic_law(oracle, ang, La, Lb, nu):
    """Isotropic Db: Zb3 in the basis normalised by (La, Lb, perimeter) and scaled by
    D is the derived law (src/FEMKernels.jl:9563-9590), whatever nu:
        R = (La-Lb)^2/(La Lb), T = (La^2-Lb^2)/(La Lb)
        a11 + a22 = 2[(1 - c + c^2)/(6s) + 1/6] + R (1 + c^2)/(6s),  a11 - a22 = T s / 6
        a12 = (-1 + 4c - c^2)/(12 s) - 1/6 + R c/(6s),  a13 = -1/6, a23 = 1/6, a33 = 1/6"""
    X = tri(ang, La, Lb, rot=0.37)
    D = 2.9e3
    Db = iso_bending(D, nu)
    parts = oracle.t3_shear_parts(X, Db, np.eye(2) * 4e7)
    assert parts["applies"]
    La_, Lb_, P, c, s = descriptors(X)
    R = (La_ - Lb_) ** 2 / (La_ * Lb_)
    T = (La_ ** 2 - Lb_ ** 2) / (La_ * Lb_)
    ssum = 2 * ((1 - c + c * c) / (6 * s) + 1.0 / 6) + R * (1 + c * c) / (6 * s)
    sdif = T * s / 6
    a = np.array([[(ssum + sdif) / 2, (-1 + 4 * c - c * c) / (12 * s) - 1.0 / 6 + R * c / (6 * s), -1.0 / 6],
                  [0.0, (ssum - sdif) / 2, 1.0 / 6],
                  [0.0, 0.0, 1.0 / 6]])
    a = np.triu(a) + np.triu(a, 1).T
    S = np.array([La_, Lb_, P])
    expected = a * np.outer(S, S) / D
    np.testing.assert_allclose(parts["Zb"], expected, rtol=1e-13, atol=1e-15 * np.abs(expected).max())


@pytest.mark.parametrize("ang,La,Lb", SHAPES[::3])
def test_t3_zb3_depends_on_directional_stiffness_only(oracle, ang, La, Lb):
    """D_dir(n) = w' Db w with w = (nx^2, ny^2, 2 nx ny): the perturbation
    D12 += d, D33 -= d/2 leaves every D_dir and D11, D22 unchanged, so Zb3 too;
    scaling Db scales Zb3 inversely."""
    X = tri(ang, La, Lb, rot=-0.2)
    Db = np.array([[3.1, 0.7, 0.2], [0.7, 1.9, -0.1], [0.2, -0.1, 0.9]]) * 1e3
    Ds = np.array([[5.0, 0.4], [0.4, 3.0]]) * 1e7
    base = oracle.t3_shear_parts(X, Db, Ds)["Zb"]
    d = 0.3e3
    pert = Db.copy()
    pert[0, 1] += d; pert[1, 0] += d; pert[2, 2] -= d / 2
    np.testing.assert_allclose(oracle.t3_shear_parts(X, pert, Ds)["Zb"], base, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(oracle.t3_shear_parts(X, 4.0 * Db, Ds)["Zb"], base / 4.0, rtol=1e-13)


def mitc3_rigidity(X, Ds, third=(1.0, 1.0), rs_moment=0.25):
    """MITC3 tying rigidity from the three edge-midpoint rule
    (src/FEMKernels.jl:9552-9560) with the third mode (u s, -v r) and the r s
    cross-moment ``rs_moment`` (A/3): Wsh = sum_pts (A/3) P' G P, P = [1 0 u s; 0 1 -v r],
    G = J^-T Ds J^-1."""
    J = np.array([X[1] - X[0], X[2] - X[0]])
    Ji = np.linalg.inv(J)
    G = Ji.T @ Ds @ Ji
    A = 0.5 * abs(np.linalg.det(J))
    u, v = third
    W = np.zeros((3, 3))
    for r, s in ((0.5, 0.0), (0.0, 0.5), (0.5, 0.5)):
        P = np.array([[1.0, 0.0, u * s], [0.0, 1.0, -v * r]])
        W += (A / 3) * P.T @ G @ P
    # the midpoint rule's r s moment is sum(r s)/3 = 1/12 = (1/4)(1/3)
    W[2, 2] += -2 * u * v * G[0, 1] * (A / 3) * (rs_moment - 0.25)
    return W


@pytest.mark.parametrize("rot", [0.0, 0.9, 2.3])
@pytest.mark.parametrize("L", [0.01, 1.0, 7.0])
def test_t3_tying_rigidity_equilateral_is_mitc3(oracle, rot, L):
    X = tri(60.0, L, L, rot=rot, shift=(0.3, -1.1))
    Ds = np.array([[6.1, 0.9], [0.9, 4.4]]) * 1e6
    W = oracle.t3_shear_parts(X, iso_bending(1.0, 0.3), Ds)["W"]
    np.testing.assert_allclose(W, mitc3_rigidity(X, Ds), rtol=1e-12, atol=1e-12 * np.abs(W).max())


@pytest.mark.parametrize("ang,La,Lb", SHAPES)
def test_t3_tying_rigidity_general(oracle, ang, La, Lb):
    X = tri(ang, La, Lb, rot=1.3)
    Ds = np.array([[6.1, 0.9], [0.9, 4.4]]) * 1e6
    La_, Lb_, P, c, s = descriptors(X)
    W = oracle.t3_shear_parts(X, iso_bending(1.0, 0.3), Ds)["W"]
    ref = mitc3_rigidity(X, Ds, third=(3 * La_ / P, 3 * Lb_ / P), rs_moment=(2 - c) / 6)
    np.testing.assert_allclose(W, ref, rtol=1e-12, atol=1e-12 * np.abs(W).max())


def test_t3_shear_enters_as_series_compliance(oracle):
    """W_total = inv(sym(inv(Wsh) + Zb3_guarded)) and the element's plate block
    carries D3' W_total D3 (MacNeal's series law on the MITC3 basis, :9725-9733)."""
    X = tri(70.0, 1.0, 1.6, rot=0.4)
    Db = iso_bending(2.0e3, 0.3)
    Ds = np.eye(2) * 3e6
    p = oracle.t3_shear_parts(X, Db, Ds)
    Zt = np.linalg.inv(p["W"]) + p["Zb_guarded"]
    np.testing.assert_allclose(p["W_total"], np.linalg.inv(0.5 * (Zt + Zt.T)), rtol=1e-12)
    Dm = np.eye(3) * 1e6
    K1 = oracle.t3_local(X, Dm, Db, Ds, k6rot=0.0)
    bi = [2, 3, 4, 8, 9, 10, 14, 15, 16]
    D3 = p["D3"]
    # bending part: constant curvature B' Db B A, independent of the shear
    A = 0.5 * abs(np.linalg.det(np.array([X[1] - X[0], X[2] - X[0]])))
    y, x = X[:, 1], X[:, 0]
    bv = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]]) / (2 * A)
    cv = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]]) / (2 * A)
    Bb = np.zeros((3, 9))
    for i in range(3):
        Bb[0, 3 * i + 2] = bv[i]; Bb[1, 3 * i + 1] = -cv[i]
        Bb[2, 3 * i + 1] = -bv[i]; Bb[2, 3 * i + 2] = cv[i]
    expected = A * Bb.T @ Db @ Bb + D3.T @ p["W_total"] @ D3
    np.testing.assert_allclose(K1[np.ix_(bi, bi)], expected, rtol=1e-11, atol=1e-11 * np.abs(expected).max())


# ---------------------------------------------------------------- CTRIA3 guard (reference tests)
def make_flexibility(P, La, Lb, c, s, Dn):
    back = np.array([[1.0, 0.0], [c, s]])
    H = back @ P @ back.T
    scale = 6 * (La / Lb) * s * Dn
    q = np.array([-0.13, 0.21]); z = 1.7
    Z = np.zeros((3, 3))
    Z[:2, :2] = H * np.outer([La, Lb], [La, Lb]) / scale + np.outer(q, q) / z
    Z[:2, 2] = q; Z[2, :2] = q; Z[2, 2] = z
    return Z


def orthogonal_block(Z, La, Lb, c, s, Dn):
    scale = 6 * (La / Lb) * s * Dn
    H = scale * (Z[:2, :2] - np.outer(Z[:2, 2], Z[2, :2]) / Z[2, 2]) / np.outer([La, Lb], [La, Lb])
    Q = np.array([[1.0, 0.0], [-c / s, 1 / s]])
    return Q @ H @ Q.T


def test_t3_guard_orthogonal_residual_energy(oracle):
    """utils/test/tria_orthogonal_residual_energy_test.jl:24-50."""
    n = 0
    for c in (-0.4, 0.0, 0.6):
        for La, Lb in ((0.03, 1.2), (1.0, 0.7), (3.0, 7.0)):
            for a in (-0.2, 1e-5, 0.3, 2.0):
                for b in (-1.1, 0.02, 0.8):
                    s = np.sqrt(1 - c * c); d = 0.7; normal = 2.3
                    P = np.array([[a, b], [b, d]])
                    Z = make_flexibility(P, La, Lb, c, s, normal)
                    Zn, changed = oracle.t3_guard(Z, La, Lb, c, s, normal)
                    got = orthogonal_block(Zn, La, Lb, c, s, normal)
                    assert got[1, 1] == pytest.approx(d, rel=2e-8, abs=2e-10)
                    assert got[0, 0] >= 1e-4 - 3e-10
                    assert got[0, 1] ** 2 <= 0.85 * got[0, 0] * got[1, 1] + 2e-10
                    assert np.all(np.linalg.eigvalsh(0.5 * (Zn + Zn.T)) > 0)
                    assert np.array_equal(Zn[2, :], Z[2, :]) and np.array_equal(Zn[:, 2], Z[:, 2])
                    if a >= 1e-4 and b * b <= 0.85 * a * d:
                        assert not changed and np.array_equal(Zn, Z)
                    else:
                        assert changed
                        assert got[0, 0] == pytest.approx(max(a, 1e-4), rel=2e-6, abs=3e-10)
                    n += 1
    assert n == 108


def test_t3_guard_inactive_and_slender(oracle):
    """utils/test/tria_bending_flexibility_guard_test.jl:7-20."""
    ordinary = np.array([[2.0, 0.1, 0.2], [0.1, 3.0, -0.1], [0.2, -0.1, 1.0]])
    Zn, changed = oracle.t3_guard(ordinary, 1.0, 1.0, 0.6, 0.8, 4.0)
    assert not changed and np.array_equal(Zn, ordinary)
    slender = np.array([[1e-8, 2e-5, 0.0], [2e-5, 2.0, 0.0], [0.0, 0.0, 1.0]])
    Zn, changed = oracle.t3_guard(slender, 0.004, 1.0, 0.8, 0.6, 3.0)
    assert changed
    assert np.array_equal(Zn[2, :], slender[2, :]) and np.array_equal(Zn[:, 2], slender[:, 2])
    assert np.linalg.eigvalsh(0.5 * (Zn + Zn.T)).min() > 0
    with pytest.raises(oracle.OracleError):
        bad = np.array([[1.0, 0.0, 0.0], [0.0, -5.0, 0.0], [0.0, 0.0, 1.0]])
        oracle.t3_guard(bad, 1.0, 1.0, 0.0, 1.0, 1.0)


@pytest.mark.parametrize("frac", [0.70, 0.80, 0.84, 0.86, 0.95, 1.5])
def test_t3_guard_coupling_clip(oracle, frac):
    """The orthogonal coupling is clipped at |b| = sqrt(0.85 a q) (:9302-9303):
    b^2 = frac * a d stays untouched up to frac = 0.85, above it the clipped block
    sits exactly on the limit, with a and d unchanged."""
    La, Lb, c, normal, a, d = 1.3, 0.8, 0.25, 1.9, 0.6, 0.7
    s = np.sqrt(1 - c * c)
    b = -np.sqrt(frac * a * d)
    Z = make_flexibility(np.array([[a, b], [b, d]]), La, Lb, c, s, normal)
    Zn, changed = oracle.t3_guard(Z, La, Lb, c, s, normal)
    got = orthogonal_block(Zn, La, Lb, c, s, normal)
    assert changed == (frac > 0.85)
    assert got[0, 0] == pytest.approx(a, rel=1e-12) and got[1, 1] == pytest.approx(d, rel=1e-9)
    expected_b = b if frac <= 0.85 else -np.sqrt(0.85 * a * d)
    assert got[0, 1] == pytest.approx(expected_b, rel=1e-9)


def rotate3(Q, angle):
    c, s = np.cos(angle), np.sin(angle)
    T = np.array([[c * c, s * s, 2 * c * s], [s * s, c * c, -2 * c * s], [-c * s, c * s, c * c - s * s]])
    Ti = np.linalg.inv(T)
    return Ti @ Q @ Ti.T


@pytest.mark.parametrize("ratio", [1.0, 2.0, 5.0])
@pytest.mark.parametrize("angle", [0.0, 0.43, 1.07])
@pytest.mark.parametrize("X", [np.array([[0.0, 0.0], [0.0042, 0.0], [0.803, 0.596]]),
                               np.array([[0.0, 0.0], [0.005, 0.0], [0.0, 1.0]])])
def test_t3_element_scaling_laws(oracle, ratio, angle, X):
    """utils/test/tria_bending_flexibility_guard_test.jl:21-62 on sliver triangles:
    symmetric, PSD, rigid motions in the null space, material units scale K,
    L' K(s X, s h) L = s^3 K (translations scale with length, rotations do not)."""
    E2, nu, h = 7e10, 0.3, 0.0025
    E1 = ratio * E2
    g12 = np.sqrt(E1 * E2) / (2 * (1 + nu))
    den = 1 - nu ** 2 * E2 / E1
    Q = np.array([[E1 / den, nu * E2 / den, 0.0], [nu * E2 / den, E2 / den, 0.0], [0.0, 0.0, g12]])
    Cs = np.array([[E1 / (2 * (1 + nu)), 0.0], [0.0, E2 / (2 * (1 + nu))]])
    Q = rotate3(Q, angle)
    c, s = np.cos(angle), np.sin(angle)
    Rm = np.array([[c, s], [-s, c]])
    Cs = Rm.T @ Cs @ Rm

    def kernel(Xs, t, ms=1.0):
        return oracle.t3_local(Xs, ms * Q * t, ms * Q * t ** 3 / 12, ms * Cs * 5 * t / 6, t, ms * E1, k6rot=100.0)

    K = kernel(X, h)
    nK = np.linalg.norm(K, 2)
    np.testing.assert_allclose(K, K.T, rtol=1e-13, atol=1e-13 * nK)
    assert np.linalg.eigvalsh(0.5 * (K + K.T)).min() >= -2e-12 * nK
    for axis in range(3):
        for rot in (False, True):
            d = np.zeros(3); d[axis] = 1
            u = np.zeros(18)
            for i in range(3):
                p = np.array([X[i, 0], X[i, 1], 0.0])
                u[6 * i:6 * i + 3] = np.cross(d, p) if rot else d
                if rot:
                    u[6 * i + 3:6 * i + 6] = d
            assert np.linalg.norm(K @ u) <= 2e-11 * nK * np.linalg.norm(u)
    np.testing.assert_allclose(kernel(X, h, 7.3), 7.3 * K, rtol=2e-9, atol=2e-9 * nK)
    sc = 9.0
    Ld = np.diag(np.tile([sc, sc, sc, 1.0, 1.0, 1.0], 3))
    np.testing.assert_allclose(Ld.T @ kernel(sc * X, sc * h) @ Ld, sc ** 3 * K, rtol=2e-9, atol=2e-9 * sc ** 3 * nK)


# ---------------------------------------------------------------- CQUAD4 Zb
CB = np.array([[38.0, 9.0, -3.0], [9.0, 119.0, 2.0], [-3.0, 2.0, 24.0]])
CS = np.array([[8.1e6, 0.27e6], [0.27e6, 9.4e6]])


@pytest.mark.parametrize("g", range(10))
def test_q4_zb_zs_scaling(oracle, g):
    """Zb is homogeneous of degree -1 in Cb, Zs in Cs; a thickness scaling s
    (Cb ~ s^3, Cs ~ s) gives Z_total = Zb/s^3 + Zs/s
    (natural_shear_work_transport_test.jl:146-157)."""
    xy = macneal_geometries()[g]
    base = oracle.q4_macneal_parts(xy, CB, CS)
    np.testing.assert_allclose(oracle.q4_macneal_parts(xy, 2.5 * CB, CS)["Zb"], base["Zb"] / 2.5, rtol=1e-13,
                               atol=1e-20)
    np.testing.assert_allclose(oracle.q4_macneal_parts(xy, CB, 2.5 * CS)["Zs"], base["Zs"] / 2.5, rtol=1e-12,
                               atol=1e-20)
    sc = 1.7
    sZ = oracle.q4_macneal_parts(xy, sc ** 3 * CB, sc * CS)["Z"]
    np.testing.assert_allclose(sZ, base["Zb"] / sc ** 3 + base["Zs"] / sc, rtol=1e-12, atol=1e-20)


@pytest.mark.parametrize("g", range(10))
def test_q4_cyclic_relabelling_preserves_shear_energy(oracle, g):
    """A cyclic change of the corners with the matching 90-degree rotation of the
    coordinates, of Cs and (with the bending axes swapped) of Cb preserves the
    elemental shear energies of Zs and of Z_total
    (natural_shear_work_transport_test.jl:158-166)."""
    xy = macneal_geometries()[g]
    R = np.array([[0.0, 1.0], [-1.0, 0.0]])
    perm = [1, 2, 3, 0]
    rxy = xy[perm] @ R.T
    rCs = R @ CS @ R.T
    B = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    base = oracle.q4_macneal_parts(xy, CB, CS)
    rot = oracle.q4_macneal_parts(rxy, B @ CB @ B.T, rCs)
    D, Dr = tying_rows(xy), tying_rows(rxy)
    v = np.array([0.1, 0.3, -0.2, -0.4, 0.7, 0.2, 0.9, -0.6, 0.5, 0.3, 0.2, -0.8])
    vr = np.concatenate([np.concatenate([[v[3 * j]], R @ v[3 * j + 1:3 * j + 3]]) for j in perm])
    for key in ("Zs", "Z"):
        e0 = (D @ v) @ np.linalg.solve(base[key], D @ v)
        e1 = (Dr @ vr) @ np.linalg.solve(rot[key], Dr @ vr)
        assert e1 == pytest.approx(e0, rel=4e-11), key


def test_q4_zero_bending_and_rigid_shear_limits(oracle):
    for xy in macneal_geometries():
        D = tying_rows(xy)
        nb = oracle.q4_macneal_parts(xy, np.zeros((3, 3)), CS)
        K0 = D.T @ np.linalg.solve(nb["Z"], D)
        assert np.all(np.isfinite(K0)) and np.linalg.norm(K0) < 1e-20
        rg = oracle.q4_macneal_parts(xy, CB, CS, rigid=True)
        assert not np.any(rg["Zs"])
        np.testing.assert_allclose(rg["Z"], oracle.q4_macneal_parts(xy, CB, CS)["Zb"], rtol=3e-13, atol=1e-18)
