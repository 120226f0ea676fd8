/*
 * jfem_oracle.c -- CPU restatement of OpenJFEM's SOL 101 hot path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Each function names the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic follows the reference's statement
 * order "as written" (accumulation order of ts_mul!/ts_mul_At_add!), built with
 * -ffp-contract=off.  Defaults are those of an unset JFEM_* environment.
 */
#include "jfem_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <stdint.h>

#define M(A, i, j, ld) (A)[(i) + (ld) * (j)]
static const double EPS = 2.220446049250313e-16; /* eps(Float64) */

void or_free(void *p) { free(p); }

/* ------------------------------------------------------------------------- */
/* small dense helpers                                                        */
/* ------------------------------------------------------------------------- */

/* shape_derivs_quad, src/FEMKernels.jl:762-766 */
static void shape_derivs_quad(double r, double s, double dNr[4], double dNs[4]) {
    dNr[0] = -0.25 * (1 - s); dNr[1] = 0.25 * (1 - s);
    dNr[2] = 0.25 * (1 + s);  dNr[3] = -0.25 * (1 + s);
    dNs[0] = -0.25 * (1 - r); dNs[1] = -0.25 * (1 + r);
    dNs[2] = 0.25 * (1 + r);  dNs[3] = 0.25 * (1 - r);
}

static void jac4(const double dNr[4], const double dNs[4], const double lc[8],
                 double *J11, double *J12, double *J21, double *J22) {
    /* coords[k,1] = lc[2k], coords[k,2] = lc[2k+1] */
    *J11 = dNr[0] * lc[0] + dNr[1] * lc[2] + dNr[2] * lc[4] + dNr[3] * lc[6];
    *J12 = dNr[0] * lc[1] + dNr[1] * lc[3] + dNr[2] * lc[5] + dNr[3] * lc[7];
    *J21 = dNs[0] * lc[0] + dNs[1] * lc[2] + dNs[2] * lc[4] + dNs[3] * lc[6];
    *J22 = dNs[0] * lc[1] + dNs[1] * lc[3] + dNs[2] * lc[5] + dNs[3] * lc[7];
}

/* _quad4_jacobian_measure, src/FEMKernels.jl:2862-2867 */
static int jac_measure(double detj, double j11, double j12, double j21, double j22, double *out) {
    double scale = fabs(j11 * j22) + fabs(j12 * j21);
    if (!(isfinite(detj) && isfinite(scale) && fabs(detj) > 64 * EPS * scale)) return OR_ERR_DEGENERATE;
    *out = fabs(detj);
    return OR_OK;
}

/* C = A*B, ts_mul! order (src/FEMKernels.jl:664-675): C[i,j] += A[i,l]*B[l,j], l ascending */
static void ts_mul(double *C, const double *A, const double *B, int m, int k, int n) {
    for (int q = 0; q < m * n; ++q) C[q] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int l = 0; l < k; ++l) {
            double v = B[l + k * j];
            for (int i = 0; i < m; ++i) C[i + m * j] += A[i + m * l] * v;
        }
}
/* C += alpha*A'*B, ts_mul_At_add! order (src/FEMKernels.jl:649-661); A is k x m */
static void ts_mul_At_add(double *C, const double *A, const double *B, double alpha,
                          int k, int m, int n) {
    for (int j = 0; j < n; ++j)
        for (int l = 0; l < k; ++l) {
            double v = alpha * B[l + k * j];
            for (int i = 0; i < m; ++i) C[i + m * j] += A[l + k * i] * v;
        }
}

/* 4x4 inverse by cofactors (stands in for StaticArrays inv(SMatrix{4,4})) */
static int inv4(const double *a, double *out) {
    double inv[16];
    inv[0] = a[5]*a[10]*a[15] - a[5]*a[11]*a[14] - a[9]*a[6]*a[15] + a[9]*a[7]*a[14] + a[13]*a[6]*a[11] - a[13]*a[7]*a[10];
    inv[4] = -a[4]*a[10]*a[15] + a[4]*a[11]*a[14] + a[8]*a[6]*a[15] - a[8]*a[7]*a[14] - a[12]*a[6]*a[11] + a[12]*a[7]*a[10];
    inv[8] = a[4]*a[9]*a[15] - a[4]*a[11]*a[13] - a[8]*a[5]*a[15] + a[8]*a[7]*a[13] + a[12]*a[5]*a[11] - a[12]*a[7]*a[9];
    inv[12] = -a[4]*a[9]*a[14] + a[4]*a[10]*a[13] + a[8]*a[5]*a[14] - a[8]*a[6]*a[13] - a[12]*a[5]*a[10] + a[12]*a[6]*a[9];
    inv[1] = -a[1]*a[10]*a[15] + a[1]*a[11]*a[14] + a[9]*a[2]*a[15] - a[9]*a[3]*a[14] - a[13]*a[2]*a[11] + a[13]*a[3]*a[10];
    inv[5] = a[0]*a[10]*a[15] - a[0]*a[11]*a[14] - a[8]*a[2]*a[15] + a[8]*a[3]*a[14] + a[12]*a[2]*a[11] - a[12]*a[3]*a[10];
    inv[9] = -a[0]*a[9]*a[15] + a[0]*a[11]*a[13] + a[8]*a[1]*a[15] - a[8]*a[3]*a[13] - a[12]*a[1]*a[11] + a[12]*a[3]*a[9];
    inv[13] = a[0]*a[9]*a[14] - a[0]*a[10]*a[13] - a[8]*a[1]*a[14] + a[8]*a[2]*a[13] + a[12]*a[1]*a[10] - a[12]*a[2]*a[9];
    inv[2] = a[1]*a[6]*a[15] - a[1]*a[7]*a[14] - a[5]*a[2]*a[15] + a[5]*a[3]*a[14] + a[13]*a[2]*a[7] - a[13]*a[3]*a[6];
    inv[6] = -a[0]*a[6]*a[15] + a[0]*a[7]*a[14] + a[4]*a[2]*a[15] - a[4]*a[3]*a[14] - a[12]*a[2]*a[7] + a[12]*a[3]*a[6];
    inv[10] = a[0]*a[5]*a[15] - a[0]*a[7]*a[13] - a[4]*a[1]*a[15] + a[4]*a[3]*a[13] + a[12]*a[1]*a[7] - a[12]*a[3]*a[5];
    inv[14] = -a[0]*a[5]*a[14] + a[0]*a[6]*a[13] + a[4]*a[1]*a[14] - a[4]*a[2]*a[13] - a[12]*a[1]*a[6] + a[12]*a[2]*a[5];
    inv[3] = -a[1]*a[6]*a[11] + a[1]*a[7]*a[10] + a[5]*a[2]*a[11] - a[5]*a[3]*a[10] - a[9]*a[2]*a[7] + a[9]*a[3]*a[6];
    inv[7] = a[0]*a[6]*a[11] - a[0]*a[7]*a[10] - a[4]*a[2]*a[11] + a[4]*a[3]*a[10] + a[8]*a[2]*a[7] - a[8]*a[3]*a[6];
    inv[11] = -a[0]*a[5]*a[11] + a[0]*a[7]*a[9] + a[4]*a[1]*a[11] - a[4]*a[3]*a[9] - a[8]*a[1]*a[7] + a[8]*a[3]*a[5];
    inv[15] = a[0]*a[5]*a[10] - a[0]*a[6]*a[9] - a[4]*a[1]*a[10] + a[4]*a[2]*a[9] + a[8]*a[1]*a[6] - a[8]*a[2]*a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    if (det == 0.0 || !isfinite(det)) return OR_ERR_DEGENERATE;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] / det;
    return OR_OK;
}

/* LU with partial pivoting, solve A X = B (n x n, nrhs columns); stands in for the
 * LAPACK getrf/getrs replicas of src/FEMKernels.jl:450-561 */
static int lu_solve(const double *Ain, int n, double *B, int nrhs) {
    double A[64]; int piv[8];
    memcpy(A, Ain, sizeof(double) * n * n);
    for (int k = 0; k < n; ++k) {
        int p = k; double mx = fabs(M(A, k, k, n));
        for (int i = k + 1; i < n; ++i) if (fabs(M(A, i, k, n)) > mx) { mx = fabs(M(A, i, k, n)); p = i; }
        piv[k] = p;
        if (mx == 0.0) return OR_ERR_DEGENERATE;
        if (p != k) for (int j = 0; j < n; ++j) { double t = M(A, k, j, n); M(A, k, j, n) = M(A, p, j, n); M(A, p, j, n) = t; }
        for (int i = k + 1; i < n; ++i) {
            M(A, i, k, n) /= M(A, k, k, n);
            for (int j = k + 1; j < n; ++j) M(A, i, j, n) -= M(A, i, k, n) * M(A, k, j, n);
        }
    }
    for (int c = 0; c < nrhs; ++c) {
        double *b = B + n * c;
        for (int k = 0; k < n; ++k) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
        for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) b[i] -= M(A, i, j, n) * b[j];
        for (int i = n - 1; i >= 0; --i) { for (int j = i + 1; j < n; ++j) b[i] -= M(A, i, j, n) * b[j]; b[i] /= M(A, i, i, n); }
    }
    return OR_OK;
}

/* symmetric 3x3 inverse via adjugate (stands in for inv(Symmetric(...))) */
static int inv3(const double *a, double *o) {
    double c00 = a[4] * a[8] - a[7] * a[5];
    double c01 = a[7] * a[2] - a[1] * a[8];
    double c02 = a[1] * a[5] - a[4] * a[2];
    double det = a[0] * c00 + a[3] * c01 + a[6] * c02;
    if (det == 0.0 || !isfinite(det)) return OR_ERR_DEGENERATE;
    o[0] = c00 / det; o[3] = (a[6] * a[5] - a[3] * a[8]) / det; o[6] = (a[3] * a[7] - a[6] * a[4]) / det;
    o[1] = c01 / det; o[4] = (a[0] * a[8] - a[6] * a[2]) / det; o[7] = (a[6] * a[1] - a[0] * a[7]) / det;
    o[2] = c02 / det; o[5] = (a[3] * a[2] - a[0] * a[5]) / det; o[8] = (a[0] * a[4] - a[3] * a[1]) / det;
    return OR_OK;
}

static double maxabs(const double *a, int n) { double m = 0; for (int i = 0; i < n; ++i) if (fabs(a[i]) > m) m = fabs(a[i]); return m; }

/* ------------------------------------------------------------------------- */
/* CQUAD4: MacNeal transverse shear                                           */
/* add_quad4_macneal_shear_rbf! default slice, src/FEMKernels.jl:4481-4834,   */
/* 6188-6285; quad4_natural_shear_flexibility!, :4223-4352                    */
/* ------------------------------------------------------------------------- */
int or_q4_macneal_parts(const double lc[8], const double Cb[9], const double Cs[4],
                        int rigid, double D[48], double Zb[16], double Zs[16],
                        double Z[16], double Kp[144]) {
    const double pt = 1.0 / sqrt(3.0);
    /* sample points (xi, eta, comp) :4509; edge rows at (0,-+1), (-+1,0) :4688-4712 */
    const double sxi[4] = {0.0, 0.0, -pt, pt}, seta[4] = {-pt, pt, 0.0, 0.0};
    const int comp[4] = {1, 1, 2, 2};
    double Ce[4 * 12], Je[4][4];
    for (int sp = 0; sp < 4; ++sp) {
        double xi = sxi[sp], eta = seta[sp];
        double xr = xi == 0.0 ? 0.0 : copysign(1.0, xi);
        double er = eta == 0.0 ? 0.0 : copysign(1.0, eta);
        double dNre[4], dNse[4], J11e, J12e, J21e, J22e;
        shape_derivs_quad(xr, er, dNre, dNse);
        jac4(dNre, dNse, lc, &J11e, &J12e, &J21e, &J22e);
        double Ne[4] = {0.25 * (1 - xr) * (1 - er), 0.25 * (1 + xr) * (1 - er),
                        0.25 * (1 + xr) * (1 + er), 0.25 * (1 - xr) * (1 + er)};
        Je[sp][0] = J11e; Je[sp][1] = J12e; Je[sp][2] = J21e; Je[sp][3] = J22e;
        double txe = comp[sp] == 1 ? J11e : J21e, tye = comp[sp] == 1 ? J12e : J22e;
        for (int k = 0; k < 4; ++k) {
            int col = 3 * k;
            M(Ce, sp, col + 0, 4) = comp[sp] == 1 ? dNre[k] : dNse[k];
            M(Ce, sp, col + 1, 4) = -Ne[k] * tye;
            M(Ce, sp, col + 2, 4) = Ne[k] * txe;
        }
    }
    /* interpolate covariant rows and map with the edge-point inverse Jacobian :4758-4797 */
    for (int sp = 0; sp < 4; ++sp) {
        double xi = sxi[sp], eta = seta[sp];
        double J11 = Je[sp][0], J12 = Je[sp][1], J21 = Je[sp][2], J22 = Je[sp][3];
        double detJ = J11 * J22 - J12 * J21;
        double adet = fabs(detJ) < 1e-14 ? (detJ < 0 ? -1e-14 : 1e-14) : detJ;
        double i11 = J22 / adet, i12 = -J12 / adet, i21 = -J21 / adet, i22 = J11 / adet;
        double w1 = 0.5 * (1.0 - eta / pt), w2 = 0.5 * (1.0 + eta / pt);
        double w3 = 0.5 * (1.0 - xi / pt), w4 = 0.5 * (1.0 + xi / pt);
        for (int j = 0; j < 12; ++j) {
            double gxie = w1 * M(Ce, 0, j, 4) + w2 * M(Ce, 1, j, 4);
            double getae = w3 * M(Ce, 2, j, 4) + w4 * M(Ce, 3, j, 4);
            M(D, sp, j, 4) = sp < 2 ? (i11 * gxie + i12 * getae) : (i21 * gxie + i22 * getae);
        }
    }
    /* natural flexibility */
    const double rn[4] = {-1.0, 1.0, 1.0, -1.0}, sn[4] = {-1.0, -1.0, 1.0, 1.0};
    double tr[16] = {0};
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) {
            int base = 3 * k;
            M(tr, i, 0, 4) += M(D, i, base + 2, 4);
            M(tr, i, 1, 4) += sn[k] * M(D, i, base + 2, 4);
            M(tr, i, 2, 4) -= M(D, i, base + 1, 4);
            M(tr, i, 3, 4) -= rn[k] * M(D, i, base + 1, 4);
        }
    const double x1 = lc[0], x2 = lc[2], x3 = lc[4], x4 = lc[6];
    const double y1 = lc[1], y2 = lc[3], y3 = lc[5], y4 = lc[7];
    double ax = 0.25 * (-x1 + x2 + x3 - x4), ay = 0.25 * (-y1 + y2 + y3 - y4);
    double bx = 0.25 * (-x1 - x2 + x3 + x4), by = 0.25 * (-y1 - y2 + y3 + y4);
    double cx = 0.25 * (x1 - x2 + x3 - x4), cy = 0.25 * (y1 - y2 + y3 - y4);
    double det0 = ax * by - ay * bx;
    double det_scale = hypot(ax, ay) * hypot(bx, by);
    if (det_scale < EPS) det_scale = EPS;
    if (!(fabs(det0) > 64 * EPS * det_scale)) return OR_ERR_DEGENERATE;
    double orient = det0 > 0 ? 1.0 : (det0 < 0 ? -1.0 : 0.0);
    const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
    for (int q = 0; q < 4; ++q) {
        double r = cr[q], s = cs_[q];
        double dj = (ax + cx * s) * (by + cy * r) - (ay + cy * s) * (bx + cx * r);
        if (!(orient * dj > 64 * EPS * det_scale)) return OR_ERR_DEGENERATE;
    }
    double area = 4 * fabs(det0);
    double dx2 = (2 * ax) * (2 * ax), dy2 = (2 * by) * (2 * by);
    double dxr = M(Cb, 0, 0, 3), dyr = M(Cb, 1, 1, 3);
    double fx = fabs(2 * ax) > 1e-30 ? (cx / (2 * ax)) * (cx / (2 * ax)) : 0.0;
    double fy = fabs(2 * by) > 1e-30 ? (cy / (2 * by)) * (cy / (2 * by)) : 0.0;
    if (fx + fy > 1e-24 && ((dxr == 0.0) != (dyr == 0.0))) return OR_ERR_BAD_ARG;
    double ux = dx2 / ((12 * area) * fmax(fabs(dxr), 1e-30));
    double uy = dy2 / ((12 * area) * fmax(fabs(dyr), 1e-30));
    double denx = dy2 + 39 * dx2, deny = dx2 + 39 * dy2;
    if (!(denx > 0 && deny > 0)) return OR_ERR_DEGENERATE;
    double zb[4] = {ux, ((3 * dy2) / denx) * ux, uy, ((3 * dx2) / deny) * uy};
    for (int q = 0; q < 16; ++q) { Zb[q] = 0; Zs[q] = 0; }
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) M(Zb, i, j, 4) += (M(tr, i, k, 4) * zb[k]) * M(tr, j, k, 4);
    if (!rigid) {
        double gp = 1.0 / sqrt(3.0);
#define JAC(r, s) fabs((ax + cx * (s)) * (by + cy * (r)) - (ay + cy * (s)) * (bx + cx * (r)))
        double va = sqrt(2 * JAC(0.0, -gp)), vb = sqrt(2 * JAC(0.0, gp));
        double vc = sqrt(2 * JAC(-gp, 0.0)), vd = sqrt(2 * JAC(gp, 0.0));
#undef JAC
        double pxa[4] = {va, -gp * va, 0, 0}, pxb[4] = {vb, gp * vb, 0, 0};
        double pyc[4] = {0, 0, vc, -gp * vc}, pyd[4] = {0, 0, vd, gp * vd};
        double H[16], Hi[16];
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i)
                M(H, i, j, 4) = M(Cs, 0, 0, 2) * (pxa[i] * pxa[j] + pxb[i] * pxb[j]) +
                                M(Cs, 1, 1, 2) * (pyc[i] * pyc[j] + pyd[i] * pyd[j]) +
                                M(Cs, 0, 1, 2) * (pxa[i] + pxb[i]) * (pyc[j] + pyd[j]) / 2 +
                                M(Cs, 1, 0, 2) * (pyc[i] + pyd[i]) * (pxa[j] + pxb[j]) / 2;
        for (int q = 0; q < 16; ++q) Hi[q] = (q % 5 == 0) ? 1.0 : 0.0;
        if (lu_solve(H, 4, Hi, 4) != OR_OK) return OR_ERR_DEGENERATE;
        double t4[16];
        ts_mul(t4, tr, Hi, 4, 4, 4);
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                double s = 0; for (int k = 0; k < 4; ++k) s += M(t4, i, k, 4) * M(tr, j, k, 4);
                M(Zs, i, j, 4) = s;
            }
    }
    for (int q = 0; q < 16; ++q) Z[q] = Zb[q] + Zs[q];
    /* K_plate = sym(D' * (Z \ D)) :6190-6200 */
    double sol[48];
    memcpy(sol, D, sizeof(sol));
    if (lu_solve(Z, 4, sol, 12) != OR_OK) return OR_ERR_DEGENERATE;
    double raw[144];
    for (int j = 0; j < 12; ++j)
        for (int i = 0; i < 12; ++i) {
            double s = 0; for (int k = 0; k < 4; ++k) s += M(D, k, i, 4) * M(sol, k, j, 4);
            M(raw, i, j, 12) = s;
        }
    for (int j = 0; j < 12; ++j)
        for (int i = 0; i < 12; ++i) M(Kp, i, j, 12) = 0.5 * (M(raw, i, j, 12) + M(raw, j, i, 12));
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* CQUAD4 core, _stiffness_quad4_core! default branch set,                    */
/* src/FEMKernels.jl:2874-4160 (see the per-step citations below)             */
/* ------------------------------------------------------------------------- */
/* inverse of K[active, active] scattered into a 4x4 (zeros elsewhere):
 * q4_internal_mode_compliance with binary weights, src/FEMKernels.jl:1918-1955 */
static int active_inverse(const double *K, int mask, double *C) {
    int act[4], na = 0;
    for (int i = 0; i < 4; ++i) if (mask & (1 << i)) act[na++] = i;
    for (int q = 0; q < 16; ++q) C[q] = 0.0;
    if (na == 0) return OR_OK;
    if (na == 4) return inv4(K, C);
    double sub[16], inv[16];
    for (int j = 0; j < na; ++j) for (int i = 0; i < na; ++i) sub[i + na * j] = M(K, act[i], act[j], 4);
    for (int j = 0; j < na; ++j) for (int i = 0; i < na; ++i) inv[i + na * j] = (i == j) ? 1.0 : 0.0;
    if (lu_solve(sub, na, inv, na) != OR_OK) return OR_ERR_DEGENERATE;
    for (int j = 0; j < na; ++j) for (int i = 0; i < na; ++i) M(C, act[i], act[j], 4) = inv[i + na * j];
    return OR_OK;
}

int or_q4_local(const double lc[8], const double Cm[9], const double Cb[9],
                const double Cs[4], double h, double E_ref, double k6rot,
                double drill_scale, int rigid, double Ke[576]) {
    return or_q4_local_modes(lc, Cm, Cb, Cs, h, E_ref, k6rot, drill_scale, rigid, 15, Ke);
}

int or_q4_local_modes(const double lc[8], const double Cm[9], const double Cb[9],
                      const double Cs[4], double h, double E_ref, double k6rot,
                      double drill_scale, int rigid, int membrane_modes, double Ke[576]) {
    (void)E_ref;
    double K_ab[96] = {0}, K_bb[16] = {0}, K_abb[96] = {0}, K_bbb[16] = {0};
    double K_abs[96] = {0}, K_bbs[16] = {0};
    for (int q = 0; q < 576; ++q) Ke[q] = 0.0;
    /* centre Jacobian :3136-3163 */
    const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac4(dNr_c, dNs_c, lc, &J11c, &J12c, &J21c, &J22c);
    double detJc = J11c * J22c - J12c * J21c, abs_detJc = fabs(detJc);
    double inv_detc = 1.0 / detJc;
    double iJ11c = J22c * inv_detc, iJ12c = -J12c * inv_detc, iJ21c = -J21c * inv_detc, iJ22c = J11c * inv_detc;
    const double pt = 1.0 / sqrt(3.0);
    const double gr[4] = {-pt, pt, pt, -pt}, gs[4] = {-pt, -pt, pt, pt};
    double Bm[72], Bb[72], Bi[12], Bib[12], tmp[72], t34[12];
    const int shape_cm = maxabs(Cm, 9) > 0;
    for (int gp = 0; gp < 4; ++gp) {  /* :3230-3621 */
        double r = gr[gp], s = gs[gp];
        double dNr[4], dNs[4], J11, J12, J21, J22, w;
        shape_derivs_quad(r, s, dNr, dNs);
        jac4(dNr, dNs, lc, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        if (jac_measure(detJ, J11, J12, J21, J22, &w) != OR_OK) return OR_ERR_DEGENERATE;
        double inv_det = 1.0 / detJ;
        double iJ11 = J22 * inv_det, iJ12 = -J12 * inv_det, iJ21 = -J21 * inv_det, iJ22 = J11 * inv_det;
        memset(Bm, 0, sizeof(Bm)); memset(Bb, 0, sizeof(Bb));
        for (int k = 0; k < 4; ++k) {  /* :3299-3352 */
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k];
            double dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            int idx = 6 * k;
            M(Bm, 0, idx + 0, 3) = dx; M(Bm, 1, idx + 1, 3) = dy;
            M(Bm, 2, idx + 0, 3) = dy; M(Bm, 2, idx + 1, 3) = dx;
            M(Bb, 0, idx + 4, 3) = dx; M(Bb, 1, idx + 3, 3) = -dy;
            M(Bb, 2, idx + 4, 3) = dy; M(Bb, 2, idx + 3, 3) = -dx;
        }
        /* membrane internal modes, patch-consistent Jacobian :1847-1896 */
        double ratio = (iJ11 * iJ22 - iJ12 * iJ21) / (iJ11c * iJ22c - iJ12c * iJ21c);
        double m11 = ratio * iJ11c, m12 = ratio * iJ12c, m21 = ratio * iJ21c, m22 = ratio * iJ22c;
        double p1x = m11 * (-2.0 * r), p1y = m21 * (-2.0 * r), p2x = m12 * (-2.0 * s), p2y = m22 * (-2.0 * s);
        memset(Bi, 0, sizeof(Bi));
        M(Bi, 0, 0, 3) = p1x; M(Bi, 2, 0, 3) = p1y; M(Bi, 1, 1, 3) = p1y; M(Bi, 2, 1, 3) = p1x;
        M(Bi, 0, 2, 3) = p2x; M(Bi, 2, 2, 3) = p2y; M(Bi, 1, 3, 3) = p2y; M(Bi, 2, 3, 3) = p2x;
        /* bending internal modes, qm6 Jacobian :3478-3488 */
        double g = detJc / detJ;
        double b1x = (iJ11c * (-2.0 * r)) * g, b1y = (iJ21c * (-2.0 * r)) * g;
        double b2x = (iJ12c * (-2.0 * s)) * g, b2y = (iJ22c * (-2.0 * s)) * g;
        memset(Bib, 0, sizeof(Bib));
        M(Bib, 1, 0, 3) = -b1y; M(Bib, 1, 1, 3) = -b2y;
        M(Bib, 0, 2, 3) = b1x;  M(Bib, 0, 3, 3) = b2x;
        M(Bib, 2, 0, 3) = -b1x; M(Bib, 2, 1, 3) = -b2x;
        M(Bib, 2, 2, 3) = b1y;  M(Bib, 2, 3, 3) = b2y;
        /* accumulations :3528-3610 */
        ts_mul(tmp, Cm, Bm, 3, 3, 24); ts_mul_At_add(Ke, Bm, tmp, w, 3, 24, 24);
        ts_mul(tmp, Cb, Bb, 3, 3, 24); ts_mul_At_add(Ke, Bb, tmp, w, 3, 24, 24);
        ts_mul(t34, Cm, Bi, 3, 3, 4);
        ts_mul_At_add(K_ab, Bm, t34, w, 3, 24, 4);
        ts_mul_At_add(K_bb, Bi, t34, w, 3, 4, 4);
        ts_mul(t34, shape_cm ? Cm : Cb, Bib, 3, 3, 4);
        ts_mul_At_add(K_abs, Bb, t34, w, 3, 24, 4);
        ts_mul_At_add(K_bbs, Bib, t34, w, 3, 4, 4);
        ts_mul(t34, Cb, Bib, 3, 3, 4);
        ts_mul_At_add(K_abb, Bb, t34, w, 3, 24, 4);
        ts_mul_At_add(K_bbb, Bib, t34, w, 3, 4, 4);
    }
    /* MacNeal shear into plate DOFs :3737-3747, skip rule :4496-4498 */
    if (!(h < 1e-30 || (!rigid && maxabs(Cs, 4) < 1e-30))) {
        double D[48], Zb[16], Zs[16], Z[16], Kp[144];
        int st = or_q4_macneal_parts(lc, Cb, Cs, rigid, D, Zb, Zs, Z, Kp);
        if (st != OR_OK) return st;
        static const int pd[12] = {2, 3, 4, 8, 9, 10, 14, 15, 16, 20, 21, 22};
        for (int j = 0; j < 12; ++j)
            for (int i = 0; i < 12; ++i) M(Ke, pd[i], pd[j], 24) += M(Kp, i, j, 12);
    }
    /* lumped K6ROT drilling with corner Jacobian weights :3772-3817 */
    {
        double A_drill = 4.0 * abs_detJc;
        double alpha_L = ((drill_scale * (k6rot * 1e-6)) * M(Cm, 2, 2, 3)) * A_drill;
        const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
        double Bd[24];
        for (int i = 0; i < 4; ++i) {
            double dNr[4], dNs[4], J11, J12, J21, J22, wdum;
            shape_derivs_quad(cr[i], cs_[i], dNr, dNs);
            jac4(dNr, dNs, lc, &J11, &J12, &J21, &J22);
            double detJ = J11 * J22 - J12 * J21;
            if (jac_measure(detJ, J11, J12, J21, J22, &wdum) != OR_OK) return OR_ERR_DEGENERATE;
            double inv_det = 1.0 / detJ;
            double iJ11 = J22 * inv_det, iJ12 = -J12 * inv_det, iJ21 = -J21 * inv_det, iJ22 = J11 * inv_det;
            memset(Bd, 0, sizeof(Bd));
            for (int k = 0; k < 4; ++k) {
                double dx = iJ11 * dNr[k] + iJ12 * dNs[k];
                double dy = iJ21 * dNr[k] + iJ22 * dNs[k];
                Bd[6 * k + 0] = 0.5 * dy; Bd[6 * k + 1] = -0.5 * dx;
            }
            Bd[6 * i + 5] = 1.0;
            double wc = fabs(detJ) / fmax(abs_detJc, 1e-30);
            ts_mul_At_add(Ke, Bd, Bd, alpha_L * wc, 1, 24, 24);
        }
    }
    /* section kinematic enrichment :1989-2024 (binary unit weights :1939-1967) */
    double Pm[96] = {0}, Pb[96] = {0};
    if (maxabs(K_bb, 16) > 0) {
        double C[16];
        if (active_inverse(K_bb, membrane_modes, C) != OR_OK) return OR_ERR_DEGENERATE;
        for (int j = 0; j < 24; ++j) for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k)
            M(Pm, i, j, 4) -= M(C, i, k, 4) * M(K_ab, j, k, 24);
    }
    if (maxabs(K_bbs, 16) > 0) {
        double C[16];
        if (inv4(K_bbs, C) != OR_OK) return OR_ERR_DEGENERATE;
        for (int j = 0; j < 24; ++j) for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k)
            M(Pb, i, j, 4) -= M(C, i, k, 4) * M(K_abs, j, k, 24);
    }
    for (int j = 0; j < 24; ++j)
        for (int i = 0; i < 24; ++i) {
            double v = 0.0;
            for (int k = 0; k < 4; ++k) {
                double qim = M(K_ab, i, k, 24), qjm = M(K_ab, j, k, 24);
                double qib = M(K_abb, i, k, 24), qjb = M(K_abb, j, k, 24);
                v += qim * M(Pm, k, j, 4) + M(Pm, k, i, 4) * qjm + qib * M(Pb, k, j, 4) + M(Pb, k, i, 4) * qjb;
                for (int l = 0; l < 4; ++l)
                    v += M(Pm, k, i, 4) * M(K_bb, k, l, 4) * M(Pm, l, j, 4) +
                         M(Pb, k, i, 4) * M(K_bbb, k, l, 4) * M(Pb, l, j, 4);
            }
            M(Ke, i, j, 24) += v;
        }
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* shell frames, constitutive, transform                                      */
/* ------------------------------------------------------------------------- */
static void v3sub(const double *a, const double *b, double *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static double v3dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void v3cross(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
}
static void v3normalize(double *a) { double n = sqrt(v3dot(a, a)); a[0] /= n; a[1] /= n; a[2] /= n; }

/* shell_element_frame_quad4 (diag mode), src/solver/helpers.jl:1422-1481 */
static void frame_quad4(const double *p1, const double *p2, const double *p3, const double *p4,
                        double *v1, double *v2, double *v3) {
    double gxi[3], geta[3];
    for (int c = 0; c < 3; ++c) {
        gxi[c] = 0.25 * (((-p1[c]) + p2[c]) + p3[c] - p4[c]);
        geta[c] = 0.25 * (((-p1[c]) - p2[c]) + p3[c] + p4[c]);
    }
    double vr[3];
    v3cross(gxi, geta, vr);
    double d13[3], d24[3];
    v3sub(p3, p1, d13); v3sub(p4, p2, d24);
    if (v3dot(vr, vr) <= 1e-24) v3cross(d13, d24, vr);
    v3[0] = vr[0]; v3[1] = vr[1]; v3[2] = vr[2]; v3normalize(v3);
    double d13n[3] = {d13[0], d13[1], d13[2]}, d24n[3] = {d24[0], d24[1], d24[2]};
    v3normalize(d13n); v3normalize(d24n);
    double xr[3] = {d13n[0] - d24n[0], d13n[1] - d24n[1], d13n[2] - d24n[2]};
    if (v3dot(xr, xr) < 1e-20) { xr[0] = d13n[0] + d24n[0]; xr[1] = d13n[1] + d24n[1]; xr[2] = d13n[2] + d24n[2]; }
    double dd = v3dot(xr, v3);
    double xp[3] = {xr[0] - dd * v3[0], xr[1] - dd * v3[1], xr[2] - dd * v3[2]};
    if (v3dot(xp, xp) <= 1e-24) {
        double de = v3dot(geta, v3);
        double xa[3] = {geta[0] - de * v3[0], geta[1] - de * v3[1], geta[2] - de * v3[2]};
        if (v3dot(xa, xa) <= 1e-24) {
            double e[3]; v3sub(p2, p1, e); double d2 = v3dot(e, v3);
            xa[0] = e[0] - d2 * v3[0]; xa[1] = e[1] - d2 * v3[1]; xa[2] = e[2] - d2 * v3[2];
        }
        xp[0] = xa[0]; xp[1] = xa[1]; xp[2] = xa[2];
    }
    v1[0] = xp[0]; v1[1] = xp[1]; v1[2] = xp[2]; v3normalize(v1);
    v3cross(v3, v1, v2);
}

/* isotropic PSHELL section, src/solver/assembly.jl:3380-3404, 2085-2129, 2174-2233 */
static void shell_section(const double prop[8], double Cm[9], double Cb[9], double Cs[4],
                          double *Gref, int *membrane_only) {
    double h = prop[OR_SP_H], E = prop[OR_SP_E], nu = prop[OR_SP_NU], G = prop[OR_SP_G];
    double br = prop[OR_SP_BR], tst = prop[OR_SP_TST];
    if (!(G > 0.0)) G = E > 0.0 ? E / (2 * (1 + nu)) : 0.0;
    double cm = (E * h) / (1 - nu * nu);
    double Mi[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
    for (int q = 0; q < 9; ++q) Cm[q] = cm * Mi[q];
    if (br > 1e-12) {
        double pc = E / fmax(1.0 - nu * nu, 1e-30);
        double Mb[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1.0 - nu) / 2.0};
        double f = ((br * (h * h * h)) / 12.0);
        for (int q = 0; q < 9; ++q) Cb[q] = (pc * Mb[q]) * f;
        double ts = tst * h;
        Cs[0] = ts * G; Cs[1] = 0.0; Cs[2] = 0.0; Cs[3] = ts * G;
        *membrane_only = 0;
    } else {
        for (int q = 0; q < 9; ++q) Cb[q] = 0.0;
        for (int q = 0; q < 4; ++q) Cs[q] = 0.0;
        *membrane_only = 1;
    }
    *Gref = G;
}

/* _rotate_constitutive_3x3!, src/solver/assembly.jl:2440-2462 (C <- T' C T) */
static void rot3(double *C, double T11, double T12, double T13, double T21, double T22, double T23,
                 double T31, double T32, double T33) {
    double c11 = C[0], c21 = C[1], c31 = C[2], c12 = C[3], c22 = C[4], c32 = C[5], c13 = C[6], c23 = C[7], c33 = C[8];
    double t11 = c11 * T11 + c12 * T21 + c13 * T31, t12 = c11 * T12 + c12 * T22 + c13 * T32, t13 = c11 * T13 + c12 * T23 + c13 * T33;
    double t21 = c21 * T11 + c22 * T21 + c23 * T31, t22 = c21 * T12 + c22 * T22 + c23 * T32, t23 = c21 * T13 + c22 * T23 + c23 * T33;
    double t31 = c31 * T11 + c32 * T21 + c33 * T31, t32 = c31 * T12 + c32 * T22 + c33 * T32, t33 = c31 * T13 + c32 * T23 + c33 * T33;
    C[0] = T11 * t11 + T21 * t21 + T31 * t31; C[3] = T11 * t12 + T21 * t22 + T31 * t32; C[6] = T11 * t13 + T21 * t23 + T31 * t33;
    C[1] = T12 * t11 + T22 * t21 + T32 * t31; C[4] = T12 * t12 + T22 * t22 + T32 * t32; C[7] = T12 * t13 + T22 * t23 + T32 * t33;
    C[2] = T13 * t11 + T23 * t21 + T33 * t31; C[5] = T13 * t12 + T23 * t22 + T33 * t32; C[8] = T13 * t13 + T23 * t23 + T33 * t33;
}

/* R' * blk * R for every 3x3 block (_quad4_transform_emit!, src/solver/assembly.jl:359-440) */
static void transform_blocks(const double *Kl, int n, const double R[9] /* rows v1,v2,v3: R[r*3+c] */, double *Kg) {
    int nb = n / 3;
    for (int B = 0; B < nb; ++B)
        for (int A = 0; A < nb; ++A) {
            double t[9];
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    double s = 0;
                    for (int l = 0; l < 3; ++l) s += M(Kl, 3 * A + i, 3 * B + l, n) * R[l * 3 + j];
                    t[i + 3 * j] = s;
                }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    double s = 0;
                    for (int l = 0; l < 3; ++l) s += R[l * 3 + i] * t[l + 3 * j];
                    M(Kg, 3 * A + i, 3 * B + j, n) = s;
                }
        }
}

/* quad4_local_z_from_coords3d + warp gate, src/FEMKernels.jl:853-865, 1100-1102 */
static int quad4_is_warped(const double *p, const double lc[8]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    double c[3], a[3], b[3], n[3];
    for (int k = 0; k < 3; ++k) c[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4;
    v3sub(p3, p1, a); v3sub(p4, p2, b); v3cross(a, b, n);
    double nr = sqrt(v3dot(n, n));
    double v3[3] = {0, 0, 1};
    if (nr > 1e-30) { v3[0] = n[0] / nr; v3[1] = n[1] / nr; v3[2] = n[2] / nr; }
    double zmax = 0, xmax = 0;
    for (int k = 0; k < 4; ++k) {
        double d[3]; v3sub(p + 3 * k, c, d);
        double z = fabs(v3dot(d, v3)); if (z > zmax) zmax = z;
        if (fabs(lc[2 * k]) > xmax) xmax = fabs(lc[2 * k]);
    }
    if (xmax < 1e-30) xmax = 1e-30;
    return zmax > 1e-12 * xmax;
}

/* local corner heights about the centroid along normalize(d13 x d24)
 * (quad4_local_z_from_coords3d, src/FEMKernels.jl:853-865) */
static void quad4_local_z(const double *p, double zl[4]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    double c[3], a[3], b[3], n[3];
    for (int k = 0; k < 3; ++k) c[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4;
    v3sub(p3, p1, a); v3sub(p4, p2, b); v3cross(a, b, n);
    double nr = sqrt(v3dot(n, n));
    double v3[3] = {0, 0, 1};
    if (nr > 1e-30) { v3[0] = n[0] / nr; v3[1] = n[1] / nr; v3[2] = n[2] / nr; }
    for (int k = 0; k < 4; ++k) { double d[3]; v3sub(p + 3 * k, c, d); zl[k] = v3dot(d, v3); }
}

/* the finite-warp displacement map Tg (24 x 24, column-major)
 * (quad4_finite_warp_displacement_map / _quad4_finite_warp_fill!, src/FEMKernels.jl:866-945) */
static void quad4_warp_map(const double lc[8], const double *p, double Tg[576]) {
    double zl[4];
    quad4_local_z(p, zl);
    for (int q = 0; q < 576; ++q) Tg[q] = (q % 25 == 0) ? 1.0 : 0.0;
    const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; ++i) {
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs_quad(cr[i], cs_[i], dNr, dNs);
        jac4(dNr, dNs, lc, &J11, &J12, &J21, &J22);
        double dj = J11 * J22 - J12 * J21;
        if (fabs(dj) < 1e-30) continue;
        double i11 = J22 / dj, i12 = -J12 / dj, i21 = -J21 / dj, i22 = J11 / dj;
        double zr = dNr[0] * zl[0] + dNr[1] * zl[1] + dNr[2] * zl[2] + dNr[3] * zl[3];
        double zs = dNs[0] * zl[0] + dNs[1] * zl[1] + dNs[2] * zl[2] + dNs[3] * zl[3];
        double gx = i11 * zr + i12 * zs, gy = i21 * zr + i22 * zs;
        int rxi = 6 * i + 3, ryi = 6 * i + 4;
        M(Tg, rxi, 6 * i + 5, 24) += gx;
        M(Tg, ryi, 6 * i + 5, 24) += gy;
        int a = (i % 2 == 0) ? 0 : 1, b = (i % 2 == 0) ? 2 : 3;
        double dx = lc[2 * b] - lc[2 * a], dy = lc[2 * b + 1] - lc[2 * a + 1];
        double dl2 = dx * dx + dy * dy;
        if (dl2 > 1e-30) {
            int js[2] = {a, b};
            double cu[2] = {dy / dl2, -dy / dl2}, cv[2] = {-dx / dl2, dx / dl2};
            for (int t = 0; t < 2; ++t) {
                int j = js[t];
                M(Tg, rxi, 6 * j + 0, 24) -= gx * cu[t];
                M(Tg, rxi, 6 * j + 1, 24) -= gx * cv[t];
                M(Tg, ryi, 6 * j + 0, 24) -= gy * cu[t];
                M(Tg, ryi, 6 * j + 1, 24) -= gy * cv[t];
            }
        }
        for (int j = 0; j < 4; ++j) {
            double dNdx = i11 * dNr[j] + i12 * dNs[j];
            double dNdy = i21 * dNr[j] + i22 * dNs[j];
            M(Tg, 6 * i + 0, 6 * j + 2, 24) += zl[i] * dNdx;
            M(Tg, 6 * i + 1, 6 * j + 2, 24) += zl[i] * dNdy;
        }
    }
}

/* finite-warp work-dual congruence Ke <- sym(Tg' Ke Tg)
 * (apply_quad4_finite_warp_equilibrium!, src/FEMKernels.jl:1082-1116) */
static void quad4_finite_warp(double *Ke, const double lc[8], const double *p) {
    double Tg[576];
    quad4_warp_map(lc, p, Tg);
    double W[576], Kt[576];
    for (int j = 0; j < 24; ++j) for (int i = 0; i < 24; ++i) {
        double v = 0; for (int m = 0; m < 24; ++m) v += M(Ke, i, m, 24) * M(Tg, m, j, 24); M(W, i, j, 24) = v; }
    for (int j = 0; j < 24; ++j) for (int i = 0; i < 24; ++i) {
        double v = 0; for (int l = 0; l < 24; ++l) v += M(Tg, l, i, 24) * M(W, l, j, 24); M(Kt, i, j, 24) = v; }
    for (int j = 0; j < 24; ++j) for (int i = 0; i < 24; ++i) M(Ke, i, j, 24) = 0.5 * (M(Kt, i, j, 24) + M(Kt, j, i, 24));
}

int or_q4_global(const double p[12], const double prop[8], double k6rot, double Ke[576]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    double v1[3], v2[3], v3[3];
    frame_quad4(p1, p2, p3, p4, v1, v2, v3);
    /* local coordinates about the centroid, src/solver/assembly.jl:3686-3691 */
    double c[3], lc[8];
    for (int k = 0; k < 3; ++k) c[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4.0;
    for (int k = 0; k < 4; ++k) {
        double d[3]; v3sub(p + 3 * k, c, d);
        lc[2 * k] = v3dot(d, v1); lc[2 * k + 1] = v3dot(d, v2);
    }
    int warped = quad4_is_warped(p, lc);
    double Cm[9], Cb[9], Cs[4], G; int mo;
    shell_section(prop, Cm, Cb, Cs, &G, &mo);
    /* PSHELL material axis from edge 1-2, helpers.jl:1749-1759, assembly.jl:4400-4425 */
    {
        double e[3]; v3sub(p2, p1, e);
        double de = v3dot(e, v3);
        double ep[3] = {e[0] - de * v3[0], e[1] - de * v3[1], e[2] - de * v3[2]};
        double l2 = v3dot(ep, ep);
        if (l2 > 1e-24) {
            double l = sqrt(l2); double xm[3] = {ep[0] / l, ep[1] / l, ep[2] / l};
            double beta = 0.0 + atan2(v3dot(xm, v2), v3dot(xm, v1));
            if (fabs(beta) > 1e-10) {
                double cb = cos(beta), sb = sin(beta);
                double c2 = cb * cb, s2 = sb * sb, cs = cb * sb;
                rot3(Cm, c2, s2, cs, s2, c2, -cs, -2 * cs, 2 * cs, c2 - s2);
                rot3(Cb, c2, s2, cs, s2, c2, -cs, -2 * cs, 2 * cs, c2 - s2);
                double a11 = Cs[0], a12 = Cs[2], a22 = Cs[3];
                Cs[0] = cb * cb * a11 - 2 * cb * sb * a12 + sb * sb * a22;
                Cs[2] = cb * sb * a11 + (cb * cb - sb * sb) * a12 - cb * sb * a22;
                Cs[1] = Cs[2];
                Cs[3] = sb * sb * a11 + 2 * cb * sb * a12 + cb * cb * a22;
            }
        }
    }
    int rigid = (prop[OR_SP_RIGID] != 0.0) && !mo;
    double k6 = mo ? 0.0 : k6rot, ds = mo ? 0.0 : 1.0;
    double Kl[576];
    int modes = (int)prop[OR_SP_MODES];
    if (modes <= 0 || modes > 15) modes = 15;
    int st = or_q4_local_modes(lc, Cm, Cb, Cs, prop[OR_SP_H], prop[OR_SP_E], k6, ds, rigid, modes, Kl);
    if (st != OR_OK) return st;
    if (warped) quad4_finite_warp(Kl, lc, p);
    double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    transform_blocks(Kl, 24, R, Ke);
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* CTRIA3, stiffness_tria3_matrices_generic default branches,                 */
/* src/FEMKernels.jl:9375-9809, guard :9281-9312                               */
/* ------------------------------------------------------------------------- */
static int tria3_guard(double *Zb, double La, double Lb, double c, double s, double Dn) {
    double scale = 6 * (La / Lb) * s * Dn;
    if (!(scale > 0 && M(Zb, 2, 2, 3) > 0)) return 0;
    double z33 = M(Zb, 2, 2, 3);
    double h11 = scale * (M(Zb, 0, 0, 3) - M(Zb, 0, 2, 3) * M(Zb, 0, 2, 3) / z33) / (La * La);
    double h22 = scale * (M(Zb, 1, 1, 3) - M(Zb, 1, 2, 3) * M(Zb, 1, 2, 3) / z33) / (Lb * Lb);
    double h12 = scale * (M(Zb, 0, 1, 3) - M(Zb, 0, 2, 3) * M(Zb, 1, 2, 3) / z33) / (La * Lb);
    double p11 = h11, p12 = (h12 - c * h11) / s, p22 = (h22 - 2 * c * h12 + c * c * h11) / (s * s);
    double a = fmax(p11, 1.0 / 10000), q = fmax(p22, a / 10000);
    double limit = sqrt((85.0 / 100) * a * q);
    double b = p12 < -limit ? -limit : (p12 > limit ? limit : p12);
    if (a == p11 && b == p12 && q == p22) return 0;
    double n11 = a, n12 = c * a + s * b, n22 = c * c * a + 2 * c * s * b + s * s * q;
    M(Zb, 0, 0, 3) += (n11 - h11) * La * La / scale;
    M(Zb, 0, 1, 3) += (n12 - h12) * La * Lb / scale;
    M(Zb, 1, 0, 3) = M(Zb, 0, 1, 3);
    M(Zb, 1, 1, 3) += (n22 - h22) * Lb * Lb / scale;
    return 1;
}

int or_t3_guard(double Zb[9], double La, double Lb, double c, double s, double Dn) {
    double scale = 6 * (La / Lb) * s * Dn;
    if (scale > 0 && Zb[8] > 0) {
        double z33 = Zb[8];
        double h11 = scale * (Zb[0] - Zb[6] * Zb[6] / z33) / (La * La);
        double h22 = scale * (Zb[4] - Zb[7] * Zb[7] / z33) / (Lb * Lb);
        double h12 = scale * (Zb[3] - Zb[6] * Zb[7] / z33) / (La * Lb);
        double p22 = (h22 - 2 * c * h12 + c * c * h11) / (s * s);
        if (!(isfinite(p22) && p22 > 0)) return -1;   /* ArgumentError, :9297-9298 */
    }
    return tria3_guard(Zb, La, Lb, c, s, Dn);
}

/* MITC3 shear of the CTRIA3 plate part: tying rows D3 (grA, gsB, cvec; 9 entries
 * each, w / rx / ry per node), tying rigidity W (:9541-9550), residual bending
 * flexibility Zb3 with directional pairing before (Zb) and after (Zg) the guard
 * (:9578-9724); W_out receives the rigidity that enters Ks = D3' W D3, i.e.
 * inv(sym(inv(W) + Zg)) when Zb3 applies.  Returns 1 when Zb3 applies, 0 when not,
 * a negative status on a degenerate inverse. */
static int t3_shear(const double lc[6], const double Db[9], const double Ds[4], double A,
                    double grA[9], double gsB[9], double cvec[9], double Wrig[9], double Zb_raw[9],
                    double Zb_guarded[9], double W_out[9]) {
    double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    /* MITC3 tying rows :9484-9508 */
    double Jm11 = x2 - x1, Jm12 = y2 - y1, Jm21 = x3 - x1, Jm22 = y3 - y1;
    double dj = Jm11 * Jm22 - Jm12 * Jm21;
    double Ji11 = Jm22 / dj, Ji12 = -Jm12 / dj, Ji21 = -Jm21 / dj, Ji22 = Jm11 / dj;
    double grC[9], gsC[9];
    {
        const double rr[3] = {0.5, 0.0, 0.5}, ss[3] = {0.0, 0.5, 0.5};
        const double dwdr[3] = {-1, 1, 0}, dwds[3] = {-1, 0, 1};
        for (int pnt = 0; pnt < 3; ++pnt) {
            double N[3] = {1 - rr[pnt] - ss[pnt], rr[pnt], ss[pnt]};
            double Br[9], Bs[9];
            for (int i = 0; i < 3; ++i) {
                Br[3 * i] = dwdr[i]; Bs[3 * i] = dwds[i];
                Br[3 * i + 2] = N[i] * Jm11; Bs[3 * i + 2] = N[i] * Jm21;
                Br[3 * i + 1] = -N[i] * Jm12; Bs[3 * i + 1] = -N[i] * Jm22;
            }
            if (pnt == 0) memcpy(grA, Br, sizeof(Br));
            else if (pnt == 1) memcpy(gsB, Bs, sizeof(Bs));
            else { memcpy(grC, Br, sizeof(Br)); memcpy(gsC, Bs, sizeof(Bs)); }
        }
    }
    for (int i = 0; i < 9; ++i) cvec[i] = (grC[i] - grA[i]) - (gsC[i] - gsB[i]);
    double La = hypot(x2 - x1, y2 - y1), Lb = hypot(x3 - x1, y3 - y1), Lc = hypot(x3 - x2, y3 - y2);
    double P = La + Lb + Lc;
    double cth = ((x2 - x1) * (x3 - x1) + (y2 - y1) * (y3 - y1)) / fmax(La * Lb, 1e-30);
    if (cth < -1) cth = -1; if (cth > 1) cth = 1;
    double sth = sqrt(fmax(1 - cth * cth, 0.0));
    /* reference tying rigidity :9541-9550; G = Jinv' Ds Jinv */
    double G11, G12, G22;
    {
        /* t = Ds * Jinv ; G = Jinv' * t */
        double t11 = M(Ds, 0, 0, 2) * Ji11 + M(Ds, 0, 1, 2) * Ji21, t12 = M(Ds, 0, 0, 2) * Ji12 + M(Ds, 0, 1, 2) * Ji22;
        double t21 = M(Ds, 1, 0, 2) * Ji11 + M(Ds, 1, 1, 2) * Ji21, t22 = M(Ds, 1, 0, 2) * Ji12 + M(Ds, 1, 1, 2) * Ji22;
        G11 = Ji11 * t11 + Ji21 * t21; G12 = Ji11 * t12 + Ji21 * t22; G22 = Ji12 * t12 + Ji22 * t22;
    }
    double uu = 3 * La / P, vv = 3 * Lb / P;
    double *W = W_out;
    M(W, 0, 0, 3) = G11 * A; M(W, 0, 1, 3) = G12 * A; M(W, 1, 0, 3) = M(W, 0, 1, 3); M(W, 1, 1, 3) = G22 * A;
    M(W, 0, 2, 3) = (A / 3) * (uu * G11 - vv * G12); M(W, 2, 0, 3) = M(W, 0, 2, 3);
    M(W, 1, 2, 3) = (A / 3) * (uu * G12 - vv * G22); M(W, 2, 1, 3) = M(W, 1, 2, 3);
    M(W, 2, 2, 3) = uu * uu * G11 * (A / 6) - 2 * uu * vv * G12 * (A * (2 - cth) / 18) + vv * vv * G22 * (A / 6);
    memcpy(Wrig, W, 9 * sizeof(double));
    for (int q = 0; q < 9; ++q) { Zb_raw[q] = 0; Zb_guarded[q] = 0; }
    /* residual bending flexibility Zb3 with directional pairing :9578-9728 */
    double Dbend = fabs(M(Db, 0, 0, 3));
    if (sth > 1e-9 && Dbend > 1e-30 && La > 1e-30 && Lb > 1e-30) {
        double S1 = La, S2 = Lb, S3 = P, s6 = 1.0 / 6;
        double D11e = Dbend, D22e = fabs(M(Db, 1, 1, 3));
        int dir_pair = D11e > 1e-30 && D22e > 1e-30;
        double rho = La / Lb, Rr2 = rho + 1 / rho - 2, Tt2 = rho - 1 / rho;
        double fr = ((1 - cth) * (1 - cth) - (rho - cth)) / (6 * sth) + Rr2 * (1 + cth * cth) / (12 * sth) + Tt2 * sth / 12;
        double fri = ((1 - cth) * (1 - cth) - (1 / rho - cth)) / (6 * sth) + Rr2 * (1 + cth * cth) / (12 * sth) - Tt2 * sth / 12;
        double w12 = (1 / rho - cth) / (6 * sth), w13 = (rho - cth) / (6 * sth), w23 = cth / (6 * sth);
        double o12 = (Rr2 * cth - (1 - cth) * (1 - cth) - Tt2 * cth) / (12 * sth);
        double o13 = (Rr2 * cth - (1 - cth) * (1 - cth) + Tt2 * cth) / (12 * sth);
        double o23 = (1 + cth * cth) / (12 * sth);
        double Rr = (La - Lb) * (La - Lb) / (La * Lb), Tt = (La * La - Lb * Lb) / (La * Lb);
        double sum_d = 2 * ((1 - cth + cth * cth) / (6 * sth) + 1.0 / 6) + Rr * (1 + cth * cth) / (6 * sth);
        double dif_d = Tt * sth / 6;
        double a11 = (sum_d + dif_d) / 2, a22 = (sum_d - dif_d) / 2;
        double a12 = (-1 + 4 * cth - cth * cth) / (12 * sth) - 1.0 / 6 + Rr * cth / (6 * sth);
        double ev[3][2] = {{x2 - x1, y2 - y1}, {x3 - x1, y3 - y1}, {x3 - x2, y3 - y2}};
        double Dn[3];
        for (int e = 0; e < 3; ++e) {
            double vx = ev[e][0], vy = ev[e][1], n = hypot(vx, vy);
            if (!dir_pair) { Dn[e] = Dbend; continue; }
            if (n < 1e-30) { Dn[e] = Dbend; continue; }
            double nx = -vy / n, ny = vx / n;
            double wv[3] = {nx * nx, ny * ny, 2 * nx * ny};
            double d = 0;
            for (int ii = 0; ii < 3; ++ii) for (int jj = 0; jj < 3; ++jj) d += wv[ii] * M(Db, ii, jj, 3) * wv[jj];
            Dn[e] = fabs(d) > 1e-30 ? d : Dbend;
        }
        double Gd = dir_pair ? sqrt(D11e * D22e) : Dbend;
        double c11 = dir_pair ? s6 / Gd + fr / Dn[0] + w13 / Dn[1] + w23 / Dn[2] : a11 / Dbend;
        double c22 = dir_pair ? s6 / Gd + w12 / Dn[0] + fri / Dn[1] + w23 / Dn[2] : a22 / Dbend;
        double c12 = dir_pair ? -s6 / Gd + o12 / Dn[0] + o13 / Dn[1] + o23 / Dn[2] : a12 / Dbend;
        double Zb[9];
        M(Zb, 0, 0, 3) = c11 * S1 * S1; M(Zb, 0, 1, 3) = c12 * S1 * S2;
        M(Zb, 1, 0, 3) = M(Zb, 0, 1, 3); M(Zb, 1, 1, 3) = c22 * S2 * S2;
        M(Zb, 0, 2, 3) = -s6 / Gd * S1 * S3; M(Zb, 2, 0, 3) = M(Zb, 0, 2, 3);
        M(Zb, 1, 2, 3) = s6 / Gd * S2 * S3; M(Zb, 2, 1, 3) = M(Zb, 1, 2, 3);
        M(Zb, 2, 2, 3) = s6 / Gd * S3 * S3;
        memcpy(Zb_raw, Zb, sizeof(Zb));
        tria3_guard(Zb, La, Lb, cth, sth, dir_pair ? Dn[0] : Dbend);
        memcpy(Zb_guarded, Zb, sizeof(Zb));
        double Wi[9], Zt[9], Zs[9];
        if (inv3(W, Wi) != OR_OK) return -OR_ERR_DEGENERATE;
        for (int q = 0; q < 9; ++q) Zt[q] = Wi[q] + Zb[q];
        for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) Zs[i + 3 * j] = 0.5 * (Zt[i + 3 * j] + Zt[j + 3 * i]);
        if (inv3(Zs, W) != OR_OK) return -OR_ERR_DEGENERATE;
        return 1;
    }
    return 0;
}

int or_t3_shear_parts(const double lc[6], const double Db[9], const double Ds[4], double D3[27], double Wrig[9],
                      double Zb[9], double Zg[9], double W[9]) {
    double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    double A = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
    return t3_shear(lc, Db, Ds, A, D3, D3 + 9, D3 + 18, Wrig, Zb, Zg, W);
}

int or_t3_local(const double lc[6], const double Dm[9], const double Db[9],
                const double Ds[4], double h, double G_ref, double k6rot, double Ke[324]) {
    (void)h; (void)G_ref;
    for (int q = 0; q < 324; ++q) Ke[q] = 0;
    double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    double A2 = x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2);
    double A = 0.5 * fabs(A2);
    if (A < 1e-12) return OR_OK; /* zero matrix, :9381 */
    double bv[3] = {(y2 - y3) / (2 * A), (y3 - y1) / (2 * A), (y1 - y2) / (2 * A)};
    double cv[3] = {(x3 - x2) / (2 * A), (x1 - x3) / (2 * A), (x2 - x1) / (2 * A)};
    /* CST membrane :9390-9396 */
    double Bm[18] = {0};
    for (int i = 0; i < 3; ++i) {
        M(Bm, 0, 2 * i, 3) = bv[i]; M(Bm, 1, 2 * i + 1, 3) = cv[i];
        M(Bm, 2, 2 * i, 3) = cv[i]; M(Bm, 2, 2 * i + 1, 3) = bv[i];
    }
    static const int mi[6] = {0, 1, 6, 7, 12, 13};
    for (int j = 0; j < 6; ++j)
        for (int i = 0; i < 6; ++i) {
            double s = 0;
            for (int a = 0; a < 3; ++a) { double t = 0; for (int b = 0; b < 3; ++b) t += M(Bm, b, i, 3) * M(Dm, b, a, 3); s += t * M(Bm, a, j, 3); }
            M(Ke, mi[i], mi[j], 18) = s * A;
        }
    /* constant-curvature bending :9399-9408 */
    double Bb[27] = {0};
    for (int i = 0; i < 3; ++i) {
        M(Bb, 0, 3 * i + 2, 3) = bv[i]; M(Bb, 1, 3 * i + 1, 3) = -cv[i];
        M(Bb, 2, 3 * i + 1, 3) = -bv[i]; M(Bb, 2, 3 * i + 2, 3) = cv[i];
    }
    double Kb[81];
    for (int j = 0; j < 9; ++j)
        for (int i = 0; i < 9; ++i) {
            double s = 0;
            for (int a = 0; a < 3; ++a) { double t = 0; for (int b = 0; b < 3; ++b) t += M(Bb, b, i, 3) * M(Db, b, a, 3); s += t * M(Bb, a, j, 3); }
            M(Kb, i, j, 9) = s * A;
        }
    double grA[9], gsB[9], cvec[9], Wrig[9], Zb_raw[9], Zb_g[9], W[9];
    int sh = t3_shear(lc, Db, Ds, A, grA, gsB, cvec, Wrig, Zb_raw, Zb_g, W);
    if (sh < 0) return -sh;
    /* Ks = D3' W D3, D3 rows grA, gsB, cvec :9731-9733 */
    const double *D3[3] = {grA, gsB, cvec};
    static const int bi[9] = {2, 3, 4, 8, 9, 10, 14, 15, 16};
    for (int j = 0; j < 9; ++j)
        for (int i = 0; i < 9; ++i) {
            double s = 0;
            for (int a = 0; a < 3; ++a) { double t = 0; for (int b = 0; b < 3; ++b) t += D3[b][i] * M(W, b, a, 3); s += t * D3[a][j]; }
            M(Ke, bi[i], bi[j], 18) += M(Kb, i, j, 9) + s;
        }
    /* lumped drilling :9780-9795 */
    double alpha_L = (k6rot * 1e-6) * M(Dm, 2, 2, 3) * (2 * A);
    for (int i = 0; i < 3; ++i) {
        double Bd[18] = {0};
        for (int j = 0; j < 3; ++j) { Bd[6 * j] = 0.5 * cv[j]; Bd[6 * j + 1] = -0.5 * bv[j]; }
        Bd[6 * i + 5] = 1.0;
        for (int b = 0; b < 18; ++b) for (int a = 0; a < 18; ++a) M(Ke, a, b, 18) += alpha_L * (Bd[a] * Bd[b]);
    }
    return OR_OK;
}

int or_t3_global(const double p[9], const double prop[8], double k6rot, double Ke[324]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6;
    /* shell_element_frame_fast n=3, src/solver/helpers.jl:1486-1490 */
    double v1[3], v2[3], v3[3], e13[3];
    v3sub(p2, p1, v1); v3normalize(v1);
    v3sub(p3, p1, e13); v3cross(v1, e13, v3); v3normalize(v3);
    v3cross(v3, v1, v2);
    double c[3], lc[6];
    for (int k = 0; k < 3; ++k) c[k] = ((p1[k] + p2[k]) + p3[k]) / 3.0;
    for (int k = 0; k < 3; ++k) { double d[3]; v3sub(p + 3 * k, c, d); lc[2 * k] = v3dot(d, v1); lc[2 * k + 1] = v3dot(d, v2); }
    double Cm[9], Cb[9], Cs[4], G; int mo;
    shell_section(prop, Cm, Cb, Cs, &G, &mo);
    double Kl[324];
    int st = or_t3_local(lc, Cm, Cb, Cs, prop[OR_SP_H], G, mo ? 0.0 : k6rot, Kl);
    if (st != OR_OK) return st;
    double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    transform_blocks(Kl, 18, R, Ke);
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* shell force / stress / strain recovery, isotropic PSHELL default branches  */
/* (recover_shell_stresses!, src/solver/stress_recovery.jl:505-1268)           */
/* ------------------------------------------------------------------------- */
/* bilinear extrapolation of 2 x 2 Gauss-point rows G[g*3+c] to (xi, eta):
 * interp_2x2_gauss_sigma, src/experimental/plate_kernels.jl:53-70 */
static void interp_gauss(const double G[12], double xi, double eta, double out[3]) {
    const double pt = 1.0 / sqrt(3.0);
    double lx1 = (pt - xi) / (2.0 * pt), lx2 = (xi + pt) / (2.0 * pt);
    double ly1 = (pt - eta) / (2.0 * pt), ly2 = (eta + pt) / (2.0 * pt);
    double w1 = lx1 * ly1, w2 = lx2 * ly1, w3 = lx2 * ly2, w4 = lx1 * ly2;
    for (int c = 0; c < 3; ++c) out[c] = w1 * G[c] + w2 * G[3 + c] + w3 * G[6 + c] + w4 * G[9 + c];
}

/* element-frame displacements: per node R u (translations) and R r (rotations),
 * R rows (v1, v2, v3), src/solver/stress_recovery.jl:609-617 / :1003-1009 */
static void element_displacements(const double *ug, int nn, const double *v1, const double *v2,
                                  const double *v3, double *u) {
    const double *R[3] = {v1, v2, v3};
    for (int k = 0; k < nn; ++k)
        for (int hf = 0; hf < 2; ++hf)
            for (int i = 0; i < 3; ++i) {
                const double *s = ug + 6 * k + 3 * hf;
                u[6 * k + 3 * hf + i] = R[i][0] * s[0] + R[i][1] * s[1] + R[i][2] * s[2];
            }
}

/* MITC4 centre transverse resultants with the phi2 factor (PHI2_ALPHA = 10):
 * quad4_mitc4_center_shear_resultant, src/FEMKernels.jl:6515-6599 (no SNORM) */
static void mitc4_center_shear(const double lc[8], const double u[24], double G, double h, double Q[2]) {
    const double tr[4] = {0.0, 0.0, -1.0, 1.0}, ts[4] = {-1.0, 1.0, 0.0, 0.0};
    double Bs[4][24];
    memset(Bs, 0, sizeof(Bs));
    for (int t = 0; t < 4; ++t) {
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs_quad(tr[t], ts[t], dNr, dNs);
        jac4(dNr, dNs, lc, &J11, &J12, &J21, &J22);
        double Nt[4] = {0.25 * (1.0 - tr[t]) * (1.0 - ts[t]), 0.25 * (1.0 + tr[t]) * (1.0 - ts[t]),
                        0.25 * (1.0 + tr[t]) * (1.0 + ts[t]), 0.25 * (1.0 - tr[t]) * (1.0 + ts[t])};
        for (int k = 0; k < 4; ++k) {
            if (t < 2) { Bs[t][6 * k + 2] = dNr[k]; Bs[t][6 * k + 3] = -J12 * Nt[k]; Bs[t][6 * k + 4] = J11 * Nt[k]; }
            else       { Bs[t][6 * k + 2] = dNs[k]; Bs[t][6 * k + 3] = -J22 * Nt[k]; Bs[t][6 * k + 4] = J21 * Nt[k]; }
        }
    }
    const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac4(dNr_c, dNs_c, lc, &J11c, &J12c, &J21c, &J22c);
    double detJc = J11c * J22c - J12c * J21c, abs_detJc = fmax(fabs(detJc), 1e-30);
    double inv_detc = 1.0 / detJc;
    double iJ11c = J22c * inv_detc, iJ12c = -J12c * inv_detc, iJ21c = -J21c * inv_detc, iJ22c = J11c * inv_detc;
    double g1 = 0.0, g2 = 0.0;
    for (int j = 0; j < 24; ++j) {
        g1 += (0.5 * (Bs[0][j] + Bs[1][j])) * u[j];
        g2 += (0.5 * (Bs[2][j] + Bs[3][j])) * u[j];
    }
    double phi2 = fmin(1.0, 10.0 * (h * h) / fmax(4.0 * abs_detJc, 1e-30));
    double k = ((5.0 / 6.0) * G) * h;
    double t11 = k * iJ11c, t12 = k * iJ12c, t21 = k * iJ21c, t22 = k * iJ22c;
    double q1 = (phi2 * (iJ11c * t11 + iJ21c * t21)) * g1 + (phi2 * (iJ11c * t12 + iJ21c * t22)) * g2;
    double q2 = (phi2 * (iJ12c * t11 + iJ22c * t21)) * g1 + (phi2 * (iJ12c * t12 + iJ22c * t22)) * g2;
    Q[0] = J11c * q1 + J21c * q2;
    Q[1] = J12c * q1 + J22c * q2;
}

/* equilibrium shear (dMx/dx, dMy/dy) of bilinear corner moments and the
 * smaller-magnitude blend: _quad4_equilibrium_shear_from_bending /
 * _quad4_blend_recovered_shear, src/solver/stress_recovery.jl:126-168 */
static void blend_equilibrium_shear(const double lc[8], const double Mc[12], double Q[2]) {
    const double dNr[4] = {-0.25, 0.25, 0.25, -0.25}, dNs[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11, J12, J21, J22;
    jac4(dNr, dNs, lc, &J11, &J12, &J21, &J22);
    double detJ = J11 * J22 - J12 * J21;
    if (fabs(detJ) < 1e-12) return;
    double i11 = J22 / detJ, i12 = -J12 / detJ, i21 = -J21 / detJ, i22 = J11 / detJ;
    double eq[2];
    for (int comp = 0; comp < 2; ++comp) {
        double dr = 0.0, ds = 0.0;
        for (int k = 0; k < 4; ++k) { dr += dNr[k] * Mc[3 * k + comp]; ds += dNs[k] * Mc[3 * k + comp]; }
        eq[comp] = comp == 0 ? i11 * dr + i12 * ds : i21 * dr + i22 * ds;
    }
    for (int i = 0; i < 2; ++i)
        if (isfinite(eq[i]) && fabs(eq[i]) < fabs(Q[i])) Q[i] = eq[i];
}

int or_q4_recover(const double p[12], const double prop[8], const double ug[24], int mean_shear,
                  double out[OR_Q4R_N]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    double v1[3], v2[3], v3[3];
    frame_quad4(p1, p2, p3, p4, v1, v2, v3);
    double c[3], lc[8];
    for (int k = 0; k < 3; ++k) c[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4.0;
    for (int k = 0; k < 4; ++k) {
        double d[3]; v3sub(p + 3 * k, c, d);
        lc[2 * k] = v3dot(d, v1); lc[2 * k + 1] = v3dot(d, v2);
    }
    double u0[24], u[24];
    element_displacements(ug, 4, v1, v2, v3, u0);
    /* physical -> projected displacement (finite-warp map of a warped cell):
     * quad4_physical_displacement_map, src/FEMKernels.jl:1062-1079, applied by
     * stress_strain_quad4 :6768-6771 and by the section fields :7160-7161 */
    if (quad4_is_warped(p, lc)) {
        double Tg[576];
        quad4_warp_map(lc, p, Tg);
        for (int i = 0; i < 24; ++i) {
            double s = 0.0;
            for (int j = 0; j < 24; ++j) s += M(Tg, i, j, 24) * u0[j];
            u[i] = s;
        }
    } else {
        memcpy(u, u0, sizeof(u));
    }
    const double h = prop[OR_SP_H], E = prop[OR_SP_E], nu = prop[OR_SP_NU], br = prop[OR_SP_BR];
    const double tst = prop[OR_SP_TST];
    double G = prop[OR_SP_G];
    if (!(G > 0.0)) G = E > 0.0 ? E / (2 * (1 + nu)) : 0.0;
    /* stress_strain_quad4's plane-stress law and membrane matrix, :6759-6763 */
    const double cmem = E / (1 - nu * nu);
    const double Dm3[9] = {cmem * 1.0, cmem * nu, cmem * 0.0, cmem * nu, cmem * 1.0, cmem * 0.0,
                           cmem * 0.0, cmem * 0.0, cmem * ((1 - nu) / 2)};
    double Cm[9];
    for (int q = 0; q < 9; ++q) Cm[q] = Dm3[q] * h;
    /* section law of the MacNeal mean-shear data (A = h C, D = br h^3/12 C,
     * Cs = ts h G): _pshell_macneal_mean_shear_data, src/solver/stress_recovery.jl:248-292,
     * pshell_plane_stress_matrix src/solver/assembly.jl:2183-2233 */
    const double pc = E / fmax(1.0 - nu * nu, 1e-30);
    double Csec[9] = {pc * 1.0, pc * nu, pc * 0.0, pc * nu, pc * 1.0, pc * 0.0, pc * 0.0, pc * 0.0, pc * ((1.0 - nu) / 2.0)};
    double Asec[9], Dsec[9];
    for (int q = 0; q < 9; ++q) { Asec[q] = h * Csec[q]; Dsec[q] = ((br * (h * h * h)) / 12) * Csec[q]; }
    double Cs[4] = {(tst * h) * G, 0.0, 0.0, (tst * h) * G};
    /* centre Jacobian, :6774-6778 */
    const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac4(dNr_c, dNs_c, lc, &J11c, &J12c, &J21c, &J22c);
    double detJc = J11c * J22c - J12c * J21c;
    double iJ11c = J22c / detJc, iJ12c = -J12c / detJc, iJ21c = -J21c / detJc, iJ22c = J11c / detJc;
    const double pt = 1.0 / sqrt(3.0);
    const double gr[4] = {-pt, pt, pt, -pt}, gs[4] = {-pt, -pt, pt, pt};
    /* per Gauss point: compatible membrane strain Bm u, curvature Bb u, the
     * membrane / bending internal-mode B matrices, weight */
    double em[4][3], kb[4][3], Bi[4][12], Bib[4][12], wg[4];
    double Kabu[4] = {0}, Kbb[16] = {0}, Kabsu[4] = {0}, Kbbs[16] = {0};
    for (int g = 0; g < 4; ++g) {
        double r = gr[g], s = gs[g], dNr[4], dNs[4], J11, J12, J21, J22, w;
        shape_derivs_quad(r, s, dNr, dNs);
        jac4(dNr, dNs, lc, &J11, &J12, &J21, &J22);
        double detJ = J11 * J22 - J12 * J21;
        if (jac_measure(detJ, J11, J12, J21, J22, &w) != OR_OK) return OR_ERR_DEGENERATE;
        double iJ11 = J22 / detJ, iJ12 = -J12 / detJ, iJ21 = -J21 / detJ, iJ22 = J11 / detJ;
        double e0 = 0, e1 = 0, e2 = 0, k0 = 0, k1 = 0, k2 = 0;
        for (int k = 0; k < 4; ++k) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k], dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            const double *uk = u + 6 * k;
            e0 += dx * uk[0]; e1 += dy * uk[1]; e2 += dy * uk[0] + dx * uk[1];
            k0 += dx * uk[4]; k1 += -dy * uk[3]; k2 += dy * uk[4] - dx * uk[3];
        }
        em[g][0] = e0; em[g][1] = e1; em[g][2] = e2;
        kb[g][0] = k0; kb[g][1] = k1; kb[g][2] = k2;
        wg[g] = w;
        /* membrane internal modes (fill_quad4_membrane_incompatible_B!, :1868-1896) */
        double ratio = (iJ11 * iJ22 - iJ12 * iJ21) / (iJ11c * iJ22c - iJ12c * iJ21c);
        double m11 = ratio * iJ11c, m12 = ratio * iJ12c, m21 = ratio * iJ21c, m22 = ratio * iJ22c;
        double p1x = m11 * (-2.0 * r), p1y = m21 * (-2.0 * r), p2x = m12 * (-2.0 * s), p2y = m22 * (-2.0 * s);
        double *B = Bi[g];
        memset(B, 0, 12 * sizeof(double));
        M(B, 0, 0, 3) = p1x; M(B, 2, 0, 3) = p1y; M(B, 1, 1, 3) = p1y; M(B, 2, 1, 3) = p1x;
        M(B, 0, 2, 3) = p2x; M(B, 2, 2, 3) = p2y; M(B, 1, 3, 3) = p2y; M(B, 2, 3, 3) = p2x;
        /* bending internal modes of the section fields' stiffness kernel (:3478-3488) */
        double gq = detJc / detJ;
        double b1x = (iJ11c * (-2.0 * r)) * gq, b1y = (iJ21c * (-2.0 * r)) * gq;
        double b2x = (iJ12c * (-2.0 * s)) * gq, b2y = (iJ22c * (-2.0 * s)) * gq;
        double *Bb = Bib[g];
        memset(Bb, 0, 12 * sizeof(double));
        M(Bb, 1, 0, 3) = -b1y; M(Bb, 1, 1, 3) = -b2y;
        M(Bb, 0, 2, 3) = b1x;  M(Bb, 0, 3, 3) = b2x;
        M(Bb, 2, 0, 3) = -b1x; M(Bb, 2, 1, 3) = -b2x;
        M(Bb, 2, 2, 3) = b1y;  M(Bb, 2, 3, 3) = b2y;
        /* K_ab' u, K_bb (membrane, Cm; :6824-6871) and K_ab' u, K_bb of the bending
         * modes with the shape metric A (the enrichment map of the section fields) */
        double tm[12], tb[12];
        ts_mul(tm, Cm, B, 3, 3, 4);
        ts_mul(tb, Asec, Bb, 3, 3, 4);
        for (int m = 0; m < 4; ++m) {
            Kabu[m] += (e0 * M(tm, 0, m, 3) + e1 * M(tm, 1, m, 3) + e2 * M(tm, 2, m, 3)) * w;
            Kabsu[m] += (k0 * M(tb, 0, m, 3) + k1 * M(tb, 1, m, 3) + k2 * M(tb, 2, m, 3)) * w;
        }
        ts_mul_At_add(Kbb, B, tm, w, 3, 4, 4);
        ts_mul_At_add(Kbbs, Bb, tb, w, 3, 4, 4);
    }
    /* alpha = -(K_bb \ K_ab' u) (:6873); bending enrichment ab = -K_bbs^{-1} K_abs' u
     * (section_kinematic enrichment of quad4_condensed_section_fields :7173-7175) */
    double Ci[16], alpha[4] = {0}, ab[4] = {0};
    if (inv4(Kbb, Ci) != OR_OK) return OR_ERR_DEGENERATE;
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k) alpha[i] -= M(Ci, i, k, 4) * Kabu[k];
    if (maxabs(Kbbs, 16) > 0) {
        if (inv4(Kbbs, Ci) != OR_OK) return OR_ERR_DEGENERATE;
        for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k) ab[i] -= M(Ci, i, k, 4) * Kabsu[k];
    }
    /* Gauss-point fields and area averages (:6896-6955); the section moments
     * M = D kappa with the bending enrichment (mean-shear branch) or the
     * compatible -Cb kappa of stress_strain_quad4 (:6898-6900, :6944-6947) */
    double Cbc[9];
    for (int q = 0; q < 9; ++q) Cbc[q] = (br * Dm3[q]) * ((h * h * h) / 12.0);
    double Ngp[12], Msec[12], eavg[3] = {0}, kavg[3] = {0}, area = 0.0;
    for (int g = 0; g < 4; ++g) {
        double eps[3], kap[3];
        for (int l = 0; l < 3; ++l) {
            double v = 0.0;
            for (int m = 0; m < 4; ++m) v += M(Bi[g], l, m, 3) * alpha[m];
            eps[l] = em[g][l] + v;
            double vb = 0.0;
            for (int m = 0; m < 4; ++m) vb += M(Bib[g], l, m, 3) * ab[m];
            kap[l] = kb[g][l] + vb;
        }
        for (int l = 0; l < 3; ++l) {
            Ngp[3 * g + l] = M(Cm, l, 0, 3) * eps[0] + M(Cm, l, 1, 3) * eps[1] + M(Cm, l, 2, 3) * eps[2];
            if (mean_shear)
                Msec[3 * g + l] = M(Dsec, l, 0, 3) * kap[0] + M(Dsec, l, 1, 3) * kap[1] + M(Dsec, l, 2, 3) * kap[2];
            else
                Msec[3 * g + l] = -(M(Cbc, l, 0, 3) * kb[g][0] + M(Cbc, l, 1, 3) * kb[g][1] + M(Cbc, l, 2, 3) * kb[g][2]);
            eavg[l] += eps[l] * wg[g];
            kavg[l] += kb[g][l] * wg[g];
        }
        area += wg[g];
    }
    for (int l = 0; l < 3; ++l) { eavg[l] /= area; kavg[l] /= area; }
    double *N = out + OR_Q4R_N0, *Mo = out + OR_Q4R_M0, *Q = out + OR_Q4R_Q0;
    double *sz1 = out + OR_Q4R_S1, *sz2 = out + OR_Q4R_S2, *ez1 = out + OR_Q4R_E1, *ez2 = out + OR_Q4R_E2;
    double *Nc = out + OR_Q4R_NC, *Mc = out + OR_Q4R_MC;
    /* resultants and fibre fields at -h/2, +h/2 (:6956-6966) */
    for (int l = 0; l < 3; ++l) {
        N[l] = M(Cm, l, 0, 3) * eavg[0] + M(Cm, l, 1, 3) * eavg[1] + M(Cm, l, 2, 3) * eavg[2];
        double dk = M(Dm3, l, 0, 3) * kavg[0] + M(Dm3, l, 1, 3) * kavg[1] + M(Dm3, l, 2, 3) * kavg[2];
        Mo[l] = ((-br) * dk) * ((h * h * h) / 12.0);
    }
    const double z1 = -h / 2.0, z2 = h / 2.0;
    for (int l = 0; l < 3; ++l) { ez1[l] = eavg[l] + z1 * kavg[l]; ez2[l] = eavg[l] + z2 * kavg[l]; }
    for (int l = 0; l < 3; ++l) {
        sz1[l] = M(Dm3, l, 0, 3) * ez1[0] + M(Dm3, l, 1, 3) * ez1[1] + M(Dm3, l, 2, 3) * ez1[2];
        sz2[l] = M(Dm3, l, 0, 3) * ez2[0] + M(Dm3, l, 1, 3) * ez2[1] + M(Dm3, l, 2, 3) * ez2[2];
    }
    /* corner resultants: N from the condensed Gauss-point membrane field, M
     * (printed sign) from the section fields (stress_recovery.jl:933-946) or from
     * stress_strain_quad4's Gauss points (:6967-6974) */
    const double cr[4] = {-1.0, 1.0, 1.0, -1.0}, cs[4] = {-1.0, -1.0, 1.0, 1.0};
    const int bending = br > 1e-12;
    for (int k = 0; k < 4; ++k) {
        double t[3];
        interp_gauss(Ngp, cr[k], cs[k], Nc + 3 * k);
        interp_gauss(Msec, cr[k], cs[k], t);
        for (int l = 0; l < 3; ++l) Mc[3 * k + l] = mean_shear ? (bending ? -t[l] : 0.0) : t[l];
    }
    Q[0] = Q[1] = 0.0;
    if (!mean_shear) {
        /* other solution sequences: MITC4 centre shear (:6953-6954) blended with the
         * equilibrium shear of the corner moments (stress_recovery.jl:948-952) */
        if (bending) {
            mitc4_center_shear(lc, u, E / (2 * (1 + nu)), h, Q);
            blend_equilibrium_shear(lc, Mc, Q);
        }
        for (int k = 0; k < 4; ++k) Mc[3 * k + 2] = Mo[2];
        return OR_OK;
    }
    /* twist held at the centre value (:960) */
    for (int k = 0; k < 4; ++k) Mc[3 * k + 2] = Mo[2];
    /* area-mean MacNeal transverse resultants (quad4_macneal_mean_shear_resultant
     * :6623-6645): work of the plate operator against unit shear states */
    const int rigid = prop[OR_SP_RIGID] != 0.0;
    if (bending && tst > 0 && h > 0 && !(!rigid && maxabs(Cs, 4) < 1e-30)) {
        double D[48], Zb[16], Zs[16], Z[16], Kp[144];
        int st = or_q4_macneal_parts(lc, Dsec, Cs, rigid, D, Zb, Zs, Z, Kp);
        if (st != OR_OK) return st;
        double qx = 0.0, qy = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int cc = 0; cc < 3; ++cc) {
                    qx += M(Kp, 3 * i + 2, 3 * j + cc, 12) * u[6 * j + cc + 2];
                    qy -= M(Kp, 3 * i + 1, 3 * j + cc, 12) * u[6 * j + cc + 2];
                }
        double ar = 4 * fabs(detJc);
        Q[0] = qx / ar; Q[1] = qy / ar;
    }
    return OR_OK;
}

/* CTRIA3: stress_strain_tria3, src/FEMKernels.jl:9895-9937, with the MITC3
 * centroid shear tria3_mitc3_shear_resultant :9859-9893; frame and element
 * displacements src/solver/stress_recovery.jl:985-1009 */
int or_t3_recover(const double p[9], const double prop[8], const double ug[18], double out[OR_T3R_N]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6;
    double v1[3], v2[3], v3[3], e13[3];
    v3sub(p2, p1, v1); v3normalize(v1);
    v3sub(p3, p1, e13); v3cross(v1, e13, v3); v3normalize(v3);
    v3cross(v3, v1, v2);
    double c[3], lc[6];
    for (int k = 0; k < 3; ++k) c[k] = ((p1[k] + p2[k]) + p3[k]) / 3.0;
    for (int k = 0; k < 3; ++k) { double d[3]; v3sub(p + 3 * k, c, d); lc[2 * k] = v3dot(d, v1); lc[2 * k + 1] = v3dot(d, v2); }
    double u[18];
    element_displacements(ug, 3, v1, v2, v3, u);
    for (int q = 0; q < OR_T3R_N; ++q) out[q] = 0.0;
    const double h = prop[OR_SP_H], E = prop[OR_SP_E], nu = prop[OR_SP_NU], br = prop[OR_SP_BR];
    const double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    double A = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
    if (A < 1e-12) return OR_OK;
    const double b[3] = {(y2 - y3) / (2 * A), (y3 - y1) / (2 * A), (y1 - y2) / (2 * A)};
    const double cc[3] = {(x3 - x2) / (2 * A), (x1 - x3) / (2 * A), (x2 - x1) / (2 * A)};
    const double cm = E / (1 - nu * nu);
    const double D[9] = {cm * 1.0, cm * nu, cm * 0.0, cm * nu, cm * 1.0, cm * 0.0, cm * 0.0, cm * 0.0, cm * ((1 - nu) / 2)};
    double eps[3] = {0, 0, 0}, kap[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {   /* Bm u_mem, columns (u_i, v_i); Bb u_rot, columns (rx_i, ry_i) */
        const double *ui = u + 6 * i;
        eps[0] += b[i] * ui[0]; eps[1] += cc[i] * ui[1]; eps[2] += cc[i] * ui[0] + b[i] * ui[1];
        kap[0] += b[i] * ui[4]; kap[1] += -cc[i] * ui[3]; kap[2] += -b[i] * ui[3] + cc[i] * ui[4];
    }
    double *N = out + OR_T3R_N0, *Mo = out + OR_T3R_M0, *Q = out + OR_T3R_Q0;
    double *sz1 = out + OR_T3R_S1, *sz2 = out + OR_T3R_S2, *ez1 = out + OR_T3R_E1, *ez2 = out + OR_T3R_E2;
    for (int l = 0; l < 3; ++l) {
        double de = M(D, l, 0, 3) * eps[0] + M(D, l, 1, 3) * eps[1] + M(D, l, 2, 3) * eps[2];
        double dk = M(D, l, 0, 3) * kap[0] + M(D, l, 1, 3) * kap[1] + M(D, l, 2, 3) * kap[2];
        N[l] = de * h;
        Mo[l] = ((-br) * dk) * ((h * h * h) / 12.0);
    }
    if (br > 1e-12) {
        double Jm11 = x2 - x1, Jm12 = y2 - y1, Jm21 = x3 - x1, Jm22 = y3 - y1;
        double detJ = Jm11 * Jm22 - Jm12 * Jm21;
        if (fabs(detJ) >= 1e-30) {
            /* covariant shear rows at the tying points A (1/2, 0), B (0, 1/2), C (1/2, 1/2) */
            double grA[9], gsB[9], grC[9], gsC[9], tmp[9];
            const double rr[3] = {0.5, 0.0, 0.5}, ss[3] = {0.0, 0.5, 0.5};
            for (int t = 0; t < 3; ++t) {
                double Nn[3] = {1.0 - rr[t] - ss[t], rr[t], ss[t]};
                const double dwdr[3] = {-1.0, 1.0, 0.0}, dwds[3] = {-1.0, 0.0, 1.0};
                double Br[9], Bs[9];
                for (int i = 0; i < 3; ++i) {
                    Br[3 * i] = dwdr[i]; Bs[3 * i] = dwds[i];
                    Br[3 * i + 2] = Nn[i] * Jm11; Bs[3 * i + 2] = Nn[i] * Jm21;
                    Br[3 * i + 1] = -Nn[i] * Jm12; Bs[3 * i + 1] = -Nn[i] * Jm22;
                }
                double *dr_ = t == 0 ? grA : (t == 2 ? grC : tmp);
                double *ds_ = t == 1 ? gsB : (t == 2 ? gsC : tmp);
                memcpy(dr_, Br, sizeof(Br));
                if (t != 0) memcpy(ds_, Bs, sizeof(Bs));
            }
            double up[9];
            for (int i = 0; i < 3; ++i) { up[3 * i] = u[6 * i + 2]; up[3 * i + 1] = u[6 * i + 3]; up[3 * i + 2] = u[6 * i + 4]; }
            double a = 0, bb = 0, cv = 0;
            for (int q = 0; q < 9; ++q) {
                a += grA[q] * up[q];
                bb += gsB[q] * up[q];
                cv += ((grC[q] - grA[q]) - (gsC[q] - gsB[q])) * up[q];
            }
            double g1 = a + cv / 3, g2 = bb - cv / 3;
            double i11 = Jm22 / detJ, i12 = -Jm12 / detJ, i21 = -Jm21 / detJ, i22 = Jm11 / detJ;
            double Gs = E / (2 * (1 + nu));
            double f = ((5.0 / 6.0) * Gs) * h;
            Q[0] = f * (i11 * g1 + i12 * g2);
            Q[1] = f * (i21 * g1 + i22 * g2);
        }
    }
    const double z1 = -h / 2.0, z2 = h / 2.0;
    for (int l = 0; l < 3; ++l) { ez1[l] = eps[l] + z1 * kap[l]; ez2[l] = eps[l] + z2 * kap[l]; }
    for (int l = 0; l < 3; ++l) {
        sz1[l] = M(D, l, 0, 3) * ez1[0] + M(D, l, 1, 3) * ez1[1] + M(D, l, 2, 3) * ez1[2];
        sz2[l] = M(D, l, 0, 3) * ez2[0] + M(D, l, 1, 3) * ez2[1] + M(D, l, 2, 3) * ez2[2];
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* CQUAD4 geometric (differential) stiffness for SOL 105: the default operator */
/* of a flat isotropic CQUAD4, geometric_stiffness_quad4_reference_kdjj_iso    */
/* (src/FEMKernels.jl:8810-8942), as assembled by assemble_geometric_stiffness */
/* (src/solver/assembly.jl:6575-6594 flatness, 6985-6998 element displacements,*/
/* 8063-8075 preload map, 8264-8320 drilling, warp congruence and transform).  */
/* ------------------------------------------------------------------------- */

/* grad_p of the kernel: bilinear gradients at (r, s) in the diagonal-bisector
 * frame (ce, se) and |detJ|; 0 when |detJ| < 1e-14 (src/FEMKernels.jl:8857-8876) */
static int kg_grad(const double X[4], const double Y[4], double ce, double se, double r, double s,
                   double dNx[4], double dNy[4], double *adet) {
    const double dNr[4] = {-(1.0 - s) * 0.25, (1.0 - s) * 0.25, (1.0 + s) * 0.25, -(1.0 + s) * 0.25};
    const double dNs[4] = {-(1.0 - r) * 0.25, -(1.0 + r) * 0.25, (1.0 + r) * 0.25, (1.0 - r) * 0.25};
    double J11 = dNr[0] * X[0] + dNr[1] * X[1] + dNr[2] * X[2] + dNr[3] * X[3];
    double J12 = dNr[0] * Y[0] + dNr[1] * Y[1] + dNr[2] * Y[2] + dNr[3] * Y[3];
    double J21 = dNs[0] * X[0] + dNs[1] * X[1] + dNs[2] * X[2] + dNs[3] * X[3];
    double J22 = dNs[0] * Y[0] + dNs[1] * Y[1] + dNs[2] * Y[2] + dNs[3] * Y[3];
    double detJ = J11 * J22 - J12 * J21;
    if (fabs(detJ) < 1e-14) return 0;
    double i11 = J22 / detJ, i12 = -J12 / detJ, i21 = -J21 / detJ, i22 = J11 / detJ;
    for (int k = 0; k < 4; ++k) {
        double gx = i11 * dNr[k] + i12 * dNs[k];
        double gy = i21 * dNr[k] + i22 * dNs[k];
        dNx[k] = ce * gx + se * gy;
        dNy[k] = -se * gx + ce * gy;
    }
    *adet = fabs(detJ);
    return 1;
}

int or_q4_kg_local(const double lc[8], const double ue[24], double E, double nu, double h, int assumed_transverse,
                   double Kg[576]) {
    for (int q = 0; q < 576; ++q) Kg[q] = 0.0;
    if (h < 1e-30) return OR_OK;
    const double X[4] = {lc[0], lc[2], lc[4], lc[6]}, Y[4] = {lc[1], lc[3], lc[5], lc[7]};
    /* an exact rectangle keeps the centre-gradient transverse form (:8824-8837) */
    double e12x = X[1] - X[0], e12y = Y[1] - Y[0], e14x = X[3] - X[0], e14y = Y[3] - Y[0];
    int exact_rectangle = (e12x != 0.0 || e12y != 0.0) && (e14x != 0.0 || e14y != 0.0) &&
                          X[2] - X[3] == e12x && Y[2] - Y[3] == e12y && X[2] - X[1] == e14x &&
                          Y[2] - Y[1] == e14y && e12x * e14x + e12y * e14y == 0.0;
    assumed_transverse = assumed_transverse && !exact_rectangle;
    /* element frame: bisector of the diagonals G1->G3 and G4->G2 (:8838-8846) */
    double d13x = X[2] - X[0], d13y = Y[2] - Y[0], l13 = hypot(d13x, d13y);
    double d42x = X[1] - X[3], d42y = Y[1] - Y[3], l42 = hypot(d42x, d42y);
    if (l13 < 1e-12 || l42 < 1e-12) return OR_OK;
    double bx = d13x / l13 + d42x / l42, by = d13y / l13 + d42y / l42, lb = hypot(bx, by);
    if (lb < 1e-12) return OR_OK;
    double ce = bx / lb, se = by / lb;
    double up[4], vp[4];
    for (int k = 0; k < 4; ++k) {
        double ux = ue[6 * k], uy = ue[6 * k + 1];
        up[k] = ce * ux + se * uy;
        vp[k] = -se * ux + ce * uy;
    }
    double d11 = E / (1.0 - nu * nu), d12 = nu * d11, d33 = E / (2.0 * (1.0 + nu));
    double dNx0[4], dNy0[4], a0;
    if (!kg_grad(X, Y, ce, se, 0.0, 0.0, dNx0, dNy0, &a0)) return OR_OK;
    /* centre-sampled membrane shear (:8881-8885) */
    double gam0 = 0.0;
    for (int k = 0; k < 4; ++k) gam0 += dNy0[k] * up[k] + dNx0[k] * vp[k];
    const double gp = 1.0 / sqrt(3.0);
    /* tying gradients: w,x at (0, -+gp), w,y at (-+gp, 0) (:8887-8899) */
    double tx[4][4], ty[4][4], ta;
    if (assumed_transverse) {
        const double tr[4] = {0.0, 0.0, -gp, gp}, ts[4] = {-gp, gp, 0.0, 0.0};
        for (int t = 0; t < 4; ++t)
            if (!kg_grad(X, Y, ce, se, tr[t], ts[t], tx[t], ty[t], &ta)) return OR_OK;
    }
    const double gr[4] = {-gp, gp, gp, -gp}, gs[4] = {-gp, -gp, gp, gp};
    for (int q = 0; q < 4; ++q) {
        double dNx[4], dNy[4], adet;
        if (!kg_grad(X, Y, ce, se, gr[q], gs[q], dNx, dNy, &adet)) continue;
        double w = h * adet, epx = 0.0, epy = 0.0;
        for (int k = 0; k < 4; ++k) { epx += dNx[k] * up[k]; epy += dNy[k] * vp[k]; }
        double sxx = d11 * epx + d12 * epy, syy = d12 * epx + d11 * epy, sxy = d33 * gam0;
        const double *dWx = assumed_transverse ? tx[q <= 1 ? 0 : 1] : dNx;
        const double *dWy = assumed_transverse ? ty[(q == 0 || q == 3) ? 2 : 3] : dNy;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                int r0 = 6 * i, c0 = 6 * j;
                /* in-plane component string rule in the primed frame (:8912-8920) */
                double sym_xy = 0.5 * (dNx[i] * dNy[j] + dNy[i] * dNx[j]);
                double kuu = w * (syy * dNy[i] * dNy[j] + sxy * sym_xy);
                double kvv = w * (sxx * dNx[i] * dNx[j] + sxy * sym_xy);
                double kuv = -w * sxy * 0.5 * (dNx[i] * dNx[j] + dNy[i] * dNy[j]);
                M(Kg, r0, c0, 24) += ce * ce * kuu - 2.0 * ce * se * kuv + se * se * kvv;
                M(Kg, r0, c0 + 1, 24) += ce * se * kuu + (ce * ce - se * se) * kuv - se * ce * kvv;
                M(Kg, r0 + 1, c0, 24) += se * ce * kuu + (ce * ce - se * se) * kuv - ce * se * kvv;
                M(Kg, r0 + 1, c0 + 1, 24) += se * se * kuu + 2.0 * se * ce * kuv + ce * ce * kvv;
                if (assumed_transverse)
                    M(Kg, r0 + 2, c0 + 2, 24) += w * (sxx * dWx[i] * dWx[j] + syy * dWy[i] * dWy[j] +
                                                      sxy * (dWx[i] * dWy[j] + dWy[i] * dWx[j]));
                else
                    M(Kg, r0 + 2, c0 + 2, 24) += w * (sxx * dNx[i] * dNx[j] + syy * dNy[i] * dNy[j] +
                                                      sxy * (dNx0[i] * dNy0[j] + dNy0[i] * dNx0[j]));
            }
    }
    return OR_OK;
}

int or_q4_kg(const double p[12], const double prop[8], const double ug[24], double Kg[576]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    /* geometric flatness gate of the flat-element operator (assembly.jl:6575-6594) */
    {
        double d13[3], d24[3], n[3];
        v3sub(p3, p1, d13); v3sub(p4, p2, d24); v3cross(d13, d24, n);
        double nl = sqrt(v3dot(n, n));
        if (nl > 1e-12) {
            double cg[3], v3g[3] = {n[0] / nl, n[1] / nl, n[2] / nl}, dev = 0.0;
            for (int k = 0; k < 3; ++k) cg[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4.0;
            for (int k = 0; k < 4; ++k) { double d[3]; v3sub(p + 3 * k, cg, d); dev = fmax(dev, fabs(v3dot(d, v3g))); }
            double ld = fmax(sqrt(v3dot(d13, d13)), sqrt(v3dot(d24, d24)));
            if (!(dev < 1e-2 * fmax(ld, 1e-12))) return OR_ERR_WARPED;
        }
    }
    double v1[3], v2[3], v3[3];
    frame_quad4(p1, p2, p3, p4, v1, v2, v3);
    double c[3], lc[8];
    for (int k = 0; k < 3; ++k) c[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4.0;
    for (int k = 0; k < 4; ++k) {
        double d[3]; v3sub(p + 3 * k, c, d);
        lc[2 * k] = v3dot(d, v1); lc[2 * k + 1] = v3dot(d, v2);
    }
    /* element-frame displacements, then the finite-warp preload map W u */
    double ul[24], uk[24], Tg[576], Kl[576];
    element_displacements(ug, 4, v1, v2, v3, ul);
    quad4_warp_map(lc, p, Tg);
    for (int i = 0; i < 24; ++i) {
        double s = 0.0;
        for (int m = 0; m < 24; ++m) s += M(Tg, i, m, 24) * ul[m];
        uk[i] = s;
    }
    int st = or_q4_kg_local(lc, uk, prop[OR_SP_E], prop[OR_SP_NU], prop[OR_SP_H], 1, Kl);
    if (st != OR_OK) return st;
    /* the operator has no drilling rows or columns (JFEM_KG_SHELL_DRILL_ZERO);
     * a warped cell takes sym(W' Kg W), then the element frame */
    if (quad4_is_warped(p, lc)) quad4_finite_warp(Kl, lc, p);
    double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    transform_blocks(Kl, 24, R, Kg);
    return OR_OK;
}

/* principal stresses and directions of (sxx, syy, sxy): principal_stress_2d_components
 * (src/FEMKernels.jl:8090-8101) */
static void kg_principal(double sxx, double syy, double sxy, double *l1, double *c1, double *s1, double *l2,
                         double *c2, double *s2) {
    double mean_s = 0.5 * (sxx + syy), half_d = 0.5 * (sxx - syy);
    double radius = sqrt(half_d * half_d + sxy * sxy);
    if (radius <= 1e-30) { *l1 = mean_s; *c1 = 1.0; *s1 = 0.0; *l2 = mean_s; *c2 = 0.0; *s2 = 1.0; return; }
    double theta = 0.5 * atan2(2.0 * sxy, sxx - syy);
    double c = cos(theta), s = sin(theta);
    *l1 = mean_s + radius; *c1 = c; *s1 = s;
    *l2 = mean_s - radius; *c2 = -s; *s2 = c;
}

/* one principal direction's transverse pair block (add_geometric_principal_transverse_pair_direction!,
 * src/FEMKernels.jl:8205-8238, unit shear factors) into Kg (n x n column-major) */
static void kg_pt_dir(double *Kg, int n, int r0, int c0, double dix, double diy, double djx, double djy, double scale,
                      double lambda, double c, double s, double uf, double vf, double uvf, double wf) {
    if (fabs(lambda) <= 1e-30) return;
    double gi = c * dix + s * diy, gj = c * djx + s * djy;
    double val = scale * lambda * gi * gj;
    double p11 = uf * s * s, p22 = vf * 1.0 * c * c, p12 = uvf * 1.0 * -c * s;
    M(Kg, r0, c0, n) += val * p11;
    M(Kg, r0, c0 + 1, n) += val * p12;
    M(Kg, r0 + 1, c0, n) += val * p12;
    M(Kg, r0 + 1, c0 + 1, n) += val * p22;
    M(Kg, r0 + 2, c0 + 2, n) += wf * 1.0 * val;
}

/* add_geometric_principal_transverse_pair! (src/FEMKernels.jl:8240-8295): both principal directions */
static void kg_pt_pair(double *Kg, int n, int r0, int c0, double dix, double diy, double djx, double djy, double scale,
                       double sxx, double syy, double sxy, double uf, double vf, double uvf, double wf) {
    double l1, c1, s1, l2, c2, s2;
    kg_principal(sxx, syy, sxy, &l1, &c1, &s1, &l2, &c2, &s2);
    kg_pt_dir(Kg, n, r0, c0, dix, diy, djx, djy, scale, l1, c1, s1, uf, vf, uvf, wf);
    kg_pt_dir(Kg, n, r0, c0, dix, diy, djx, djy, scale, l2, c2, s2, uf, vf, uvf, wf);
}

int or_t3_kg_local(const double lc[6], const double sig[3], double h, double Kg[324]) {
    for (int q = 0; q < 324; ++q) Kg[q] = 0.0;
    if (h < 1e-30) return OR_OK;
    const double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    double A2 = x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2);
    double A = 0.5 * fabs(A2);
    if (A < 1e-12) return OR_OK;
    const double bv[3] = {(y2 - y3) / (2 * A), (y3 - y1) / (2 * A), (y1 - y2) / (2 * A)};
    const double cv[3] = {(x3 - x2) / (2 * A), (x1 - x3) / (2 * A), (x2 - x1) / (2 * A)};
    const double sxx = sig[0], syy = sig[1], sxy = sig[2], sc = h * A;
    /* add_geometric_linear_principal_transverse_pair! (:8297-8338): the in-plane work
     * per stress component, the transverse work with the whole stress */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int r0 = 6 * i, c0 = 6 * j;
            kg_pt_pair(Kg, 18, r0, c0, bv[i], cv[i], bv[j], cv[j], sc, sxx, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0);
            kg_pt_pair(Kg, 18, r0, c0, bv[i], cv[i], bv[j], cv[j], sc, 0.0, syy, 0.0, 1.0, 1.0, 1.0, 0.0);
            kg_pt_pair(Kg, 18, r0, c0, bv[i], cv[i], bv[j], cv[j], sc, 0.0, 0.0, sxy, 1.0, 1.0, 1.0, 0.0);
            kg_pt_pair(Kg, 18, r0, c0, bv[i], cv[i], bv[j], cv[j], sc, sxx, syy, sxy, 0.0, 0.0, 0.0, 1.0);
        }
    return OR_OK;
}

int or_t3_kg(const double p[9], const double prop[8], const double ug[18], double Kg[324]) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6;
    double v1[3], v2[3], v3[3], e13[3];
    v3sub(p2, p1, v1); v3normalize(v1);
    v3sub(p3, p1, e13); v3cross(v1, e13, v3); v3normalize(v3);
    v3cross(v3, v1, v2);
    double c[3], lc[6];
    for (int k = 0; k < 3; ++k) c[k] = ((p1[k] + p2[k]) + p3[k]) / 3.0;
    for (int k = 0; k < 3; ++k) { double d[3]; v3sub(p + 3 * k, c, d); lc[2 * k] = v3dot(d, v1); lc[2 * k + 1] = v3dot(d, v2); }
    /* membrane resultants of the preload: stress_strain_tria3 (the recovery's N) */
    double rec[OR_T3R_N];
    int st = or_t3_recover(p, prop, ug, rec);
    if (st != OR_OK) return st;
    const double h = prop[OR_SP_H];
    double sig[3] = {rec[OR_T3R_N0] / h, rec[OR_T3R_N0 + 1] / h, rec[OR_T3R_N0 + 2] / h};
    double Kl[324];
    st = or_t3_kg_local(lc, sig, h, Kl);
    if (st != OR_OK) return st;
    double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    transform_blocks(Kl, 18, R, Kg);
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* CTETRA10, src/tetra10.jl:6-79; iso_3d_constitutive src/FEMKernels.jl:12056 */
/* ------------------------------------------------------------------------- */
int or_tet10(const double X[30], double E, double nu, double Ke[900]) {
    if (!(isfinite(E) && E > 0 && isfinite(nu) && nu > -1 && nu < 0.5)) return OR_ERR_BAD_ARG;
    const double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
    const double pts[4][3] = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
    static const int edges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    double cc = E / ((1.0 + nu) * (1.0 - 2.0 * nu)), dd = 1.0 - nu, ss = (1.0 - 2.0 * nu) / 2.0;
    double D[36] = {0};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M(D, i, j, 6) = (i == j) ? cc * dd : cc * nu;
    for (int i = 3; i < 6; ++i) M(D, i, i, 6) = cc * ss;
    double K[900] = {0};
    for (int q = 0; q < 4; ++q) {
        double L[4] = {1.0 - pts[q][0] - pts[q][1] - pts[q][2], pts[q][0], pts[q][1], pts[q][2]};
        const double dL[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
        double dN[3][10];
        for (int i = 0; i < 4; ++i) for (int k = 0; k < 3; ++k) dN[k][i] = (4.0 * L[i] - 1.0) * dL[k][i];
        for (int e = 0; e < 6; ++e) {
            int i = edges[e][0], j = edges[e][1];
            for (int k = 0; k < 3; ++k) dN[k][e + 4] = 4.0 * (L[j] * dL[k][i] + L[i] * dL[k][j]);
        }
        double J[3][3];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) {
            double s = 0; for (int n = 0; n < 10; ++n) s += dN[r][n] * M(X, n, c, 10); J[r][c] = s;
        }
        /* det(J) = col1 . (col2 x col3) */
        double c1[3] = {J[0][0], J[1][0], J[2][0]}, c2[3] = {J[0][1], J[1][1], J[2][1]}, c3[3] = {J[0][2], J[1][2], J[2][2]};
        double cr[3]; v3cross(c2, c3, cr);
        double det = v3dot(c1, cr);
        if (!(isfinite(det) && det > 0.0)) return OR_ERR_INVERTED;
        double w = det / 24.0;
        /* J \ dN via adjugate */
        double Ji[9], Jc[9];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) Jc[r + 3 * c] = J[r][c];
        Ji[0] = (Jc[4] * Jc[8] - Jc[7] * Jc[5]) / det; Ji[3] = (Jc[6] * Jc[5] - Jc[3] * Jc[8]) / det; Ji[6] = (Jc[3] * Jc[7] - Jc[6] * Jc[4]) / det;
        Ji[1] = (Jc[7] * Jc[2] - Jc[1] * Jc[8]) / det; Ji[4] = (Jc[0] * Jc[8] - Jc[6] * Jc[2]) / det; Ji[7] = (Jc[6] * Jc[1] - Jc[0] * Jc[7]) / det;
        Ji[2] = (Jc[1] * Jc[5] - Jc[4] * Jc[2]) / det; Ji[5] = (Jc[3] * Jc[2] - Jc[0] * Jc[5]) / det; Ji[8] = (Jc[0] * Jc[4] - Jc[3] * Jc[1]) / det;
        double g[3][10];
        for (int r = 0; r < 3; ++r) for (int n = 0; n < 10; ++n)
            g[r][n] = Ji[r] * dN[0][n] + Ji[r + 3] * dN[1][n] + Ji[r + 6] * dN[2][n];
        double B[180] = {0};
        for (int i = 0; i < 10; ++i) {
            int c = 3 * i; double dx = g[0][i], dy = g[1][i], dz = g[2][i];
            M(B, 0, c, 6) = dx; M(B, 1, c + 1, 6) = dy; M(B, 2, c + 2, 6) = dz;
            M(B, 3, c, 6) = dy; M(B, 3, c + 1, 6) = dx;
            M(B, 4, c + 1, 6) = dz; M(B, 4, c + 2, 6) = dy;
            M(B, 5, c, 6) = dz; M(B, 5, c + 2, 6) = dx;
        }
        double DB[180];
        for (int j = 0; j < 30; ++j) for (int i = 0; i < 6; ++i) {
            double s = 0; for (int l = 0; l < 6; ++l) s += M(D, i, l, 6) * M(B, l, j, 6); M(DB, i, j, 6) = s;
        }
        for (int j = 0; j < 30; ++j) for (int i = 0; i < 30; ++i) {
            double s = 0; for (int l = 0; l < 6; ++l) s += M(B, l, i, 6) * M(DB, l, j, 6);
            M(K, i, j, 30) += w * s;
        }
    }
    /* Matrix(Symmetric(K)): upper triangle mirrored */
    for (int j = 0; j < 30; ++j) for (int i = 0; i < 30; ++i) M(Ke, i, j, 30) = i <= j ? M(K, i, j, 30) : M(K, j, i, 30);
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* assembly: triplets + sparse(I,J,V) (src/solver/assembly.jl:432-438,          */
/* 5506-5508, 5882-5884, 5942)                                                 */
/* ------------------------------------------------------------------------- */
typedef struct { int64_t row; int32_t col; int64_t seq; double v; } trip_t;
static int trip_cmp(const void *a, const void *b) {
    const trip_t *x = a, *y = b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

int or_assemble(int64_t n_nodes, const double *x, const double *y, const double *z,
                int64_t nq, const int32_t *qc, const double *qp,
                int64_t nt, const int32_t *tc, const double *tp,
                int64_t ntet, const int32_t *ec, const double *ep,
                double k6rot, int64_t *out_nnz, int64_t **row_ptr, int32_t **col,
                double **val, int64_t *bad_elem) {
    /* triplets in element order (CQUAD4s, CTRIA3s, CTETRA10s), column-major per
     * element: element e owns a fixed slot range, so the element loops run in
     * parallel (OpenMP) and the triplet order does not depend on the schedule */
    const int64_t ntrip = nq * 576 + nt * 324 + ntet * 900;
    const int64_t t3_base = nq * 576, tet_base = t3_base + nt * 324;
    trip_t *T = (trip_t *)malloc(sizeof(trip_t) * (ntrip > 0 ? ntrip : 1));
    if (!T) return OR_ERR_OOM;
    *bad_elem = -1;
    int64_t bad = INT64_MAX; int bad_st = OR_OK;
#pragma omp parallel for schedule(static) reduction(min : bad)
    for (int64_t e = 0; e < nq; ++e) {
        double p[12], prop[8], Ke[576]; int nd[4];
        for (int k = 0; k < 4; ++k) {
            nd[k] = qc[k * nq + e];
            p[3 * k] = x[nd[k]]; p[3 * k + 1] = y[nd[k]]; p[3 * k + 2] = z[nd[k]];
        }
        for (int c = 0; c < 8; ++c) prop[c] = qp[c * nq + e];
        if (or_q4_global(p, prop, k6rot, Ke) != OR_OK) { if (e < bad) bad = e; continue; }
        int64_t seq = e * 576;
        for (int cc = 0; cc < 24; ++cc)
            for (int rr = 0; rr < 24; ++rr, ++seq) {
                T[seq].row = 6 * (int64_t)nd[rr / 6] + rr % 6; T[seq].col = 6 * nd[cc / 6] + cc % 6;
                T[seq].v = Ke[rr + 24 * cc]; T[seq].seq = seq;
            }
    }
    if (bad != INT64_MAX) {   /* the first failing element, with its status (serial re-run) */
        double p[12], prop[8], Ke[576];
        for (int k = 0; k < 4; ++k) { int nd = qc[k * nq + bad]; p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd]; }
        for (int c = 0; c < 8; ++c) prop[c] = qp[c * nq + bad];
        bad_st = or_q4_global(p, prop, k6rot, Ke);
        free(T); *bad_elem = bad; return bad_st;
    }
#pragma omp parallel for schedule(static) reduction(min : bad)
    for (int64_t e = 0; e < nt; ++e) {
        double p[9], prop[8], Ke[324]; int nd[3];
        for (int k = 0; k < 3; ++k) {
            nd[k] = tc[k * nt + e];
            p[3 * k] = x[nd[k]]; p[3 * k + 1] = y[nd[k]]; p[3 * k + 2] = z[nd[k]];
        }
        for (int c = 0; c < 8; ++c) prop[c] = tp[c * nt + e];
        if (or_t3_global(p, prop, k6rot, Ke) != OR_OK) { if (e < bad) bad = e; continue; }
        int64_t seq = t3_base + e * 324;
        for (int cc = 0; cc < 18; ++cc)
            for (int rr = 0; rr < 18; ++rr, ++seq) {
                T[seq].row = 6 * (int64_t)nd[rr / 6] + rr % 6; T[seq].col = 6 * nd[cc / 6] + cc % 6;
                T[seq].v = Ke[rr + 18 * cc]; T[seq].seq = seq;
            }
    }
    if (bad != INT64_MAX) {
        double p[9], prop[8], Ke[324];
        for (int k = 0; k < 3; ++k) { int nd = tc[k * nt + bad]; p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd]; }
        for (int c = 0; c < 8; ++c) prop[c] = tp[c * nt + bad];
        bad_st = or_t3_global(p, prop, k6rot, Ke);
        free(T); *bad_elem = nq + bad; return bad_st;
    }
#pragma omp parallel for schedule(static) reduction(min : bad)
    for (int64_t e = 0; e < ntet; ++e) {
        double X[30], Ke[900]; int nd[10];
        for (int k = 0; k < 10; ++k) {
            nd[k] = ec[k * ntet + e];
            X[k] = x[nd[k]]; X[k + 10] = y[nd[k]]; X[k + 20] = z[nd[k]];
        }
        if (or_tet10(X, ep[e], ep[ntet + e], Ke) != OR_OK) { if (e < bad) bad = e; continue; }
        int64_t seq = tet_base + e * 900;
        for (int cc = 0; cc < 30; ++cc)
            for (int rr = 0; rr < 30; ++rr, ++seq) {
                T[seq].row = 6 * (int64_t)nd[rr / 3] + rr % 3; T[seq].col = 6 * nd[cc / 3] + cc % 3;
                T[seq].v = Ke[rr + 30 * cc]; T[seq].seq = seq;
            }
    }
    if (bad != INT64_MAX) {
        double X[30], Ke[900];
        for (int k = 0; k < 10; ++k) { int nd = ec[k * ntet + bad]; X[k] = x[nd]; X[k + 10] = y[nd]; X[k + 20] = z[nd]; }
        bad_st = or_tet10(X, ep[bad], ep[ntet + bad], Ke);
        free(T); *bad_elem = nq + nt + bad; return bad_st;
    }
    /* sparse(I,J,V): order by (row, col, triplet order) -- a stable counting sort
     * by row, then per row a sort by (col, seq) -- and sum duplicates in that order */
    const int64_t ndof = 6 * n_nodes;
    int64_t *rstart = (int64_t *)calloc((size_t)ndof + 1, sizeof(int64_t));
    trip_t *S = (trip_t *)malloc(sizeof(trip_t) * (ntrip > 0 ? ntrip : 1));
    int64_t *rnnz = (int64_t *)calloc((size_t)ndof + 1, sizeof(int64_t));
    if (!rstart || !S || !rnnz) { free(T); free(rstart); free(S); free(rnnz); return OR_ERR_OOM; }
    for (int64_t i = 0; i < ntrip; ++i) rstart[T[i].row + 1]++;
    for (int64_t r = 0; r < ndof; ++r) rstart[r + 1] += rstart[r];
    {
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (ndof > 0 ? ndof : 1));
        memcpy(fill, rstart, sizeof(int64_t) * ndof);
        for (int64_t i = 0; i < ntrip; ++i) S[fill[T[i].row]++] = T[i];
        free(fill);
    }
    free(T);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < ndof; ++r) {
        const int64_t a = rstart[r], b = rstart[r + 1];
        if (b > a) qsort(S + a, (size_t)(b - a), sizeof(trip_t), trip_cmp);
        int64_t c = 0;
        for (int64_t i = a; i < b; ++i)
            if (i == a || S[i].col != S[i - 1].col) ++c;
        rnnz[r + 1] = c;
    }
    for (int64_t r = 0; r < ndof; ++r) rnnz[r + 1] += rnnz[r];
    const int64_t nnz = rnnz[ndof];
    int32_t *ci = (int32_t *)malloc(sizeof(int32_t) * (nnz ? nnz : 1));
    double *cv = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    if (!ci || !cv) { free(S); free(rstart); free(rnnz); free(ci); free(cv); return OR_ERR_OOM; }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < ndof; ++r) {
        int64_t k = rnnz[r] - 1;
        for (int64_t i = rstart[r]; i < rstart[r + 1]; ++i) {
            if (i == rstart[r] || S[i].col != S[i - 1].col) {
                ++k; ci[k] = S[i].col; cv[k] = S[i].v;
            } else {
                cv[k] += S[i].v; /* duplicates summed in triplet order */
            }
        }
    }
    free(S); free(rstart);
    *out_nnz = nnz; *row_ptr = rnnz; *col = ci; *val = cv;
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* BC partition: SPC + AUTOSPC (zero/negative diagonal) + grid-point            */
/* singularity test (src/solver/boundary_conditions.jl:1028-1138, 745-845)     */
/* ------------------------------------------------------------------------- */
static double csr_get(const int64_t *rp, const int32_t *ci, const double *v, int64_t i, int64_t j) {
    int64_t lo = rp[i], hi = rp[i + 1] - 1;
    while (lo <= hi) {
        int64_t m = (lo + hi) / 2;
        if (ci[m] == j) return v[m];
        if (ci[m] < j) lo = m + 1; else hi = m - 1;
    }
    return 0.0;
}

/* cyclic Jacobi for symmetric 3x3 (A column-major), eigenvalues ascending */
void or_eig3(const double A[9], double w[3], double V[9]) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = A[i + 3 * j];
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        double sc = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off <= 1e-300 || off <= 1e-18 * sc) break;
        for (int p = 0; p < 2; ++p) for (int q = p + 1; q < 3; ++q) {
            if (a[p][q] == 0.0) continue;
            double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
            double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
            double c = 1 / sqrt(t * t + 1), s = t * c;
            for (int k = 0; k < 3; ++k) { double akp = a[k][p], akq = a[k][q]; a[k][p] = c * akp - s * akq; a[k][q] = s * akp + c * akq; }
            for (int k = 0; k < 3; ++k) { double apk = a[p][k], aqk = a[q][k]; a[p][k] = c * apk - s * aqk; a[q][k] = s * apk + c * aqk; }
            for (int k = 0; k < 3; ++k) { double vkp = v[k][p], vkq = v[k][q]; v[k][p] = c * vkp - s * vkq; v[k][q] = s * vkp + c * vkq; }
        }
    }
    int idx[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i) for (int j = i + 1; j < 3; ++j) if (a[idx[j]][idx[j]] < a[idx[i]][idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int k = 0; k < 3; ++k) { w[k] = a[idx[k]][idx[k]]; for (int r = 0; r < 3; ++r) V[r + 3 * k] = v[r][idx[k]]; }
}

int64_t or_bc_partition(int64_t ndof, const int64_t *rp, const int32_t *ci, const double *v,
                        uint8_t *fixed, double eps_zero) {
    int64_t added = 0;
    /* diagonal AUTOSPC with grid-point processing on: only exact-zero and negative
     * diagonals are removed here (local_relative_* = false) */
    for (int64_t i = 0; i < ndof; ++i) {
        double d = csr_get(rp, ci, v, i, i);
        if (!fixed[i] && (d == 0.0 || d < 0)) { fixed[i] = 1; ++added; }
    }
    /* grid point singularity on 3x3 translational and rotational blocks */
    int64_t nn = ndof / 6;
    for (int64_t n = 0; n < nn; ++n) {
        for (int off = 0; off <= 3; off += 3) {
            int fl[3], m = 0;
            for (int a = 0; a < 3; ++a) if (!fixed[6 * n + off + a]) fl[m++] = a;
            if (m < 1) continue;
            double S[9] = {0}, w[3], V[9];
            for (int ii = 0; ii < m; ++ii) for (int jj = 0; jj < m; ++jj) {
                double kij = csr_get(rp, ci, v, 6 * n + off + fl[ii], 6 * n + off + fl[jj]);
                double kji = csr_get(rp, ci, v, 6 * n + off + fl[jj], 6 * n + off + fl[ii]);
                S[ii + 3 * jj] = (kij + kji) / 2;
            }
            double vals[3], vecs[9];
            if (m == 3) { or_eig3(S, w, V); memcpy(vals, w, sizeof(w)); memcpy(vecs, V, sizeof(V)); }
            else if (m == 2) {
                double a = S[0], b = S[3], d = S[4];
                double tr = 0.5 * (a + d), df = 0.5 * (a - d), r = sqrt(df * df + b * b);
                vals[0] = tr - r; vals[1] = tr + r;
                for (int k = 0; k < 2; ++k) {
                    double vx, vy;
                    if (b != 0.0) { vx = b; vy = vals[k] - a; }
                    else { vx = (k == 0) == (a <= d) ? 1.0 : 0.0; vy = 1.0 - vx; }
                    double nr = hypot(vx, vy); vecs[0 + 3 * k] = vx / nr; vecs[1 + 3 * k] = vy / nr;
                }
            } else { vals[0] = S[0]; vecs[0] = 1.0; }
            double lmax = 0; for (int k = 0; k < m; ++k) if (fabs(vals[k]) > lmax) lmax = fabs(vals[k]);
            if (!(lmax > 0.0)) continue;
            int claimed[3] = {0, 0, 0};
            for (int k = 0; k < m; ++k) {
                double ratio = fabs(vals[k]) / lmax;
                if (!(ratio < eps_zero)) continue;
                int best_a = -1; double best_w = -1.0;
                for (int ii = 0; ii < m; ++ii) {
                    double wv = fabs(vecs[ii + 3 * k]); int a = fl[ii];
                    if (claimed[a]) continue;
                    if (wv > best_w + 1e-12) { best_w = wv; best_a = a; }
                }
                if (best_a < 0) continue;
                int64_t g = 6 * n + off + best_a;
                if (fixed[g]) continue;
                claimed[best_a] = 1; fixed[g] = 1; ++added;
            }
        }
    }
    return added;
}

/* ------------------------------------------------------------------------- */
/* reference sparse Cholesky: orderings + up-looking factor (textbook,          */
/* stands in for CHOLMOD; src/solver/boundary_conditions.jl:1656-1714)          */
/* ------------------------------------------------------------------------- */
typedef struct { int64_t n; int64_t *p; int32_t *i; double *x; } csc_t;

/* reverse Cuthill-McKee on a symmetric pattern (adjacency lists, no diagonal) */
static void rcm_order(int64_t n, const int64_t *ap, const int32_t *ai, int32_t *perm) {
    int64_t *deg = malloc(sizeof(int64_t) * n); uint8_t *vis = calloc(n, 1);
    int32_t *q = malloc(sizeof(int32_t) * n);
    for (int64_t v = 0; v < n; ++v) deg[v] = ap[v + 1] - ap[v];
    int64_t head = 0, tail = 0;
    for (;;) {
        int64_t start = -1;
        for (int64_t v = 0; v < n; ++v) if (!vis[v] && (start < 0 || deg[v] < deg[start])) start = v;
        if (start < 0) break;
        vis[start] = 1; q[tail++] = (int32_t)start;
        while (head < tail) {
            int32_t u = q[head++];
            int64_t b = tail;
            for (int64_t p = ap[u]; p < ap[u + 1]; ++p) { int32_t w = ai[p]; if (!vis[w]) { vis[w] = 1; q[tail++] = w; } }
            /* sort new ones by degree (insertion sort) */
            for (int64_t s = b + 1; s < tail; ++s) { int32_t t = q[s]; int64_t r = s - 1; while (r >= b && deg[q[r]] > deg[t]) { q[r + 1] = q[r]; --r; } q[r + 1] = t; }
        }
    }
    for (int64_t k = 0; k < n; ++k) perm[k] = q[n - 1 - k];
    free(deg); free(vis); free(q);
}

/* geometric nested dissection on supervariables (grid points) */
typedef struct { const double *xyz; const int64_t *ap; const int32_t *ai; int32_t *mark; int32_t *out; int64_t nout; } nd_ctx;
static int nd_cmp_axis; static const double *nd_xyz;
static int nd_cmp(const void *a, const void *b) {
    double u = nd_xyz[3 * *(const int32_t *)a + nd_cmp_axis], w = nd_xyz[3 * *(const int32_t *)b + nd_cmp_axis];
    return u < w ? -1 : (u > w);
}
/* split set at the median of axis ax; count the one-sided separator (A-side points
 * adjacent to B); returns the separator size, or -1 if the split is empty */
static int64_t nd_try(nd_ctx *c, int32_t *set, int64_t n, int ax) {
    nd_cmp_axis = ax; nd_xyz = c->xyz;
    qsort(set, (size_t)n, sizeof(int32_t), nd_cmp);
    int64_t h = n / 2;
    for (int64_t k = 0; k < n; ++k) c->mark[set[k]] = k < h ? 1 : 2;
    int64_t ns = 0;
    for (int64_t k = 0; k < h; ++k) {
        int32_t v = set[k];
        for (int64_t p = c->ap[v]; p < c->ap[v + 1]; ++p) if (c->mark[c->ai[p]] == 2) { ++ns; break; }
    }
    for (int64_t k = 0; k < n; ++k) c->mark[set[k]] = 0;
    return (ns == h) ? -1 : ns;
}

static void nd_rec(nd_ctx *c, int32_t *set, int64_t n, int32_t tag) {
    if (n <= 64) { for (int64_t k = 0; k < n; ++k) c->out[c->nout++] = set[k]; return; }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t k = 0; k < n; ++k) for (int a = 0; a < 3; ++a) {
        double v = c->xyz[3 * set[k] + a]; if (v < lo[a]) lo[a] = v; if (v > hi[a]) hi[a] = v;
    }
    /* axis with the smallest median-plane separator */
    int ax = -1; int64_t best = -1;
    for (int a = 0; a < 3; ++a) {
        if (!(hi[a] - lo[a] > 0)) continue;
        int64_t sz = nd_try(c, set, n, a);
        if (sz < 0) continue;
        if (ax < 0 || sz < best) { ax = a; best = sz; }
    }
    if (ax < 0) { for (int64_t k = 0; k < n; ++k) c->out[c->nout++] = set[k]; return; }
    nd_cmp_axis = ax; nd_xyz = c->xyz;
    qsort(set, (size_t)n, sizeof(int32_t), nd_cmp);
    int64_t h = n / 2;
    for (int64_t k = 0; k < n; ++k) c->mark[set[k]] = k < h ? 1 : 2;
    int32_t *A = malloc(sizeof(int32_t) * n), *B = malloc(sizeof(int32_t) * n), *S = malloc(sizeof(int32_t) * n);
    int64_t na = 0, nb = 0, ns = 0;
    for (int64_t k = 0; k < h; ++k) {
        int32_t v = set[k]; int sep = 0;
        for (int64_t p = c->ap[v]; p < c->ap[v + 1]; ++p) if (c->mark[c->ai[p]] == 2) { sep = 1; break; }
        if (sep) S[ns++] = v; else A[na++] = v;
    }
    for (int64_t k = h; k < n; ++k) B[nb++] = set[k];
    for (int64_t k = 0; k < n; ++k) c->mark[set[k]] = 0;
    if (na == 0 || nb == 0) { for (int64_t k = 0; k < n; ++k) c->out[c->nout++] = set[k]; free(A); free(B); free(S); return; }
    nd_rec(c, A, na, tag); nd_rec(c, B, nb, tag);
    for (int64_t k = 0; k < ns; ++k) c->out[c->nout++] = S[k];
    free(A); free(B); free(S);
}

/* y <- (L L')^{-1} y for the column-compressed Cholesky factor (diagonal first) */
static void ltsolve(int64_t n, const int64_t *Lp, const int32_t *Li, const double *Lx, double *y) {
    for (int64_t j = 0; j < n; ++j) { y[j] /= Lx[Lp[j]]; for (int64_t p = Lp[j] + 1; p < Lp[j + 1]; ++p) y[Li[p]] -= Lx[p] * y[j]; }
    for (int64_t j = n - 1; j >= 0; --j) { for (int64_t p = Lp[j] + 1; p < Lp[j + 1]; ++p) y[j] -= Lx[p] * y[Li[p]]; y[j] /= Lx[Lp[j]]; }
}

static int solve_impl(int64_t ndof, const int64_t *rp, const int32_t *ci, const double *val,
                      const uint8_t *fixed, int order_kind, const double *xyz_nodes, const int64_t *given_order,
                      int nrhs, const double *b, double *u, double *relres,
                      int64_t *nnz_L_out, double *flops_out, int64_t *n_small, int refine, double *ratios_out);

int or_solve(int64_t ndof, const int64_t *rp, const int32_t *ci, const double *val,
             const uint8_t *fixed, int order_kind, const double *xyz_nodes,
             int nrhs, const double *b, double *u, double *relres,
             int64_t *nnz_L_out, double *flops_out, int64_t *n_small, int refine) {
    return solve_impl(ndof, rp, ci, val, fixed, order_kind, xyz_nodes, NULL, nrhs, b, u, relres, nnz_L_out, flops_out,
                      n_small, refine, NULL);
}

int or_pivot_ratios(int64_t ndof, const int64_t *rp, const int32_t *ci, const double *val, const uint8_t *fixed,
                    const int64_t *order, double *ratios, int64_t *n_small) {
    int64_t nz = 0; double fl = 0;
    return solve_impl(ndof, rp, ci, val, fixed, 2, NULL, order, 0, NULL, NULL, NULL, &nz, &fl, n_small, 0, ratios);
}

static int solve_impl(int64_t ndof, const int64_t *rp, const int32_t *ci, const double *val,
                      const uint8_t *fixed, int order_kind, const double *xyz_nodes, const int64_t *given_order,
                      int nrhs, const double *b, double *u, double *relres,
                      int64_t *nnz_L_out, double *flops_out, int64_t *n_small, int refine, double *ratios_out) {
    /* free DOF numbering (ascending global DOF) */
    int64_t *loc = malloc(sizeof(int64_t) * ndof); int64_t n = 0;
    for (int64_t i = 0; i < ndof; ++i) loc[i] = fixed[i] ? -1 : n++;
    int64_t *glob = malloc(sizeof(int64_t) * (n ? n : 1));
    for (int64_t i = 0; i < ndof; ++i) if (loc[i] >= 0) glob[loc[i]] = i;
    /* symmetric adjacency (off-diagonal) of K_ff */
    int64_t *ap = calloc(n + 1, sizeof(int64_t));
    for (int64_t r = 0; r < n; ++r) { int64_t g = glob[r]; for (int64_t p = rp[g]; p < rp[g + 1]; ++p) { int64_t c = loc[ci[p]]; if (c >= 0 && c != r) ap[r + 1]++; } }
    for (int64_t r = 0; r < n; ++r) ap[r + 1] += ap[r];
    int32_t *ai = malloc(sizeof(int32_t) * (ap[n] ? ap[n] : 1));
    { int64_t *fill = malloc(sizeof(int64_t) * (n ? n : 1)); for (int64_t r = 0; r < n; ++r) fill[r] = ap[r];
      for (int64_t r = 0; r < n; ++r) { int64_t g = glob[r]; for (int64_t p = rp[g]; p < rp[g + 1]; ++p) { int64_t c = loc[ci[p]]; if (c >= 0 && c != r) ai[fill[r]++] = (int32_t)c; } }
      free(fill); }
    /* ordering */
    int32_t *perm = malloc(sizeof(int32_t) * (n ? n : 1));
    if (order_kind == 2) {
        /* the caller's elimination order, as global DOFs (position k eliminates
         * DOF given_order[k]); must be a permutation of the free DOFs */
        for (int64_t k = 0; k < n; ++k) {
            int64_t g = given_order[k];
            perm[k] = (g >= 0 && g < ndof && loc[g] >= 0) ? (int32_t)loc[g] : 0;
        }
    } else if (order_kind == 1 && xyz_nodes) {
        /* supervariables = grid points: order grid points by ND, expand to their free DOFs */
        int64_t nn = ndof / 6;
        int64_t *np_ = calloc(nn + 1, sizeof(int64_t));
        /* node adjacency from DOF adjacency */
        int32_t *mark = calloc(nn, sizeof(int32_t));
        int64_t cap = ap[n] / 2 + nn + 16, used = 0;
        int32_t *na = malloc(sizeof(int32_t) * cap);
        for (int64_t v = 0; v < nn; ++v) {
            np_[v] = used;
            int32_t stamp = (int32_t)v + 1;
            mark[v] = stamp;
            for (int a = 0; a < 6; ++a) { int64_t r = loc[6 * v + a]; if (r < 0) continue;
                for (int64_t p = ap[r]; p < ap[r + 1]; ++p) { int64_t w = glob[ai[p]] / 6; if (mark[w] != stamp) { mark[w] = stamp;
                    if (used >= cap) { cap *= 2; na = realloc(na, sizeof(int32_t) * cap); } na[used++] = (int32_t)w; } } }
        }
        np_[nn] = used;
        int32_t *nodes = malloc(sizeof(int32_t) * (nn ? nn : 1)), *out = malloc(sizeof(int32_t) * (nn ? nn : 1));
        int64_t cnt = 0;
        for (int64_t v = 0; v < nn; ++v) { int any = 0; for (int a = 0; a < 6; ++a) if (loc[6 * v + a] >= 0) any = 1; if (any) nodes[cnt++] = (int32_t)v; }
        for (int64_t v = 0; v < nn; ++v) mark[v] = 0;
        nd_ctx c = {xyz_nodes, np_, na, mark, out, 0};
        nd_rec(&c, nodes, cnt, 0);
        int64_t k = 0;
        for (int64_t t = 0; t < c.nout; ++t) for (int a = 0; a < 6; ++a) { int64_t r = loc[6 * (int64_t)out[t] + a]; if (r >= 0) perm[k++] = (int32_t)r; }
        free(np_); free(mark); free(na); free(nodes); free(out);
    } else {
        rcm_order(n, ap, ai, perm);
    }
    int32_t *pinv = malloc(sizeof(int32_t) * (n ? n : 1));
    for (int64_t k = 0; k < n; ++k) pinv[perm[k]] = (int32_t)k;
    /* C = P A P' upper triangle in CSC, values from the original upper triangle
     * (Symmetric(K_ff) reads the upper triangle) */
    int64_t *cp = calloc(n + 1, sizeof(int64_t));
    for (int64_t r = 0; r < n; ++r) { int64_t g = glob[r];
        for (int64_t p = rp[g]; p < rp[g + 1]; ++p) { int64_t c = loc[ci[p]]; if (c < r) continue; /* original upper incl diag */
            int64_t pr = pinv[r], pc = pinv[c]; int64_t col = pr > pc ? pr : pc; cp[col + 1]++; } }
    for (int64_t j = 0; j < n; ++j) cp[j + 1] += cp[j];
    int32_t *cix = malloc(sizeof(int32_t) * (cp[n] ? cp[n] : 1)); double *cx = malloc(sizeof(double) * (cp[n] ? cp[n] : 1));
    { int64_t *fill = malloc(sizeof(int64_t) * (n ? n : 1)); for (int64_t j = 0; j < n; ++j) fill[j] = cp[j];
      for (int64_t r = 0; r < n; ++r) { int64_t g = glob[r];
        for (int64_t p = rp[g]; p < rp[g + 1]; ++p) { int64_t c = loc[ci[p]]; if (c < r) continue;
            int64_t pr = pinv[r], pc = pinv[c]; int64_t col = pr > pc ? pr : pc, row = pr > pc ? pc : pr;
            cix[fill[col]] = (int32_t)row; cx[fill[col]++] = val[p]; } }
      free(fill); }
    /* elimination tree (Liu) on C upper */
    int32_t *parent = malloc(sizeof(int32_t) * (n ? n : 1)), *anc = malloc(sizeof(int32_t) * (n ? n : 1));
    for (int64_t k = 0; k < n; ++k) { parent[k] = -1; anc[k] = -1;
        for (int64_t p = cp[k]; p < cp[k + 1]; ++p) { int32_t i = cix[p];
            for (; i != -1 && i < k; ) { int32_t nx = anc[i]; anc[i] = (int32_t)k; if (nx == -1) { parent[i] = (int32_t)k; break; } i = nx; } } }
    /* column counts via row subtrees (ereach) */
    int64_t *cc = calloc(n + 1, sizeof(int64_t)); int32_t *flag = malloc(sizeof(int32_t) * (n ? n : 1)), *stk = malloc(sizeof(int32_t) * (n ? n : 1));
    for (int64_t k = 0; k < n; ++k) flag[k] = -1;
    for (int64_t k = 0; k < n; ++k) { flag[k] = (int32_t)k; cc[k]++;
        for (int64_t p = cp[k]; p < cp[k + 1]; ++p) { int32_t i = cix[p]; if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) { flag[i] = (int32_t)k; cc[i]++; } } }
    int64_t *Lp = malloc(sizeof(int64_t) * (n + 1)); Lp[0] = 0;
    double fl = 0;
    for (int64_t k = 0; k < n; ++k) { Lp[k + 1] = Lp[k] + cc[k]; fl += (double)cc[k] * (double)cc[k]; }
    int64_t nzL = Lp[n];
    int32_t *Li = malloc(sizeof(int32_t) * (nzL ? nzL : 1)); double *Lx = malloc(sizeof(double) * (nzL ? nzL : 1));
    double *xw = calloc(n ? n : 1, sizeof(double)); int64_t *cnext = malloc(sizeof(int64_t) * (n ? n : 1));
    for (int64_t k = 0; k < n; ++k) { cnext[k] = Lp[k]; flag[k] = -1; }
    int st = OR_OK;
    for (int64_t k = 0; k < n && st == OR_OK; ++k) {
        /* nonzero pattern of row k of L: ereach, topological order */
        int64_t top = n; flag[k] = (int32_t)k;
        for (int64_t p = cp[k]; p < cp[k + 1]; ++p) { int32_t i = cix[p]; if (i > k) continue; xw[i] += cx[p]; if (i == k) continue;
            int64_t len = 0; for (; flag[i] != k; i = parent[i]) { stk[len++] = i; flag[i] = (int32_t)k; }
            while (len > 0) stk[--top] = stk[--len]; }
        double d = xw[k]; xw[k] = 0;
        for (; top < n; ++top) { int32_t i = stk[top]; double lki = xw[i] / Lx[Lp[i]]; xw[i] = 0;
            for (int64_t p = Lp[i] + 1; p < cnext[i]; ++p) xw[Li[p]] -= Lx[p] * lki;
            d -= lki * lki; int64_t p2 = cnext[i]++; Li[p2] = (int32_t)k; Lx[p2] = lki; }
        if (!(d > 0)) { st = OR_ERR_NOT_SPD; break; }
        int64_t p = cnext[k]++; Li[p] = (int32_t)k; Lx[p] = sqrt(d);
    }
    /* column i of L is filled at cnext[i]; its first write is the diagonal (step
     * k == i precedes every step k > i), so Lx[Lp[i]] is L(i,i) */
    int64_t smallp = 0;
    if (st == OR_OK) {
        for (int64_t k = 0; k < n; ++k) { double lkk = Lx[Lp[k]]; int64_t g = glob[perm[k]];
            double kd = fabs(csr_get(rp, ci, val, g, g));
            double ratio = kd > 1e-30 ? lkk * lkk / kd : 1.0; if (ratio < 1e-7) ++smallp;
            if (ratios_out) ratios_out[k] = ratio; }
        double *y = malloc(sizeof(double) * (n ? n : 1));
        long double *xl = refine > 0 ? malloc(sizeof(long double) * (n ? n : 1)) : NULL;
        for (int r = 0; r < nrhs; ++r) {
            const double *br = b + (size_t)ndof * r; double *ur = u + (size_t)ndof * r;
            for (int64_t k = 0; k < n; ++k) y[k] = br[glob[perm[k]]];
            ltsolve(n, Lp, Li, Lx, y);
            for (int64_t i = 0; i < ndof; ++i) ur[i] = 0.0;
            for (int64_t k = 0; k < n; ++k) ur[glob[perm[k]]] = y[k];
            if (refine > 0) {
                /* iterative refinement with the residual and the iterate in long
                 * double (64-bit significand): converges to the solution correctly
                 * rounded to double up to a few units, independent of cond(K_ff)
                 * while cond(K_ff) * 2^-53 << 1 (Moler 1967; Higham, Accuracy and
                 * Stability of Numerical Algorithms, 2nd ed., section 12.1). */
                for (int64_t k = 0; k < n; ++k) xl[k] = y[k];
                for (int it = 0; it < refine; ++it) {
                    for (int64_t k = 0; k < n; ++k) { int64_t g = glob[perm[k]]; long double s = br[g];
                        for (int64_t p = rp[g]; p < rp[g + 1]; ++p) { int64_t c = loc[ci[p]]; if (c >= 0) s -= (long double)val[p] * xl[pinv[c]]; }
                        y[k] = (double)s; }
                    ltsolve(n, Lp, Li, Lx, y);
                    long double dmax = 0, xmax = 0;
                    for (int64_t k = 0; k < n; ++k) { xl[k] += y[k]; long double a = fabsl(y[k]), bx = fabsl(xl[k]);
                        if (a > dmax) dmax = a; if (bx > xmax) xmax = bx; }
                    if (dmax <= 1e-19L * xmax) break;
                }
                for (int64_t k = 0; k < n; ++k) ur[glob[perm[k]]] = (double)xl[k];
            }
            /* relative residual ||K_ff u - F_ff|| / max(||F_ff||, 1e-30), :1859-1868
             * (accumulated in long double when refining, so the figure is the
             * residual of the returned double vector itself) */
            double rn = 0, fn = 0;
            for (int64_t r2 = 0; r2 < n; ++r2) { int64_t g = glob[r2]; double e;
                if (refine > 0) { long double s = 0; for (int64_t p = rp[g]; p < rp[g + 1]; ++p) if (loc[ci[p]] >= 0) s += (long double)val[p] * ur[ci[p]];
                                  e = (double)(s - br[g]); }
                else { double s = 0; for (int64_t p = rp[g]; p < rp[g + 1]; ++p) if (loc[ci[p]] >= 0) s += val[p] * ur[ci[p]];
                       e = s - br[g]; }
                rn += e * e; fn += br[g] * br[g]; }
            relres[r] = sqrt(rn) / fmax(sqrt(fn), 1e-30);
        }
        free(y); free(xl);
    }
    if (nnz_L_out) *nnz_L_out = nzL;
    if (flops_out) *flops_out = fl;
    if (n_small) *n_small = smallp;
    free(loc); free(glob); free(ap); free(ai); free(perm); free(pinv); free(cp); free(cix); free(cx);
    free(parent); free(anc); free(cc); free(flag); free(stk); free(Lp); free(Li); free(Lx); free(xw); free(cnext);
    return st;
}


/* ======================================================================== */
/* Load vector of one load set: FORCE / MOMENT on grid points, uniform PLOAD4 on
 * CQUAD4 / CTRIA3, basic coordinates.  _resolve_loads!, src/solver/loads.jl. */

/* LinearAlgebra.generic_norm2 for a 3-vector (below the BLAS nrm2 cutoff):
 * maxabs = normInf; plain sum of squares unless length*maxabs^2 is not finite
 * or maxabs^2 is zero, then the sum of (|v_i|/maxabs)^2 times maxabs. */
double or_jl_norm3(const double v[3]) {
    double mx = fabs(v[0]);
    for (int i = 1; i < 3; ++i) {
        const double a = fabs(v[i]);
        if (!(isnan(mx) || mx > a)) mx = a;        /* ifelse(isnan(m) | (m > a), m, a) */
    }
    if (mx == 0.0 || isinf(mx)) return mx;
    if (isfinite(3.0 * mx * mx) && mx * mx != 0.0) {
        double sum = v[0] * v[0];
        sum += v[1] * v[1];
        sum += v[2] * v[2];
        return sqrt(sum);
    }
    double sum = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double q = fabs(v[i]) / mx;
        sum += q * q;
    }
    return mx * sqrt(sum);
}

/* LinearAlgebra.normalize -> __normalize!(a, nrm): delta = inv(prevfloat(Inf));
 * nrm >= delta: a * inv(nrm); else a * (eps/delta), then * inv(nrm * (eps/delta)). */
void or_jl_normalize3(double v[3]) {
    const double nrm = or_jl_norm3(v);
    const double delta = 1.0 / DBL_MAX;
    if (nrm >= delta) {
        const double inv = 1.0 / nrm;
        for (int i = 0; i < 3; ++i) v[i] *= inv;
    } else {
        const double ed = DBL_EPSILON / delta;
        for (int i = 0; i < 3; ++i) v[i] *= ed;
        const double inv = 1.0 / (nrm * ed);
        for (int i = 0; i < 3; ++i) v[i] *= inv;
    }
}

int or_load_vector(int64_t n_nodes, const double *x, const double *y, const double *z, int64_t nq,
                   const int32_t *quad_conn, int64_t nt, const int32_t *tria_conn, int64_t n_pload4,
                   const int32_t *pl_elem, const double *pl_p, const double *pl_dir3, int64_t n_force,
                   const int32_t *f_node, const double *f_mag, const double *f_dir3, int64_t n_moment,
                   const int32_t *m_node, const double *m_mag, const double *m_dir3, double scale,
                   const double *shell_areal, int64_t n_grav, const double *grav_a, const double *grav_n3,
                   double *F) {
    if (n_nodes <= 0) return OR_ERR_BAD_ARG;
    for (int64_t i = 0; i < 6 * n_nodes; ++i) F[i] = 0.0;
    /* PLOAD4 on shells, loads.jl:541-587 (the CSHELL branch) */
    for (int64_t i = 0; i < n_pload4; ++i) {
        const int64_t e = pl_elem[i];
        if (e < 0 || e >= nq + nt) return OR_ERR_BAD_ARG;
        int nids[4], nn;
        if (e < nq) { nn = 4; for (int k = 0; k < 4; ++k) nids[k] = quad_conn[e + nq * k]; }
        else { nn = 3; for (int k = 0; k < 3; ++k) nids[k] = tria_conn[(e - nq) + nt * k]; }
        double Xc[4][3];
        for (int k = 0; k < nn; ++k) { Xc[k][0] = x[nids[k]]; Xc[k][1] = y[nids[k]]; Xc[k][2] = z[nids[k]]; }
        double a[3], b[3], normal_vec[3];
        if (nn == 4) { v3sub(Xc[2], Xc[0], a); v3sub(Xc[3], Xc[1], b); }   /* cross(X3-X1, X4-X2) */
        else { v3sub(Xc[1], Xc[0], a); v3sub(Xc[2], Xc[0], b); }          /* cross(v1, v2) */
        v3cross(a, b, normal_vec);
        const double area = 0.5 * or_jl_norm3(normal_vec);
        if (!(area > 1e-30)) continue;                                   /* area > 1e-30 || continue */
        double load_dir[3];
        int have_dir = 0;
        if (pl_dir3) {                                                   /* haskey(pload, "N") */
            const double *dir_vec = pl_dir3 + 3 * i;
            const double n_dir = or_jl_norm3(dir_vec);
            if (n_dir > 1e-30) {
                for (int c = 0; c < 3; ++c) load_dir[c] = dir_vec[c] / n_dir;   /* dir_vec ./ n_dir */
                have_dir = 1;
            }
        }
        if (!have_dir) {
            for (int c = 0; c < 3; ++c) load_dir[c] = normal_vec[c];
            or_jl_normalize3(load_dir);
        }
        const double tf = area * pl_p[i] * scale;
        double f_node[3];
        for (int c = 0; c < 3; ++c) f_node[c] = load_dir[c] * (tf / (double)nn);
        for (int k = 0; k < nn; ++k)
            for (int c = 0; c < 3; ++c) F[6 * (int64_t)nids[k] + c] += f_node[c];
    }
    /* GRAV on shells, loads.jl:680-684, 730-772: per card the acceleration
     * (A scale) N, per shell (element order: CQUAD4s, then CTRIA3s) with a mass
     * (rho T + NSM, given per shell; 0 skips it) the lumped share
     * (areal * area / nn) .* g on each node's translations, area = norm of the
     * diagonal (quad) / edge (tria) cross product over 2 */
    for (int64_t gi = 0; gi < n_grav; ++gi) {
        double gv[3];
        for (int c = 0; c < 3; ++c) gv[c] = (grav_a[gi] * scale) * grav_n3[3 * gi + c];
        for (int64_t e = 0; e < nq + nt; ++e) {
            if (!shell_areal || shell_areal[e] == 0.0) continue;
            int nids[4], nn;
            if (e < nq) { nn = 4; for (int k = 0; k < 4; ++k) nids[k] = quad_conn[e + nq * k]; }
            else { nn = 3; for (int k = 0; k < 3; ++k) nids[k] = tria_conn[(e - nq) + nt * k]; }
            double Xc[4][3];
            for (int k = 0; k < nn; ++k) { Xc[k][0] = x[nids[k]]; Xc[k][1] = y[nids[k]]; Xc[k][2] = z[nids[k]]; }
            double a[3], b[3], cr[3];
            if (nn == 4) { v3sub(Xc[2], Xc[0], a); v3sub(Xc[3], Xc[1], b); }
            else { v3sub(Xc[1], Xc[0], a); v3sub(Xc[2], Xc[0], b); }
            v3cross(a, b, cr);
            const double area = 0.5 * or_jl_norm3(cr);
            const double total_mass = shell_areal[e] * area;
            for (int k = 0; k < nn; ++k)
                for (int c = 0; c < 3; ++c) F[6 * (int64_t)nids[k] + c] += (total_mass / nn) * gv[c];
        }
    }
    /* FORCE / MOMENT into raw_forces[gid] (zeros(6) on first use), loads.jl:519-539;
     * applied after the distributed loads, :1338-1340 */
    double *raw = (double *)calloc(6 * (size_t)n_nodes, sizeof(double));
    unsigned char *has = (unsigned char *)calloc((size_t)n_nodes, 1);
    if (!raw || !has) { free(raw); free(has); return OR_ERR_BAD_ARG; }
    for (int64_t i = 0; i < n_force; ++i) {
        const int64_t g = f_node[i];
        if (g < 0 || g >= n_nodes) { free(raw); free(has); return OR_ERR_BAD_ARG; }
        has[g] = 1;
        for (int c = 0; c < 3; ++c) raw[6 * g + c] += f_dir3[3 * i + c] * f_mag[i] * scale;
    }
    for (int64_t i = 0; i < n_moment; ++i) {
        const int64_t g = m_node[i];
        if (g < 0 || g >= n_nodes) { free(raw); free(has); return OR_ERR_BAD_ARG; }
        has[g] = 1;
        for (int c = 0; c < 3; ++c) raw[6 * g + 3 + c] += m_dir3[3 * i + c] * m_mag[i] * scale;
    }
    for (int64_t g = 0; g < n_nodes; ++g)
        if (has[g])
            for (int c = 0; c < 6; ++c) F[6 * g + c] += raw[6 * g + c];
    free(raw); free(has);
    return 0;
}
