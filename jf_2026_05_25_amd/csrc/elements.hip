// Batched element stiffness generation for gfx950 and the deterministic
// segmented-reduction assembly.
//
// One 64-lane wavefront owns one element.  Every lane owns one 3x3 block of the
// element matrix (CQUAD4: 8x8 blocks of a 24x24; CTRIA3: 6x6 of 18x18; CTETRA10:
// 10x10 of 30x30, two blocks per lane).  Gauss-point derivative rows, internal-mode
// couplings, the MacNeal plate block and the drilling rows are built once per
// element into LDS by a few lanes, then every lane forms its block, applies the
// frame rotation R' * blk * R and stages the global block in LDS so the wavefront
// writes the whole element matrix with coalesced stores.
//
// The per-entry arithmetic follows the statement order of the reference kernels
// (ts_mul! / ts_mul_At_add!, src/FEMKernels.jl:649-675); citations per step below.
#include "jfem_internal.h"

#include <cfloat>
#include <climits>

namespace {

constexpr double kEps = 2.220446049250313e-16;
constexpr int kWave = 64;
constexpr int kElemsPerBlock = 2;  // wavefronts (= elements) per workgroup

// Variants of one mesh (design variants): a launch over V section sets runs
// variant blockIdx.y with its section records at blockIdx.y * prop, its element
// matrices at blockIdx.y * Ke (doubles); a failing element is reported at index
// base + blockIdx.y * elem + e, base being the kind's first element in triplet
// order (CQUAD4s, CTRIA3s, CTETRA10s).  Zero strides for a single model.
struct ElemVar {
    int64_t prop = 0, Ke = 0, elem = 0, base = 0;
};

__device__ inline void report(int64_t *status, int64_t elem, int code) {
    // first failing element wins (smallest index); code stored by that element
    unsigned long long old = atomicMin((unsigned long long *)&status[0], (unsigned long long)elem);
    if ((long long)old > elem) atomicExch((unsigned long long *)&status[1], (unsigned long long)code);
}

__device__ inline void shape_derivs(double r, double s, double *dNr, double *dNs) {
    dNr[0] = -0.25 * (1 - s); dNr[1] = 0.25 * (1 - s); dNr[2] = 0.25 * (1 + s); dNr[3] = -0.25 * (1 + s);
    dNs[0] = -0.25 * (1 - r); dNs[1] = -0.25 * (1 + r); dNs[2] = 0.25 * (1 + r); dNs[3] = 0.25 * (1 - r);
}

__device__ inline void jac4(const double *dNr, const double *dNs, const double *lc, double &J11, double &J12,
                            double &J21, double &J22) {
    J11 = dNr[0] * lc[0] + dNr[1] * lc[2] + dNr[2] * lc[4] + dNr[3] * lc[6];
    J12 = dNr[0] * lc[1] + dNr[1] * lc[3] + dNr[2] * lc[5] + dNr[3] * lc[7];
    J21 = dNs[0] * lc[0] + dNs[1] * lc[2] + dNs[2] * lc[4] + dNs[3] * lc[6];
    J22 = dNs[0] * lc[1] + dNs[1] * lc[3] + dNs[2] * lc[5] + dNs[3] * lc[7];
}

// _quad4_jacobian_measure, src/FEMKernels.jl:2862-2867
__device__ inline bool jac_ok(double detj, double j11, double j12, double j21, double j22) {
    double scale = fabs(j11 * j22) + fabs(j12 * j21);
    return isfinite(detj) && isfinite(scale) && fabs(detj) > 64 * kEps * scale;
}

// ---- small dense helpers (single lane) ----
__device__ bool lu_solve_dev(const double *Ain, int n, double *B, int nrhs) {
    double A[16];
    int piv[4];
    for (int q = 0; q < n * n; ++q) A[q] = Ain[q];
    for (int k = 0; k < n; ++k) {
        int p = k;
        double mx = fabs(A[k + n * k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i + n * k]) > mx) { mx = fabs(A[i + n * k]); p = i; }
        piv[k] = p;
        if (mx == 0.0) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) { double t = A[k + n * j]; A[k + n * j] = A[p + n * j]; A[p + n * j] = t; }
        for (int i = k + 1; i < n; ++i) {
            A[i + n * k] /= A[k + n * k];
            for (int j = k + 1; j < n; ++j) A[i + n * j] -= A[i + n * k] * A[k + n * j];
        }
    }
    for (int c = 0; c < nrhs; ++c) {
        double *b = B + n * c;
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) b[i] -= A[i + n * j] * b[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) b[i] -= A[i + n * j] * b[j];
            b[i] /= A[i + n * i];
        }
    }
    return true;
}

__device__ bool inv4_dev(const double *a, double *out) {
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
    if (det == 0.0 || !isfinite(det)) return false;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] / det;
    return true;
}

// inverse of K[active, active] scattered into a 4x4 (binary internal-mode weights,
// q4_internal_mode_compliance, src/FEMKernels.jl:1918-1955)
__device__ bool active_inverse_dev(const double *K, int mask, double *C) {
    int act[4], na = 0;
    for (int i = 0; i < 4; ++i)
        if (mask & (1 << i)) act[na++] = i;
    for (int q = 0; q < 16; ++q) C[q] = 0.0;
    if (na == 0) return true;
    if (na == 4) return inv4_dev(K, C);
    double sub[16], inv[16];
    for (int j = 0; j < na; ++j)
        for (int i = 0; i < na; ++i) {
            sub[i + na * j] = K[act[i] + 4 * act[j]];
            inv[i + na * j] = (i == j) ? 1.0 : 0.0;
        }
    if (!lu_solve_dev(sub, na, inv, na)) return false;
    for (int j = 0; j < na; ++j)
        for (int i = 0; i < na; ++i) C[act[i] + 4 * act[j]] = inv[i + na * j];
    return true;
}

__device__ bool inv3_dev(const double *a, double *o) {
    double c00 = a[4] * a[8] - a[7] * a[5];
    double c01 = a[7] * a[2] - a[1] * a[8];
    double c02 = a[1] * a[5] - a[4] * a[2];
    double det = a[0] * c00 + a[3] * c01 + a[6] * c02;
    if (det == 0.0 || !isfinite(det)) return false;
    o[0] = c00 / det; o[3] = (a[6] * a[5] - a[3] * a[8]) / det; o[6] = (a[3] * a[7] - a[6] * a[4]) / det;
    o[1] = c01 / det; o[4] = (a[0] * a[8] - a[6] * a[2]) / det; o[7] = (a[6] * a[1] - a[0] * a[7]) / det;
    o[2] = c02 / det; o[5] = (a[3] * a[2] - a[0] * a[5]) / det; o[8] = (a[0] * a[4] - a[3] * a[1]) / det;
    return true;
}

__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline void normalize3(double *a) {
    double n = sqrt(dot3(a, a));
    a[0] /= n; a[1] /= n; a[2] /= n;
}

// isotropic PSHELL section (src/solver/assembly.jl:3380-3404, 2085-2129, 2174-2233)
__device__ void shell_section(const double *prop, double *Cm, double *Cb, double *Cs, double &G, bool &mo) {
    double h = prop[JFEM_SP_H], E = prop[JFEM_SP_E], nu = prop[JFEM_SP_NU];
    double br = prop[JFEM_SP_BEND_RATIO], tst = prop[JFEM_SP_TS_T];
    G = prop[JFEM_SP_G];
    if (!(G > 0.0)) G = E > 0.0 ? E / (2 * (1 + nu)) : 0.0;
    double cm = (E * h) / (1 - nu * nu);
    const double Mi[9] = {1.0, nu, 0.0, nu, 1.0, 0.0, 0.0, 0.0, (1 - nu) / 2};
    for (int q = 0; q < 9; ++q) Cm[q] = cm * Mi[q];
    if (br > 1e-12) {
        double pc = E / fmax(1.0 - nu * nu, 1e-30);
        double f = ((br * (h * h * h)) / 12.0);
        for (int q = 0; q < 9; ++q) Cb[q] = (pc * Mi[q]) * f;
        double ts = tst * h;
        Cs[0] = ts * G; Cs[1] = 0.0; Cs[2] = 0.0; Cs[3] = ts * G;
        mo = false;
    } else {
        for (int q = 0; q < 9; ++q) Cb[q] = 0.0;
        for (int q = 0; q < 4; ++q) Cs[q] = 0.0;
        mo = true;
    }
}

// _rotate_constitutive_3x3!, src/solver/assembly.jl:2440-2462
__device__ void rot3(double *C, double T11, double T12, double T13, double T21, double T22, double T23,
                     double T31, double T32, double T33) {
    double c11 = C[0], c21 = C[1], c31 = C[2], c12 = C[3], c22 = C[4], c32 = C[5], c13 = C[6], c23 = C[7], c33 = C[8];
    double t11 = c11 * T11 + c12 * T21 + c13 * T31, t12 = c11 * T12 + c12 * T22 + c13 * T32, t13 = c11 * T13 + c12 * T23 + c13 * T33;
    double t21 = c21 * T11 + c22 * T21 + c23 * T31, t22 = c21 * T12 + c22 * T22 + c23 * T32, t23 = c21 * T13 + c22 * T23 + c23 * T33;
    double t31 = c31 * T11 + c32 * T21 + c33 * T31, t32 = c31 * T12 + c32 * T22 + c33 * T32, t33 = c31 * T13 + c32 * T23 + c33 * T33;
    C[0] = T11 * t11 + T21 * t21 + T31 * t31; C[3] = T11 * t12 + T21 * t22 + T31 * t32; C[6] = T11 * t13 + T21 * t23 + T31 * t33;
    C[1] = T12 * t11 + T22 * t21 + T32 * t31; C[4] = T12 * t12 + T22 * t22 + T32 * t32; C[7] = T12 * t13 + T22 * t23 + T32 * t33;
    C[2] = T13 * t11 + T23 * t21 + T33 * t31; C[5] = T13 * t12 + T23 * t22 + T33 * t32; C[8] = T13 * t13 + T23 * t23 + T33 * t33;
}

// ------------------------------------------------------------------------------
// CQUAD4
// ------------------------------------------------------------------------------
struct Q4Shared {
    double lc[8], Cm[9], Cb[9], Cs[4], R[9];
    double h, k6, ds;
    int rigid, modes, do_shear, status;
    double zl[4];                      // corner heights of a warped cell (finite-warp map)
    int warped;
    // the element's working arrays are dead once its blocks are formed, when the
    // finite-warp map and K*Tg (warped cells only) take their place: one wavefront
    // per element runs these phases in program order
    union {
        struct {
            double gdx[4][4], gdy[4][4], gw[4];
            double Bi[4][12], Bib[4][12];      // per Gauss point, 3x4 column-major
            double Kab[96], Kbb[16], Kabs[96], Kbbs[16], Kabb[96], Kbbb[16];
            double Cim[16], Cib[16];
            double Pm[96], Pb[96];
            double Dm[48], Xm[48];             // MacNeal D (4x12) and Z \ D
            double Kp[144];
            double Bd[4][24], dcoef[4];
        };
        struct {
            double Tg[576], W[576];
        };
    };
    double out[576];
};

// Per-element quantities of the serial set-up (frame, section, MacNeal tying
// operators), computed one element per thread by the prep kernels and handed to
// the wavefront kernel through the first kQ4Prep doubles of the element's own
// output slot (overwritten by the element matrix at the end).
struct Q4Prep {
    double lc[8], Cm[9], Cb[9], Cs[4], R[9], zl[4];
    double h, k6, ds;
    int rigid, modes, do_shear, warped;
    double Dm[48], Xm[48];
};
// record layout: lc 0, Cm 8, Cb 17, Cs 26, R 30, zl 39, h 43, k6 44, ds 45,
// rigid 46, modes 47, do_shear 48, warped 49, status 50, Dm 51, Xm 99
constexpr int kQ4Prep = 147;

// B-matrix entries as functions of the local DOF d = 6k + c
__device__ inline double bm_entry(const Q4Shared &S, int g, int l, int d) {
    int k = d / 6, c = d % 6;
    double dx = S.gdx[g][k], dy = S.gdy[g][k];
    if (c == 0) return l == 0 ? dx : (l == 1 ? 0.0 : dy);
    if (c == 1) return l == 0 ? 0.0 : (l == 1 ? dy : dx);
    return 0.0;
}
__device__ inline double bb_entry(const Q4Shared &S, int g, int l, int d) {
    int k = d / 6, c = d % 6;
    double dx = S.gdx[g][k], dy = S.gdy[g][k];
    if (c == 3) return l == 0 ? 0.0 : (l == 1 ? -dy : -dx);
    if (c == 4) return l == 0 ? dx : (l == 1 ? 0.0 : dy);
    return 0.0;
}

// MacNeal transverse-shear block (one lane): add_quad4_macneal_shear_rbf! default slice,
// src/FEMKernels.jl:4495-4834, 6188-6200; quad4_natural_shear_flexibility! :4223-4352
template <class St>
__device__ bool q4_macneal(St &S) {
    const double *lc = S.lc;
    const double pt = 1.0 / sqrt(3.0);
    const double sxi[4] = {0.0, 0.0, -pt, pt}, seta[4] = {-pt, pt, 0.0, 0.0};
    const int comp[4] = {1, 1, 2, 2};
    double Ce[48], Je[4][4];
    for (int sp = 0; sp < 4; ++sp) {
        double xi = sxi[sp], eta = seta[sp];
        double xr = xi == 0.0 ? 0.0 : copysign(1.0, xi);
        double er = eta == 0.0 ? 0.0 : copysign(1.0, eta);
        double dNre[4], dNse[4], J11e, J12e, J21e, J22e;
        shape_derivs(xr, er, dNre, dNse);
        jac4(dNre, dNse, lc, J11e, J12e, J21e, J22e);
        double Ne[4] = {0.25 * (1 - xr) * (1 - er), 0.25 * (1 + xr) * (1 - er), 0.25 * (1 + xr) * (1 + er),
                        0.25 * (1 - xr) * (1 + er)};
        Je[sp][0] = J11e; Je[sp][1] = J12e; Je[sp][2] = J21e; Je[sp][3] = J22e;
        double txe = comp[sp] == 1 ? J11e : J21e, tye = comp[sp] == 1 ? J12e : J22e;
        for (int k = 0; k < 4; ++k) {
            Ce[sp + 4 * (3 * k + 0)] = comp[sp] == 1 ? dNre[k] : dNse[k];
            Ce[sp + 4 * (3 * k + 1)] = -Ne[k] * tye;
            Ce[sp + 4 * (3 * k + 2)] = Ne[k] * txe;
        }
    }
    double *D = S.Dm;
    for (int sp = 0; sp < 4; ++sp) {
        double xi = sxi[sp], eta = seta[sp];
        double J11 = Je[sp][0], J12 = Je[sp][1], J21 = Je[sp][2], J22 = Je[sp][3];
        double detJ = J11 * J22 - J12 * J21;
        double adet = fabs(detJ) < 1e-14 ? (detJ < 0 ? -1e-14 : 1e-14) : detJ;
        double i11 = J22 / adet, i12 = -J12 / adet, i21 = -J21 / adet, i22 = J11 / adet;
        double w1 = 0.5 * (1.0 - eta / pt), w2 = 0.5 * (1.0 + eta / pt);
        double w3 = 0.5 * (1.0 - xi / pt), w4 = 0.5 * (1.0 + xi / pt);
        for (int j = 0; j < 12; ++j) {
            double gxie = w1 * Ce[0 + 4 * j] + w2 * Ce[1 + 4 * j];
            double getae = w3 * Ce[2 + 4 * j] + w4 * Ce[3 + 4 * j];
            D[sp + 4 * j] = sp < 2 ? (i11 * gxie + i12 * getae) : (i21 * gxie + i22 * getae);
        }
    }
    const double rn[4] = {-1.0, 1.0, 1.0, -1.0}, sn[4] = {-1.0, -1.0, 1.0, 1.0};
    double tr[16];
    for (int q = 0; q < 16; ++q) tr[q] = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) {
            int base = 3 * k;
            tr[i + 0] += D[i + 4 * (base + 2)];
            tr[i + 4] += sn[k] * D[i + 4 * (base + 2)];
            tr[i + 8] -= D[i + 4 * (base + 1)];
            tr[i + 12] -= rn[k] * D[i + 4 * (base + 1)];
        }
    double x1 = lc[0], x2 = lc[2], x3 = lc[4], x4 = lc[6];
    double y1 = lc[1], y2 = lc[3], y3 = lc[5], y4 = lc[7];
    double ax = 0.25 * (-x1 + x2 + x3 - x4), ay = 0.25 * (-y1 + y2 + y3 - y4);
    double bx = 0.25 * (-x1 - x2 + x3 + x4), by = 0.25 * (-y1 - y2 + y3 + y4);
    double cx = 0.25 * (x1 - x2 + x3 - x4), cy = 0.25 * (y1 - y2 + y3 - y4);
    double det0 = ax * by - ay * bx;
    double det_scale = hypot(ax, ay) * hypot(bx, by);
    if (det_scale < kEps) det_scale = kEps;
    if (!(fabs(det0) > 64 * kEps * det_scale)) return false;
    double orient = det0 > 0 ? 1.0 : (det0 < 0 ? -1.0 : 0.0);
    const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
    for (int q = 0; q < 4; ++q) {
        double r = cr[q], s = cs_[q];
        double dj = (ax + cx * s) * (by + cy * r) - (ay + cy * s) * (bx + cx * r);
        if (!(orient * dj > 64 * kEps * det_scale)) return false;
    }
    double area = 4 * fabs(det0);
    double dx2 = (2 * ax) * (2 * ax), dy2 = (2 * by) * (2 * by);
    double dxr = S.Cb[0], dyr = S.Cb[4];
    double fx = fabs(2 * ax) > 1e-30 ? (cx / (2 * ax)) * (cx / (2 * ax)) : 0.0;
    double fy = fabs(2 * by) > 1e-30 ? (cy / (2 * by)) * (cy / (2 * by)) : 0.0;
    if (fx + fy > 1e-24 && ((dxr == 0.0) != (dyr == 0.0))) return false;
    double ux = dx2 / ((12 * area) * fmax(fabs(dxr), 1e-30));
    double uy = dy2 / ((12 * area) * fmax(fabs(dyr), 1e-30));
    double denx = dy2 + 39 * dx2, deny = dx2 + 39 * dy2;
    if (!(denx > 0 && deny > 0)) return false;
    double zb[4] = {ux, ((3 * dy2) / denx) * ux, uy, ((3 * dx2) / deny) * uy};
    double Z[16];
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            double v = 0.0;
            for (int k = 0; k < 4; ++k) v += (tr[i + 4 * k] * zb[k]) * tr[j + 4 * k];
            Z[i + 4 * j] = v;
        }
    if (!S.rigid) {
        double gp = 1.0 / sqrt(3.0);
        auto jac = [&](double r, double s) {
            return fabs((ax + cx * s) * (by + cy * r) - (ay + cy * s) * (bx + cx * r));
        };
        double va = sqrt(2 * jac(0.0, -gp)), vb = sqrt(2 * jac(0.0, gp));
        double vc = sqrt(2 * jac(-gp, 0.0)), vd = sqrt(2 * jac(gp, 0.0));
        double pxa[4] = {va, -gp * va, 0, 0}, pxb[4] = {vb, gp * vb, 0, 0};
        double pyc[4] = {0, 0, vc, -gp * vc}, pyd[4] = {0, 0, vd, gp * vd};
        double H[16], Hi[16];
        const double *Cs = S.Cs;
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                H[i + 4 * j] = Cs[0] * (pxa[i] * pxa[j] + pxb[i] * pxb[j]) + Cs[3] * (pyc[i] * pyc[j] + pyd[i] * pyd[j]) +
                               Cs[2] * (pxa[i] + pxb[i]) * (pyc[j] + pyd[j]) / 2 +
                               Cs[1] * (pyc[i] + pyd[i]) * (pxa[j] + pxb[j]) / 2;
                Hi[i + 4 * j] = (i == j) ? 1.0 : 0.0;
            }
        if (!lu_solve_dev(H, 4, Hi, 4)) return false;
        double t4[16];
        for (int q = 0; q < 16; ++q) t4[q] = 0.0;
        for (int j = 0; j < 4; ++j)
            for (int l = 0; l < 4; ++l) {
                double v = Hi[l + 4 * j];
                for (int i = 0; i < 4; ++i) t4[i + 4 * j] += tr[i + 4 * l] * v;
            }
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                double s = 0;
                for (int k = 0; k < 4; ++k) s += t4[i + 4 * k] * tr[j + 4 * k];
                Z[i + 4 * j] = Z[i + 4 * j] + s;  // Z_total = Zb + Zs
            }
    }
    for (int q = 0; q < 48; ++q) S.Xm[q] = D[q];
    return lu_solve_dev(Z, 4, S.Xm, 12);
}

__device__ inline int plate_index(int d) {  // local DOF -> MacNeal plate index (w, rx, ry) or -1
    int k = d / 6, c = d % 6;
    return (c >= 2 && c <= 4) ? 3 * k + (c - 2) : -1;
}

// Per-wavefront CQUAD4 kernel body.  p: 4 node xyz (p[3k+c]); prop: section row.
// global_frame: compute frame + section from p/prop; otherwise use lc/C* given.
__device__ void q4_element(Q4Shared &S, int lane, bool valid, int64_t elem, int64_t *status, double *Ke_out,
                           int64_t report_index) {
    // Phase 1: Gauss-point rows (lanes 0..3), drilling rows (lanes 59..62); the
    // MacNeal operators Dm, Xm come with the prep record
    const double pt = 1.0 / sqrt(3.0);
    const double gr[4] = {-pt, pt, pt, -pt}, gs[4] = {-pt, -pt, pt, pt};
    const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac4(dNr_c, dNs_c, S.lc, J11c, J12c, J21c, J22c);
    double detJc = J11c * J22c - J12c * J21c, abs_detJc = fabs(detJc);
    double inv_detc = 1.0 / detJc;
    double iJ11c = J22c * inv_detc, iJ12c = -J12c * inv_detc, iJ21c = -J21c * inv_detc, iJ22c = J11c * inv_detc;
    if (valid && lane < 4) {
        int g = lane;
        double r = gr[g], s = gs[g];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jac4(dNr, dNs, S.lc, J11, J12, J21, J22);
        double detJ = J11 * J22 - J12 * J21;
        if (!jac_ok(detJ, J11, J12, J21, J22)) S.status = JFEM_ERR_DEGENERATE;
        double inv_det = 1.0 / detJ;
        double iJ11 = J22 * inv_det, iJ12 = -J12 * inv_det, iJ21 = -J21 * inv_det, iJ22 = J11 * inv_det;
        for (int k = 0; k < 4; ++k) {
            S.gdx[g][k] = iJ11 * dNr[k] + iJ12 * dNs[k];
            S.gdy[g][k] = iJ21 * dNr[k] + iJ22 * dNs[k];
        }
        S.gw[g] = fabs(detJ);
        // membrane internal modes, patch-consistent Jacobian (src/FEMKernels.jl:1847-1896)
        double ratio = (iJ11 * iJ22 - iJ12 * iJ21) / (iJ11c * iJ22c - iJ12c * iJ21c);
        double m11 = ratio * iJ11c, m12 = ratio * iJ12c, m21 = ratio * iJ21c, m22 = ratio * iJ22c;
        double p1x = m11 * (-2.0 * r), p1y = m21 * (-2.0 * r), p2x = m12 * (-2.0 * s), p2y = m22 * (-2.0 * s);
        double *Bi = S.Bi[g];
        for (int q = 0; q < 12; ++q) Bi[q] = 0.0;
        Bi[0 + 3 * 0] = p1x; Bi[2 + 3 * 0] = p1y; Bi[1 + 3 * 1] = p1y; Bi[2 + 3 * 1] = p1x;
        Bi[0 + 3 * 2] = p2x; Bi[2 + 3 * 2] = p2y; Bi[1 + 3 * 3] = p2y; Bi[2 + 3 * 3] = p2x;
        // bending internal modes, qm6 Jacobian (src/FEMKernels.jl:3478-3488)
        double gq = detJc / detJ;
        double b1x = (iJ11c * (-2.0 * r)) * gq, b1y = (iJ21c * (-2.0 * r)) * gq;
        double b2x = (iJ12c * (-2.0 * s)) * gq, b2y = (iJ22c * (-2.0 * s)) * gq;
        double *Bb = S.Bib[g];
        for (int q = 0; q < 12; ++q) Bb[q] = 0.0;
        Bb[1 + 3 * 0] = -b1y; Bb[1 + 3 * 1] = -b2y;
        Bb[0 + 3 * 2] = b1x;  Bb[0 + 3 * 3] = b2x;
        Bb[2 + 3 * 0] = -b1x; Bb[2 + 3 * 1] = -b2x;
        Bb[2 + 3 * 2] = b1y;  Bb[2 + 3 * 3] = b2y;
    }
    if (valid && lane >= 59 && lane <= 62) {  // lumped drilling rows, src/FEMKernels.jl:3772-3817
        int i = lane - 59;
        const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(cr[i], cs_[i], dNr, dNs);
        jac4(dNr, dNs, S.lc, J11, J12, J21, J22);
        double detJ = J11 * J22 - J12 * J21;
        if (!jac_ok(detJ, J11, J12, J21, J22)) S.status = JFEM_ERR_DEGENERATE;
        double inv_det = 1.0 / detJ;
        double iJ11 = J22 * inv_det, iJ12 = -J12 * inv_det, iJ21 = -J21 * inv_det, iJ22 = J11 * inv_det;
        for (int q = 0; q < 24; ++q) S.Bd[i][q] = 0.0;
        for (int k = 0; k < 4; ++k) {
            double dx = iJ11 * dNr[k] + iJ12 * dNs[k];
            double dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            S.Bd[i][6 * k + 0] = 0.5 * dy;
            S.Bd[i][6 * k + 1] = -0.5 * dx;
        }
        S.Bd[i][6 * i + 5] = 1.0;
        double A_drill = 4.0 * abs_detJc;
        double alpha_L = ((S.ds * (S.k6 * 1e-6)) * S.Cm[8]) * A_drill;
        S.dcoef[i] = alpha_L * (fabs(detJ) / fmax(abs_detJc, 1e-30));
    }
    __syncthreads();
    // Phase 2: internal-mode couplings K_ab, K_bb (membrane), K_ab/K_bb with the
    // shape metric and with Cb (bending), src/FEMKernels.jl:3565-3610
    if (valid) {
        for (int idx = lane; idx < 336; idx += kWave) {
            int which, i, m;
            double *dst;
            if (idx < 96) { which = 0; i = idx % 24; m = idx / 24; dst = &S.Kab[idx]; }
            else if (idx < 112) { which = 1; i = (idx - 96) % 4; m = (idx - 96) / 4; dst = &S.Kbb[idx - 96]; }
            else if (idx < 208) { which = 2; i = (idx - 112) % 24; m = (idx - 112) / 24; dst = &S.Kabs[idx - 112]; }
            else if (idx < 224) { which = 3; i = (idx - 208) % 4; m = (idx - 208) / 4; dst = &S.Kbbs[idx - 208]; }
            else if (idx < 320) { which = 4; i = (idx - 224) % 24; m = (idx - 224) / 24; dst = &S.Kabb[idx - 224]; }
            else { which = 5; i = (idx - 320) % 4; m = (idx - 320) / 4; dst = &S.Kbbb[idx - 320]; }
            const double *C = (which <= 1) ? S.Cm : ((which <= 3) ? S.Cm : S.Cb);
            double acc = 0.0;
            for (int g = 0; g < 4; ++g) {
                const double *Bx = (which <= 1) ? S.Bi[g] : S.Bib[g];
                double t[3];
                for (int l = 0; l < 3; ++l) {
                    double v = 0.0;
                    for (int lp = 0; lp < 3; ++lp) v += C[l + 3 * lp] * Bx[lp + 3 * m];
                    t[l] = v;
                }
                double w = S.gw[g];
                for (int l = 0; l < 3; ++l) {
                    double a;
                    if (which == 0) a = bm_entry(S, g, l, i);
                    else if (which == 2 || which == 4) a = bb_entry(S, g, l, i);
                    else if (which == 1) a = S.Bi[g][l + 3 * i];
                    else a = S.Bib[g][l + 3 * i];
                    acc += a * (w * t[l]);
                }
            }
            *dst = acc;
        }
    }
    __syncthreads();
    // Phase 3: internal-mode compliances (lanes 0, 1), MacNeal plate block (others)
    if (valid) {
        if (lane == 0) {
            double mx = 0.0;
            for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(S.Kbb[q]));
            if (mx > 0) {
                if (!active_inverse_dev(S.Kbb, S.modes, S.Cim)) S.status = JFEM_ERR_DEGENERATE;
            } else {
                for (int q = 0; q < 16; ++q) S.Cim[q] = 0.0;
            }
        } else if (lane == 1) {
            double mx = 0.0;
            for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(S.Kbbs[q]));
            if (mx > 0) {
                if (!inv4_dev(S.Kbbs, S.Cib)) S.status = JFEM_ERR_DEGENERATE;
            } else {
                for (int q = 0; q < 16; ++q) S.Cib[q] = 0.0;
            }
        } else if (S.do_shear) {
            for (int idx = lane - 2; idx < 144; idx += kWave - 2) {
                int i = idx % 12, j = idx / 12;
                double rij = 0.0, rji = 0.0;
                for (int k = 0; k < 4; ++k) rij += S.Dm[k + 4 * i] * S.Xm[k + 4 * j];
                for (int k = 0; k < 4; ++k) rji += S.Dm[k + 4 * j] * S.Xm[k + 4 * i];
                S.Kp[idx] = 0.5 * (rij + rji);
            }
        }
    }
    __syncthreads();
    // Phase 4: enrichment maps Pm = -Cm_inv K_ab', Pb = -Cb_inv K_abs' (src/FEMKernels.jl:1989-2003)
    if (valid) {
        for (int idx = lane; idx < 192; idx += kWave) {
            bool bend = idx >= 96;
            int q = bend ? idx - 96 : idx;
            int i = q % 4, j = q / 4;
            const double *C = bend ? S.Cib : S.Cim;
            const double *Ka = bend ? S.Kabs : S.Kab;
            double v = 0.0;
            for (int k = 0; k < 4; ++k) v -= C[i + 4 * k] * Ka[j + 24 * k];
            (bend ? S.Pb : S.Pm)[q] = v;
        }
    }
    __syncthreads();
    // Phase 5: every lane forms its 3x3 block (A, B) in S.out, one entry at a time
    // (the entry loops stay rolled: unrolled, they pushed the kernel past the
    // register file).  Rows 3A..3A+2 are the translations (A even) or rotations
    // (A odd) of node A/2: membrane B-rows and the membrane enrichment are nonzero
    // only for translations u, v, the bending ones only for rotations rx, ry, so a
    // lane adds the membrane terms when A and B are both even, the bending terms
    // when both are odd, in the order of the full sums (the skipped terms are
    // exact zeros).  The Gauss-point derivatives of its two nodes sit in registers.
    if (valid) {
        const int A = lane / 8, B = lane % 8;
        const int kA = A >> 1, kB = B >> 1;
        const bool mem = !(A & 1) && !(B & 1), bend = (A & 1) && (B & 1);
        double ax[4], ay[4], bx[4], by[4], gw[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            ax[g] = S.gdx[g][kA]; ay[g] = S.gdy[g][kA];
            bx[g] = S.gdx[g][kB]; by[g] = S.gdy[g][kB];
            gw[g] = S.gw[g];
        }
        const double *C = mem ? S.Cm : S.Cb;
        const double *Ka = mem ? S.Kab : S.Kabb, *Pa = mem ? S.Pm : S.Pb, *Kb = mem ? S.Kbb : S.Kbbb;
#pragma unroll 1
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll 1
            for (int ii = 0; ii < 3; ++ii) {
                const int i = 3 * A + ii, j = 3 * B + jj;
                double v = 0.0;
                if (mem || bend) {
                    // B-row entries (l = 0..2) of DOF i / j at Gauss point g: membrane
                    // u: (dx, 0, dy), v: (0, dy, dx); bending rx: (0, -dy, -dx), ry: (dx, 0, dy)
                    for (int g = 0; g < 4; ++g) {
                        double bi[3], bj[3];
                        const int ci = mem ? ii : ii + 3, cj = mem ? jj : jj + 3;
                        const double dxi = ax[g], dyi = ay[g], dxj = bx[g], dyj = by[g];
                        if (ci == 0 || ci == 4) { bi[0] = dxi; bi[1] = 0.0; bi[2] = dyi; }
                        else if (ci == 1) { bi[0] = 0.0; bi[1] = dyi; bi[2] = dxi; }
                        else if (ci == 3) { bi[0] = 0.0; bi[1] = -dyi; bi[2] = -dxi; }
                        else { bi[0] = 0.0; bi[1] = 0.0; bi[2] = 0.0; }
                        if (cj == 0 || cj == 4) { bj[0] = dxj; bj[1] = 0.0; bj[2] = dyj; }
                        else if (cj == 1) { bj[0] = 0.0; bj[1] = dyj; bj[2] = dxj; }
                        else if (cj == 3) { bj[0] = 0.0; bj[1] = -dyj; bj[2] = -dxj; }
                        else { bj[0] = 0.0; bj[1] = 0.0; bj[2] = 0.0; }
                        const double w = gw[g];
                        double t[3];
                        for (int l = 0; l < 3; ++l) {
                            double a = 0.0;
                            for (int lp = 0; lp < 3; ++lp) a += C[l + 3 * lp] * bj[lp];
                            t[l] = a;
                        }
                        for (int l = 0; l < 3; ++l) v += bi[l] * (w * t[l]);
                    }
                }
                // MacNeal plate block, src/FEMKernels.jl:6279-6285
                int pi = plate_index(i), pj = plate_index(j);
                if (S.do_shear && pi >= 0 && pj >= 0) v += S.Kp[pi + 12 * pj];
                // lumped drilling
                for (int c = 0; c < 4; ++c) v += S.Bd[c][i] * (S.dcoef[c] * S.Bd[c][j]);
                // section enrichment, src/FEMKernels.jl:2005-2021
                double e = 0.0;
                if (mem || bend) {
                    for (int k = 0; k < 4; ++k) {
                        const double qi = Ka[i + 24 * k], qj = Ka[j + 24 * k];
                        const double pik = Pa[k + 4 * i];
                        e += qi * Pa[k + 4 * j] + pik * qj;
                        for (int l = 0; l < 4; ++l) e += pik * Kb[k + 4 * l] * Pa[l + 4 * j];
                    }
                }
                S.out[i + 24 * j] = v + e;
            }
    }
    // finite-warp work-dual congruence Ke <- sym(Tg' Ke Tg) for warped cells
    // (apply_quad4_finite_warp_equilibrium!, src/FEMKernels.jl:1082-1116; map
    // _quad4_finite_warp_fill!, :880-945).  Every lane passes the barriers.
    if (valid && S.warped)
        for (int q = lane; q < 576; q += kWave) S.Tg[q] = (q % 25 == 0) ? 1.0 : 0.0;
    __syncthreads();
    if (valid && S.warped && lane < 4) {
        const int i = lane;
        const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(cr[i], cs_[i], dNr, dNs);
        jac4(dNr, dNs, S.lc, J11, J12, J21, J22);
        double dj = J11 * J22 - J12 * J21;
        if (fabs(dj) >= 1e-30) {
            double i11 = J22 / dj, i12 = -J12 / dj, i21 = -J21 / dj, i22 = J11 / dj;
            double zr = dNr[0] * S.zl[0] + dNr[1] * S.zl[1] + dNr[2] * S.zl[2] + dNr[3] * S.zl[3];
            double zs = dNs[0] * S.zl[0] + dNs[1] * S.zl[1] + dNs[2] * S.zl[2] + dNs[3] * S.zl[3];
            double gx = i11 * zr + i12 * zs, gy = i21 * zr + i22 * zs;
            int rxi = 6 * i + 3, ryi = 6 * i + 4;
            S.Tg[rxi + 24 * (6 * i + 5)] += gx;
            S.Tg[ryi + 24 * (6 * i + 5)] += gy;
            int a = (i % 2 == 0) ? 0 : 1, b = (i % 2 == 0) ? 2 : 3;
            double dx = S.lc[2 * b] - S.lc[2 * a], dy = S.lc[2 * b + 1] - S.lc[2 * a + 1];
            double dl2 = dx * dx + dy * dy;
            if (dl2 > 1e-30) {
                int js[2] = {a, b};
                double cu[2] = {dy / dl2, -dy / dl2}, cv[2] = {-dx / dl2, dx / dl2};
                for (int t = 0; t < 2; ++t) {
                    int j = js[t];
                    S.Tg[rxi + 24 * (6 * j + 0)] -= gx * cu[t];
                    S.Tg[rxi + 24 * (6 * j + 1)] -= gx * cv[t];
                    S.Tg[ryi + 24 * (6 * j + 0)] -= gy * cu[t];
                    S.Tg[ryi + 24 * (6 * j + 1)] -= gy * cv[t];
                }
            }
            for (int j = 0; j < 4; ++j) {
                double dNdx = i11 * dNr[j] + i12 * dNs[j];
                double dNdy = i21 * dNr[j] + i22 * dNs[j];
                S.Tg[(6 * i + 0) + 24 * (6 * j + 2)] += S.zl[i] * dNdx;
                S.Tg[(6 * i + 1) + 24 * (6 * j + 2)] += S.zl[i] * dNdy;
            }
        }
    }
    __syncthreads();
    if (valid && S.warped) {   // W = Ke * Tg
        const int A = lane / 8, B = lane % 8;
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const int i = 3 * A + ii, j = 3 * B + jj;
                double v = 0;
                for (int m = 0; m < 24; ++m) v += S.out[i + 24 * m] * S.Tg[m + 24 * j];
                S.W[i + 24 * j] = v;
            }
    }
    __syncthreads();
    if (valid && S.warped) {   // Kt = Tg' * W  (written over the local Ke)
        const int A = lane / 8, B = lane % 8;
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const int i = 3 * A + ii, j = 3 * B + jj;
                double v = 0;
                for (int l = 0; l < 24; ++l) v += S.Tg[l + 24 * i] * S.W[l + 24 * j];
                S.out[i + 24 * j] = v;
            }
    }
    __syncthreads();
    if (valid) {
        const int A = lane / 8, B = lane % 8;
        double blk[9];
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const int i = 3 * A + ii, j = 3 * B + jj;
                blk[ii + 3 * jj] = S.warped ? 0.5 * (S.out[i + 24 * j] + S.out[j + 24 * i]) : S.out[i + 24 * j];
            }
        // R' * blk * R, R rows = (v1, v2, v3) (src/solver/assembly.jl:392-424)
        const double *R = S.R;
        double t[9];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double s = 0;
                for (int l = 0; l < 3; ++l) s += blk[i + 3 * l] * R[l * 3 + j];
                t[i + 3 * j] = s;
            }
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double s = 0;
                for (int l = 0; l < 3; ++l) s += R[l * 3 + i] * t[l + 3 * j];
                S.out[(3 * A + i) + 24 * (3 * B + j)] = s;
            }
    }
    __syncthreads();
    if (valid) {
        if (S.status != 0) {
            if (lane == 0) report(status, report_index, S.status);
        }
        double *dst = Ke_out + elem * 576;
        for (int q = lane; q < 576; q += kWave) dst[q] = S.out[q];
    }
}

// frame + local coordinates + section for a CQUAD4 in global coordinates
// (shell_element_frame_quad4 diag mode, src/solver/helpers.jl:1422-1481;
//  src/solver/assembly.jl:3686-3691, 4392-4426)
template <class St>
__device__ void q4_setup(St &S, const double *p, const double *prop, double k6rot) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6, *p4 = p + 9;
    double gxi[3], geta[3], vr[3], d13[3], d24[3], v1[3], v2[3], v3[3];
    for (int c = 0; c < 3; ++c) {
        gxi[c] = 0.25 * (((-p1[c]) + p2[c]) + p3[c] - p4[c]);
        geta[c] = 0.25 * (((-p1[c]) - p2[c]) + p3[c] + p4[c]);
        d13[c] = p3[c] - p1[c];
        d24[c] = p4[c] - p2[c];
    }
    cross3(gxi, geta, vr);
    if (dot3(vr, vr) <= 1e-24) cross3(d13, d24, vr);
    v3[0] = vr[0]; v3[1] = vr[1]; v3[2] = vr[2];
    normalize3(v3);
    double a13[3] = {d13[0], d13[1], d13[2]}, a24[3] = {d24[0], d24[1], d24[2]};
    normalize3(a13); normalize3(a24);
    double xr[3] = {a13[0] - a24[0], a13[1] - a24[1], a13[2] - a24[2]};
    if (dot3(xr, xr) < 1e-20) { xr[0] = a13[0] + a24[0]; xr[1] = a13[1] + a24[1]; xr[2] = a13[2] + a24[2]; }
    double dd = dot3(xr, v3);
    double xp[3] = {xr[0] - dd * v3[0], xr[1] - dd * v3[1], xr[2] - dd * v3[2]};
    if (dot3(xp, xp) <= 1e-24) {
        double de = dot3(geta, v3);
        double xa[3] = {geta[0] - de * v3[0], geta[1] - de * v3[1], geta[2] - de * v3[2]};
        if (dot3(xa, xa) <= 1e-24) {
            double e[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
            double d2 = dot3(e, v3);
            xa[0] = e[0] - d2 * v3[0]; xa[1] = e[1] - d2 * v3[1]; xa[2] = e[2] - d2 * v3[2];
        }
        xp[0] = xa[0]; xp[1] = xa[1]; xp[2] = xa[2];
    }
    v1[0] = xp[0]; v1[1] = xp[1]; v1[2] = xp[2];
    normalize3(v1);
    cross3(v3, v1, v2);
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = (((p1[k] + p2[k]) + p3[k]) + p4[k]) / 4.0;
    double lc[8];
    for (int k = 0; k < 4; ++k) {
        double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = dot3(d, v1);
        lc[2 * k + 1] = dot3(d, v2);
    }
    // warp gate (quad4_local_z_from_coords3d, src/FEMKernels.jl:853-865, 1100-1102)
    {
        double a[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
        double b[3] = {p4[0] - p2[0], p4[1] - p2[1], p4[2] - p2[2]};
        double n[3];
        cross3(a, b, n);
        double nr = sqrt(dot3(n, n));
        double w3[3] = {0, 0, 1};
        if (nr > 1e-30) { w3[0] = n[0] / nr; w3[1] = n[1] / nr; w3[2] = n[2] / nr; }
        double zmax = 0, xmax = 0;
        for (int k = 0; k < 4; ++k) {
            double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
            zmax = fmax(zmax, fabs(dot3(d, w3)));
            xmax = fmax(xmax, fabs(lc[2 * k]));
        }
        if (xmax < 1e-30) xmax = 1e-30;
        S.warped = zmax > 1e-12 * xmax;
        for (int k = 0; k < 4; ++k) {
            double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
            S.zl[k] = dot3(d, w3);
        }
    }
    double Cm[9], Cb[9], Cs[4], G;
    bool mo;
    shell_section(prop, Cm, Cb, Cs, G, mo);
    {
        double e[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
        double de = dot3(e, v3);
        double ep[3] = {e[0] - de * v3[0], e[1] - de * v3[1], e[2] - de * v3[2]};
        double l2 = dot3(ep, ep);
        if (l2 > 1e-24) {
            double l = sqrt(l2);
            double xm[3] = {ep[0] / l, ep[1] / l, ep[2] / l};
            double beta = 0.0 + atan2(dot3(xm, v2), dot3(xm, v1));
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
    for (int q = 0; q < 8; ++q) S.lc[q] = lc[q];
    for (int q = 0; q < 9; ++q) { S.Cm[q] = Cm[q]; S.Cb[q] = Cb[q]; }
    for (int q = 0; q < 4; ++q) S.Cs[q] = Cs[q];
    S.R[0] = v1[0]; S.R[1] = v1[1]; S.R[2] = v1[2];
    S.R[3] = v2[0]; S.R[4] = v2[1]; S.R[5] = v2[2];
    S.R[6] = v3[0]; S.R[7] = v3[1]; S.R[8] = v3[2];
    S.h = prop[JFEM_SP_H];
    S.rigid = (prop[JFEM_SP_RIGID_SHEAR] != 0.0) && !mo;
    S.k6 = mo ? 0.0 : k6rot;
    S.ds = mo ? 0.0 : 1.0;
    int modes = (int)prop[JFEM_SP_MEMBRANE_MODES];
    S.modes = (modes <= 0 || modes > 15) ? 15 : modes;
    double mcs = fmax(fmax(fabs(Cs[0]), fabs(Cs[1])), fmax(fabs(Cs[2]), fabs(Cs[3])));
    S.do_shear = !(S.h < 1e-30 || (!S.rigid && mcs < 1e-30));
}

__device__ inline void q4_prep_store(const Q4Prep &P, int st, double *d) {
    for (int q = 0; q < 8; ++q) d[q] = P.lc[q];
    for (int q = 0; q < 9; ++q) { d[8 + q] = P.Cm[q]; d[17 + q] = P.Cb[q]; }
    for (int q = 0; q < 4; ++q) d[26 + q] = P.Cs[q];
    for (int q = 0; q < 9; ++q) d[30 + q] = P.R[q];
    for (int q = 0; q < 4; ++q) d[39 + q] = P.zl[q];
    d[43] = P.h; d[44] = P.k6; d[45] = P.ds;
    d[46] = P.rigid; d[47] = P.modes; d[48] = P.do_shear; d[49] = P.warped; d[50] = st;
    for (int q = 0; q < 48; ++q) { d[51 + q] = P.Dm[q]; d[99 + q] = P.Xm[q]; }
}

// the wavefront's lanes copy an element's prep record into its LDS state
__device__ inline void q4_prep_load(Q4Shared &S, const double *d, int lane) {
    for (int q = lane; q < kQ4Prep; q += kWave) {
        const double v = d[q];
        if (q < 8) S.lc[q] = v;
        else if (q < 17) S.Cm[q - 8] = v;
        else if (q < 26) S.Cb[q - 17] = v;
        else if (q < 30) S.Cs[q - 26] = v;
        else if (q < 39) S.R[q - 30] = v;
        else if (q < 43) S.zl[q - 39] = v;
        else if (q == 43) S.h = v;
        else if (q == 44) S.k6 = v;
        else if (q == 45) S.ds = v;
        else if (q == 46) S.rigid = (int)v;
        else if (q == 47) S.modes = (int)v;
        else if (q == 48) S.do_shear = (int)v;
        else if (q == 49) S.warped = (int)v;
        else if (q == 50) S.status = (int)v;
        else if (q < 99) S.Dm[q - 51] = v;
        else S.Xm[q - 99] = v;
    }
}

// MacNeal tying operators of a prepared element (zero when it has no shear block)
__device__ inline int q4_prep_shear(Q4Prep &P) {
    for (int q = 0; q < 48; ++q) { P.Dm[q] = 0.0; P.Xm[q] = 0.0; }
    return (P.do_shear && !q4_macneal(P)) ? JFEM_ERR_DEGENERATE : 0;
}

// serial set-up of CQUAD4s in global coordinates, one element per thread
__global__ void __launch_bounds__(64)
q4_prep_global_kernel(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
                      const double *__restrict__ z, const int32_t *__restrict__ conn,
                      const double *__restrict__ prop, double k6rot, double *__restrict__ Ke, ElemVar ev) {
    const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (e >= n) return;
    prop += blockIdx.y * ev.prop;
    Ke += blockIdx.y * ev.Ke;
    double p[12], pr[8];
    for (int k = 0; k < 4; ++k) {
        int nd = conn[k * n + e];
        p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd];
    }
    for (int c = 0; c < 8; ++c) pr[c] = prop[c * n + e];
    Q4Prep P;
    q4_setup(P, p, pr, k6rot);
    const int st = q4_prep_shear(P);
    q4_prep_store(P, st, Ke + e * 576);
}

// serial set-up of CQUAD4s given in local coordinates with their sections
__global__ void __launch_bounds__(64)
q4_prep_local_kernel(int64_t n, const double *__restrict__ lc, const double *__restrict__ Cm,
                     const double *__restrict__ Cb, const double *__restrict__ Cs,
                     const double *__restrict__ h, const double *__restrict__ k6,
                     const double *__restrict__ ds, const int32_t *__restrict__ flags,
                     double *__restrict__ Ke) {
    const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (e >= n) return;
    Q4Prep P;
    for (int q = 0; q < 8; ++q) P.lc[q] = lc[8 * e + q];
    for (int q = 0; q < 9; ++q) { P.Cm[q] = Cm[9 * e + q]; P.Cb[q] = Cb[9 * e + q]; }
    for (int q = 0; q < 4; ++q) P.Cs[q] = Cs[4 * e + q];
    for (int q = 0; q < 9; ++q) P.R[q] = (q % 4 == 0) ? 1.0 : 0.0;
    for (int q = 0; q < 4; ++q) P.zl[q] = 0.0;
    P.h = h[e]; P.k6 = k6[e]; P.ds = ds[e];
    P.warped = 0;
    const int f = flags[e];
    P.rigid = f & 1;
    const int modes = (f >> 4) & 15;
    P.modes = modes == 0 ? 15 : modes;
    const double mcs = fmax(fmax(fabs(P.Cs[0]), fabs(P.Cs[1])), fmax(fabs(P.Cs[2]), fabs(P.Cs[3])));
    P.do_shear = !(P.h < 1e-30 || (!P.rigid && mcs < 1e-30));
    const int st = q4_prep_shear(P);
    q4_prep_store(P, st, Ke + e * 576);
}

// the lane-parallel part: one wavefront per prepared element, result over the record
__global__ void __launch_bounds__(kWave * kElemsPerBlock, 2)
q4_elem_kernel(int64_t n, double *__restrict__ Ke, int64_t *status, ElemVar ev) {
    __shared__ Q4Shared sh[kElemsPerBlock];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t e = (int64_t)blockIdx.x * kElemsPerBlock + wave;
    const bool valid = e < n;
    Ke += blockIdx.y * ev.Ke;
    Q4Shared &S = sh[wave];
    if (valid) q4_prep_load(S, Ke + e * 576, lane);
    __syncthreads();
    q4_element(S, lane, valid, e, status, Ke, ev.base + blockIdx.y * ev.elem + e);
}

// ------------------------------------------------------------------------------
// CTRIA3 (stiffness_tria3_matrices_generic defaults, src/FEMKernels.jl:9375-9809)
// ------------------------------------------------------------------------------
struct T3Shared {
    double R[9], Dm[9], Db[9];
    double A, bv[3], cv[3], alpha;
    double W[9], D3[27], Kb[81];
    double out[324];
    int status, zero;
};

__device__ void tria3_guard(double *Zb, double La, double Lb, double c, double s, double Dn) {
    double scale = 6 * (La / Lb) * s * Dn;
    if (!(scale > 0 && Zb[8] > 0)) return;
    double z33 = Zb[8];
    double h11 = scale * (Zb[0] - Zb[6] * Zb[6] / z33) / (La * La);
    double h22 = scale * (Zb[4] - Zb[7] * Zb[7] / z33) / (Lb * Lb);
    double h12 = scale * (Zb[3] - Zb[6] * Zb[7] / z33) / (La * Lb);
    double p11 = h11, p12 = (h12 - c * h11) / s, p22 = (h22 - 2 * c * h12 + c * c * h11) / (s * s);
    double a = fmax(p11, 1.0 / 10000), q = fmax(p22, a / 10000);
    double limit = sqrt((85.0 / 100) * a * q);
    double b = p12 < -limit ? -limit : (p12 > limit ? limit : p12);
    if (a == p11 && b == p12 && q == p22) return;
    double n11 = a, n12 = c * a + s * b, n22 = c * c * a + 2 * c * s * b + s * s * q;
    Zb[0] += (n11 - h11) * La * La / scale;
    Zb[3] += (n12 - h12) * La * Lb / scale;
    Zb[1] = Zb[3];
    Zb[4] += (n22 - h22) * Lb * Lb / scale;
}

// serial part of the CTRIA3 kernel (one lane): geometry, MITC3 tying rows,
// tying rigidity with residual bending flexibility (src/FEMKernels.jl:9379-9733)
__device__ void t3_setup(T3Shared &S, const double *lc, const double *Ds, double k6) {
    double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    double A2 = x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2);
    double A = 0.5 * fabs(A2);
    S.zero = A < 1e-12;
    S.A = A;
    if (S.zero) return;
    S.bv[0] = (y2 - y3) / (2 * A); S.bv[1] = (y3 - y1) / (2 * A); S.bv[2] = (y1 - y2) / (2 * A);
    S.cv[0] = (x3 - x2) / (2 * A); S.cv[1] = (x1 - x3) / (2 * A); S.cv[2] = (x2 - x1) / (2 * A);
    // constant-curvature bending Kb = Bb' Db Bb A
    double Bb[27];
    for (int q = 0; q < 27; ++q) Bb[q] = 0.0;
    for (int i = 0; i < 3; ++i) {
        Bb[0 + 3 * (3 * i + 2)] = S.bv[i]; Bb[1 + 3 * (3 * i + 1)] = -S.cv[i];
        Bb[2 + 3 * (3 * i + 1)] = -S.bv[i]; Bb[2 + 3 * (3 * i + 2)] = S.cv[i];
    }
    for (int j = 0; j < 9; ++j)
        for (int i = 0; i < 9; ++i) {
            double s = 0;
            for (int a = 0; a < 3; ++a) {
                double t = 0;
                for (int b = 0; b < 3; ++b) t += Bb[b + 3 * i] * S.Db[b + 3 * a];
                s += t * Bb[a + 3 * j];
            }
            S.Kb[i + 9 * j] = s * A;
        }
    double Jm11 = x2 - x1, Jm12 = y2 - y1, Jm21 = x3 - x1, Jm22 = y3 - y1;
    double dj = Jm11 * Jm22 - Jm12 * Jm21;
    double Ji11 = Jm22 / dj, Ji12 = -Jm12 / dj, Ji21 = -Jm21 / dj, Ji22 = Jm11 / dj;
    double grA[9], gsB[9], grC[9], gsC[9];
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
        for (int q = 0; q < 9; ++q) {
            if (pnt == 0) grA[q] = Br[q];
            else if (pnt == 1) gsB[q] = Bs[q];
            else { grC[q] = Br[q]; gsC[q] = Bs[q]; }
        }
    }
    for (int q = 0; q < 9; ++q) {
        S.D3[0 + 3 * q] = grA[q];
        S.D3[1 + 3 * q] = gsB[q];
        S.D3[2 + 3 * q] = (grC[q] - grA[q]) - (gsC[q] - gsB[q]);
    }
    double La = hypot(x2 - x1, y2 - y1), Lb = hypot(x3 - x1, y3 - y1), Lc = hypot(x3 - x2, y3 - y2);
    double P = La + Lb + Lc;
    double cth = ((x2 - x1) * (x3 - x1) + (y2 - y1) * (y3 - y1)) / fmax(La * Lb, 1e-30);
    cth = cth < -1 ? -1 : (cth > 1 ? 1 : cth);
    double sth = sqrt(fmax(1 - cth * cth, 0.0));
    double t11 = Ds[0] * Ji11 + Ds[2] * Ji21, t12 = Ds[0] * Ji12 + Ds[2] * Ji22;
    double t21 = Ds[1] * Ji11 + Ds[3] * Ji21, t22 = Ds[1] * Ji12 + Ds[3] * Ji22;
    double G11 = Ji11 * t11 + Ji21 * t21, G12 = Ji11 * t12 + Ji21 * t22, G22 = Ji12 * t12 + Ji22 * t22;
    double uu = 3 * La / P, vv = 3 * Lb / P;
    double *W = S.W;
    W[0] = G11 * A; W[3] = G12 * A; W[1] = W[3]; W[4] = G22 * A;
    W[6] = (A / 3) * (uu * G11 - vv * G12); W[2] = W[6];
    W[7] = (A / 3) * (uu * G12 - vv * G22); W[5] = W[7];
    W[8] = uu * uu * G11 * (A / 6) - 2 * uu * vv * G12 * (A * (2 - cth) / 18) + vv * vv * G22 * (A / 6);
    double Dbend = fabs(S.Db[0]);
    if (sth > 1e-9 && Dbend > 1e-30 && La > 1e-30 && Lb > 1e-30) {
        double S1 = La, S2 = Lb, S3 = P, s6 = 1.0 / 6;
        double D11e = Dbend, D22e = fabs(S.Db[4]);
        bool dir_pair = D11e > 1e-30 && D22e > 1e-30;
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
            if (!dir_pair || n < 1e-30) { Dn[e] = Dbend; continue; }
            double nx = -vy / n, ny = vx / n;
            double wv[3] = {nx * nx, ny * ny, 2 * nx * ny};
            double d = 0;
            for (int ii = 0; ii < 3; ++ii)
                for (int jj = 0; jj < 3; ++jj) d += wv[ii] * S.Db[ii + 3 * jj] * wv[jj];
            Dn[e] = fabs(d) > 1e-30 ? d : Dbend;
        }
        double Gd = dir_pair ? sqrt(D11e * D22e) : Dbend;
        double c11 = dir_pair ? s6 / Gd + fr / Dn[0] + w13 / Dn[1] + w23 / Dn[2] : a11 / Dbend;
        double c22 = dir_pair ? s6 / Gd + w12 / Dn[0] + fri / Dn[1] + w23 / Dn[2] : a22 / Dbend;
        double c12 = dir_pair ? -s6 / Gd + o12 / Dn[0] + o13 / Dn[1] + o23 / Dn[2] : a12 / Dbend;
        double Zb[9];
        Zb[0] = c11 * S1 * S1; Zb[3] = c12 * S1 * S2; Zb[1] = Zb[3]; Zb[4] = c22 * S2 * S2;
        Zb[6] = -s6 / Gd * S1 * S3; Zb[2] = Zb[6];
        Zb[7] = s6 / Gd * S2 * S3; Zb[5] = Zb[7];
        Zb[8] = s6 / Gd * S3 * S3;
        tria3_guard(Zb, La, Lb, cth, sth, dir_pair ? Dn[0] : Dbend);
        double Wi[9], Zt[9], Zs[9];
        if (!inv3_dev(W, Wi)) { S.status = JFEM_ERR_DEGENERATE; return; }
        for (int q = 0; q < 9; ++q) Zt[q] = Wi[q] + Zb[q];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) Zs[i + 3 * j] = 0.5 * (Zt[i + 3 * j] + Zt[j + 3 * i]);
        if (!inv3_dev(Zs, W)) { S.status = JFEM_ERR_DEGENERATE; return; }
    }
    S.alpha = (k6 * 1e-6) * S.Dm[8] * (2 * A);
}

__device__ inline int t3_mem(int d) {  // local DOF -> CST membrane index or -1
    int k = d / 6, c = d % 6;
    return c < 2 ? 2 * k + c : -1;
}
__device__ inline int t3_plate(int d) {
    int k = d / 6, c = d % 6;
    return (c >= 2 && c <= 4) ? 3 * k + (c - 2) : -1;
}
__device__ inline double t3_bm(const T3Shared &S, int l, int m) {  // Bm (3x6) entry
    int i = m / 2, c = m % 2;
    if (c == 0) return l == 0 ? S.bv[i] : (l == 1 ? 0.0 : S.cv[i]);
    return l == 0 ? 0.0 : (l == 1 ? S.cv[i] : S.bv[i]);
}

// Local = false: element e of a mesh (frame, section, T'KT), as assembled;
// Local = true: the element-frame kernel on caller-given local coordinates and
// section matrices (lc n x 6, Dm / Db n x 9, Ds n x 4 column-major per element,
// k6 per element), frame = identity.
struct T3LocalIn {
    const double *lc, *Dm, *Db, *Ds, *k6;
};

template <bool Local>
__global__ void __launch_bounds__(kWave * kElemsPerBlock)
t3_kernel(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
          const double *__restrict__ z, const int32_t *__restrict__ conn,
          const double *__restrict__ prop, double k6rot, T3LocalIn loc, double *__restrict__ Ke,
          int64_t *status, ElemVar ev) {
    __shared__ T3Shared sh[kElemsPerBlock];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t e = (int64_t)blockIdx.x * kElemsPerBlock + wave;
    const bool valid = e < n;
    if (!Local) prop += blockIdx.y * ev.prop;
    Ke += blockIdx.y * ev.Ke;
    T3Shared &S = sh[wave];
    if (Local && valid && lane == 0) {
        for (int q = 0; q < 9; ++q) {
            S.Dm[q] = loc.Dm[9 * e + q];
            S.Db[q] = loc.Db[9 * e + q];
            S.R[q] = (q % 4 == 0) ? 1.0 : 0.0;
        }
        S.status = 0;
        t3_setup(S, loc.lc + 6 * e, loc.Ds + 4 * e, loc.k6[e]);
    }
    if (!Local && valid && lane == 0) {
        double p[9], pr[8];
        for (int k = 0; k < 3; ++k) {
            int nd = conn[k * n + e];
            p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd];
        }
        for (int c = 0; c < 8; ++c) pr[c] = prop[c * n + e];
        // shell_element_frame_fast n=3, src/solver/helpers.jl:1486-1490
        double v1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
        normalize3(v1);
        double e13[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]}, v3[3], v2[3];
        cross3(v1, e13, v3);
        normalize3(v3);
        cross3(v3, v1, v2);
        double c[3], lc[6];
        for (int k = 0; k < 3; ++k) c[k] = ((p[k] + p[3 + k]) + p[6 + k]) / 3.0;
        for (int k = 0; k < 3; ++k) {
            double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
            lc[2 * k] = dot3(d, v1);
            lc[2 * k + 1] = dot3(d, v2);
        }
        double Cs[4], G;
        bool mo;
        shell_section(pr, S.Dm, S.Db, Cs, G, mo);
        S.R[0] = v1[0]; S.R[1] = v1[1]; S.R[2] = v1[2];
        S.R[3] = v2[0]; S.R[4] = v2[1]; S.R[5] = v2[2];
        S.R[6] = v3[0]; S.R[7] = v3[1]; S.R[8] = v3[2];
        S.status = 0;
        t3_setup(S, lc, Cs, mo ? 0.0 : k6rot);
    }
    __syncthreads();
    if (valid && lane < 36) {
        const int A = lane / 6, B = lane % 6;
        double blk[9];
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 3; ++ii) {
                const int i = 3 * A + ii, j = 3 * B + jj;
                double v = 0.0;
                if (!S.zero) {
                    int mi = t3_mem(i), mj = t3_mem(j);
                    if (mi >= 0 && mj >= 0) {  // CST membrane, src/FEMKernels.jl:9390-9396
                        double s = 0;
                        for (int a = 0; a < 3; ++a) {
                            double t = 0;
                            for (int b = 0; b < 3; ++b) t += t3_bm(S, b, mi) * S.Dm[b + 3 * a];
                            s += t * t3_bm(S, a, mj);
                        }
                        v = s * S.A;
                    }
                    int pi = t3_plate(i), pj = t3_plate(j);
                    if (pi >= 0 && pj >= 0) {  // Kb + D3' W D3, :9731-9733
                        double s = 0;
                        for (int a = 0; a < 3; ++a) {
                            double t = 0;
                            for (int b = 0; b < 3; ++b) t += S.D3[b + 3 * pi] * S.W[b + 3 * a];
                            s += t * S.D3[a + 3 * pj];
                        }
                        v += S.Kb[pi + 9 * pj] + s;
                    }
                    // lumped drilling, :9780-9795
                    for (int c = 0; c < 3; ++c) {
                        int ki = i / 6, ci = i % 6, kj = j / 6, cj = j % 6;
                        double bi = ci == 0 ? 0.5 * S.cv[ki] : (ci == 1 ? -0.5 * S.bv[ki] : (ci == 5 && ki == c ? 1.0 : 0.0));
                        double bj = cj == 0 ? 0.5 * S.cv[kj] : (cj == 1 ? -0.5 * S.bv[kj] : (cj == 5 && kj == c ? 1.0 : 0.0));
                        v += S.alpha * (bi * bj);
                    }
                }
                blk[ii + 3 * jj] = v;
            }
        const double *R = S.R;
        double t[9];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double s = 0;
                for (int l = 0; l < 3; ++l) s += blk[i + 3 * l] * R[l * 3 + j];
                t[i + 3 * j] = s;
            }
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double s = 0;
                for (int l = 0; l < 3; ++l) s += R[l * 3 + i] * t[l + 3 * j];
                S.out[(3 * A + i) + 18 * (3 * B + j)] = s;
            }
    }
    __syncthreads();
    if (valid) {
        if (lane == 0 && S.status != 0) report(status, ev.base + blockIdx.y * ev.elem + e, S.status);
        double *dst = Ke + e * 324;
        for (int q = lane; q < 324; q += kWave) dst[q] = S.out[q];
    }
}

// ------------------------------------------------------------------------------
// CTETRA10 (stiffness_tetra10, src/tetra10.jl:6-79; iso_3d_constitutive,
// src/FEMKernels.jl:12056-12069)
// ------------------------------------------------------------------------------
struct TetShared {
    double X[30];
    double g[4][3][10];
    double w[4];
    double D[36];
    double out[900];
    int status;
};

__device__ inline double tet_b(const TetShared &S, int q, int l, int d) {  // B (6x30) entry
    int n = d / 3, c = d % 3;
    double dx = S.g[q][0][n], dy = S.g[q][1][n], dz = S.g[q][2][n];
    switch (l) {
        case 0: return c == 0 ? dx : 0.0;
        case 1: return c == 1 ? dy : 0.0;
        case 2: return c == 2 ? dz : 0.0;
        case 3: return c == 0 ? dy : (c == 1 ? dx : 0.0);
        case 4: return c == 1 ? dz : (c == 2 ? dy : 0.0);
        default: return c == 0 ? dz : (c == 2 ? dx : 0.0);
    }
}

// Direct = false: element e of a mesh (connectivity into the grid-point
// coordinates, section record (E, nu) column-major n x 2); Direct = true: the
// caller's coordinates Xd (n x 30: per element the 10 x 3 matrix, column-major)
// and E, nu given as prop[e], prop[n + e].
template <bool Direct>
__global__ void __launch_bounds__(kWave * kElemsPerBlock)
tet10_kernel(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
             const double *__restrict__ z, const int32_t *__restrict__ conn, const double *__restrict__ Xd,
             const double *__restrict__ prop, double *__restrict__ Ke, int64_t *status, ElemVar ev) {
    __shared__ TetShared sh[kElemsPerBlock];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t e = (int64_t)blockIdx.x * kElemsPerBlock + wave;
    const bool valid = e < n;
    prop += blockIdx.y * ev.prop;
    Ke += blockIdx.y * ev.Ke;
    TetShared &S = sh[wave];
    if (valid) {
        if (lane < 10) {
            if (Direct) {
                const double *Xe = Xd + 30 * e;
                S.X[lane] = Xe[lane]; S.X[lane + 10] = Xe[lane + 10]; S.X[lane + 20] = Xe[lane + 20];
            } else {
                int nd = conn[lane * n + e];
                S.X[lane] = x[nd]; S.X[lane + 10] = y[nd]; S.X[lane + 20] = z[nd];
            }
        }
        if (lane == 0) {
            S.status = 0;
            double E = prop[e], nu = prop[n + e];
            if (!(isfinite(E) && E > 0 && isfinite(nu) && nu > -1 && nu < 0.5)) S.status = JFEM_ERR_BAD_ARG;
            double cc = E / ((1.0 + nu) * (1.0 - 2.0 * nu)), dd = 1.0 - nu, ss = (1.0 - 2.0 * nu) / 2.0;
            for (int q = 0; q < 36; ++q) S.D[q] = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) S.D[i + 6 * j] = (i == j) ? cc * dd : cc * nu;
            for (int i = 3; i < 6; ++i) S.D[i + 6 * i] = cc * ss;
        }
    }
    __syncthreads();
    if (valid && lane < 4) {
        const double a = (5.0 + 3.0 * sqrt(5.0)) / 20.0, b = (5.0 - sqrt(5.0)) / 20.0;
        const double pts[4][3] = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
        const int edges[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
        const double dL[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
        int q = lane;
        double L[4] = {1.0 - pts[q][0] - pts[q][1] - pts[q][2], pts[q][0], pts[q][1], pts[q][2]};
        double dN[3][10];
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 3; ++k) dN[k][i] = (4.0 * L[i] - 1.0) * dL[k][i];
        for (int ed = 0; ed < 6; ++ed) {
            int i = edges[ed][0], j = edges[ed][1];
            for (int k = 0; k < 3; ++k) dN[k][ed + 4] = 4.0 * (L[j] * dL[k][i] + L[i] * dL[k][j]);
        }
        double J[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int nn = 0; nn < 10; ++nn) s += dN[r][nn] * S.X[nn + 10 * c];
                J[r][c] = s;
            }
        double c1[3] = {J[0][0], J[1][0], J[2][0]}, c2[3] = {J[0][1], J[1][1], J[2][1]}, c3[3] = {J[0][2], J[1][2], J[2][2]};
        double cr[3];
        cross3(c2, c3, cr);
        double det = dot3(c1, cr);
        if (!(isfinite(det) && det > 0.0)) S.status = JFEM_ERR_INVERTED;
        S.w[q] = det / 24.0;
        double Jc[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Jc[r + 3 * c] = J[r][c];
        double Ji[9];
        Ji[0] = (Jc[4] * Jc[8] - Jc[7] * Jc[5]) / det; Ji[3] = (Jc[6] * Jc[5] - Jc[3] * Jc[8]) / det; Ji[6] = (Jc[3] * Jc[7] - Jc[6] * Jc[4]) / det;
        Ji[1] = (Jc[7] * Jc[2] - Jc[1] * Jc[8]) / det; Ji[4] = (Jc[0] * Jc[8] - Jc[6] * Jc[2]) / det; Ji[7] = (Jc[6] * Jc[1] - Jc[0] * Jc[7]) / det;
        Ji[2] = (Jc[1] * Jc[5] - Jc[4] * Jc[2]) / det; Ji[5] = (Jc[3] * Jc[2] - Jc[0] * Jc[5]) / det; Ji[8] = (Jc[0] * Jc[4] - Jc[3] * Jc[1]) / det;
        for (int r = 0; r < 3; ++r)
            for (int nn = 0; nn < 10; ++nn) S.g[q][r][nn] = Ji[r] * dN[0][nn] + Ji[r + 3] * dN[1][nn] + Ji[r + 6] * dN[2][nn];
    }
    __syncthreads();
    if (valid) {
        for (int blk = lane; blk < 100; blk += kWave) {
            int a = blk / 10, b = blk % 10;
            for (int jj = 0; jj < 3; ++jj)
                for (int ii = 0; ii < 3; ++ii) {
                    int I = 3 * a + ii, J = 3 * b + jj;
                    int r = I <= J ? I : J, c = I <= J ? J : I;  // Matrix(Symmetric(K)): upper mirrored
                    double kv = 0.0;
                    for (int q = 0; q < 4; ++q) {
                        double s = 0;
                        for (int l = 0; l < 6; ++l) {
                            double db = 0;
                            for (int lp = 0; lp < 6; ++lp) db += S.D[l + 6 * lp] * tet_b(S, q, lp, c);
                            s += tet_b(S, q, l, r) * db;
                        }
                        kv += S.w[q] * s;
                    }
                    S.out[I + 30 * J] = kv;
                }
        }
    }
    __syncthreads();
    if (valid) {
        if (lane == 0 && S.status != 0) report(status, ev.base + blockIdx.y * ev.elem + e, S.status);
        double *dst = Ke + e * 900;
        for (int q = lane; q < 900; q += kWave) dst[q] = S.out[q];
    }
}

// ------------------------------------------------------------------------------
// Assembly: one thread per CSR slot sums its contributions in triplet order.
// ------------------------------------------------------------------------------
__global__ void assemble_kernel(int64_t nnz, const int64_t *__restrict__ seg_ptr,
                                const int32_t *__restrict__ seg_src, const double *__restrict__ Ke,
                                double *__restrict__ val, int64_t Ke_stride, int64_t val_stride) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    Ke += blockIdx.y * Ke_stride;
    val += blockIdx.y * val_stride;
    int64_t b = seg_ptr[s], e = seg_ptr[s + 1];
    double v = 0.0;
    for (int64_t p = b; p < e; ++p) v += Ke[(uint32_t)seg_src[p]];
    val[s] = v;
}

__global__ void gather_nodeblk_kernel(int64_t n, const int64_t *__restrict__ pos, const double *__restrict__ val,
                                      double *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t p = pos[i];
    out[i] = p >= 0 ? val[p] : 0.0;
}

// Per grid point: SPC + AUTOSPC of zero / negative diagonal entries (fixed0), and
// a flag set when a translational or rotational group of still-free DOFs could
// hold an eigenvalue below eps_zero * lambda_max.  For a symmetric positive
// definite m x m block lambda_min / lambda_max >= det / trace^m (lambda_max <=
// trace), so a group with det >= thresh * trace^m (thresh >> eps_zero) can make
// no claim and is not flagged; indefinite or NaN groups are flagged.
__global__ void bc_screen_kernel(int64_t nn, const int64_t *__restrict__ pos, const double *__restrict__ val,
                                 const uint8_t *__restrict__ spc, double thresh, uint8_t *__restrict__ fixed0,
                                 uint8_t *__restrict__ flag) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    double kv[36];
#pragma unroll
    for (int q = 0; q < 36; ++q) {
        const int64_t p = pos[36 * n + q];
        kv[q] = p >= 0 ? val[p] : 0.0;
    }
    uint8_t fx[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const double d = kv[6 * a + a];
        fx[a] = spc[6 * n + a] ? 1 : ((d == 0.0 || d < 0) ? 1 : 0);
        fixed0[6 * n + a] = fx[a];
    }
    uint8_t fl = 0;
#pragma unroll
    for (int off = 0; off <= 3; off += 3) {
        int idx[3], m = 0;
        for (int a = 0; a < 3; ++a)
            if (!fx[off + a]) idx[m++] = off + a;
        if (m < 2) continue;
        auto S = [&](int i, int j) { return 0.5 * (kv[6 * idx[i] + idx[j]] + kv[6 * idx[j] + idx[i]]); };
        double det, tr;
        if (m == 2) {
            det = S(0, 0) * S(1, 1) - S(0, 1) * S(0, 1);
            tr = S(0, 0) + S(1, 1);
            if (!(det >= thresh * tr * tr)) fl = 1;
        } else {
            det = S(0, 0) * (S(1, 1) * S(2, 2) - S(1, 2) * S(1, 2)) - S(0, 1) * (S(0, 1) * S(2, 2) - S(1, 2) * S(0, 2)) +
                  S(0, 2) * (S(0, 1) * S(1, 2) - S(1, 1) * S(0, 2));
            tr = S(0, 0) + S(1, 1) + S(2, 2);
            if (!(det >= thresh * tr * tr * tr)) fl = 1;
        }
    }
    flag[n] = fl;
}

// symmetric 3x3 eigen-decomposition by cyclic Jacobi, ascending eigenvalues,
// eigenvectors in the columns of V
__device__ void bc_eig3(const double A[9], double w[3], double V[9]) {
#pragma clang fp contract(off)
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = A[i + 3 * j];
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        const double sc = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
        if (off <= 1e-300 || off <= 1e-18 * sc) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                const double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; ++k) { const double x = a[k][p], y = a[k][q]; a[k][p] = c * x - s * y; a[k][q] = s * x + c * y; }
                for (int k = 0; k < 3; ++k) { const double x = a[p][k], y = a[q][k]; a[p][k] = c * x - s * y; a[q][k] = s * x + c * y; }
                for (int k = 0; k < 3; ++k) { const double x = v[k][p], y = v[k][q]; v[k][p] = c * x - s * y; v[k][q] = s * x + c * y; }
            }
    }
    int idx[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (a[idx[j]][idx[j]] < a[idx[i]][idx[i]]) { const int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int k = 0; k < 3; ++k) {
        w[k] = a[idx[k]][idx[k]];
        for (int r = 0; r < 3; ++r) V[r + 3 * k] = v[r][idx[k]];
    }
}

// Grid-point singularity test of the flagged grid points (src/solver/
// boundary_conditions.jl:1480-1607): per translational and rotational group of
// still-free DOFs, the symmetric part of its block; each eigenvalue below
// eps_zero * lambda_max claims the unclaimed DOF with the largest eigenvector
// component (ties within 1e-12 to the first), which becomes fixed.
__global__ void bc_decide_kernel(int64_t nn, const int64_t *__restrict__ pos, const double *__restrict__ val,
                                 const uint8_t *__restrict__ flag, double eps_zero, uint8_t *__restrict__ fixed) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn || !flag[n]) return;
    auto kv = [&](int a, int b) {
        const int64_t p = pos[36 * n + 6 * a + b];
        return p >= 0 ? val[p] : 0.0;
    };
    for (int off = 0; off <= 3; off += 3) {
        int fl[3], m = 0;
        for (int a = 0; a < 3; ++a)
            if (!fixed[6 * n + off + a]) fl[m++] = a;
        if (m < 1) continue;
        double S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, vals[3] = {0, 0, 0}, vecs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int ii = 0; ii < m; ++ii)
            for (int jj = 0; jj < m; ++jj)
                S[ii + 3 * jj] = (kv(off + fl[ii], off + fl[jj]) + kv(off + fl[jj], off + fl[ii])) / 2;
        if (m == 3) {
            bc_eig3(S, vals, vecs);
        } else if (m == 2) {
            const double a = S[0], b = S[3], d = S[4];
            const double tr = 0.5 * (a + d), df = 0.5 * (a - d), r = sqrt(df * df + b * b);
            vals[0] = tr - r; vals[1] = tr + r;
            for (int k = 0; k < 2; ++k) {
                double vx, vy;
                if (b != 0.0) { vx = b; vy = vals[k] - a; }
                else { vx = (k == 0) == (a <= d) ? 1.0 : 0.0; vy = 1.0 - vx; }
                const double nr = hypot(vx, vy);
                vecs[0 + 3 * k] = vx / nr; vecs[1 + 3 * k] = vy / nr;
            }
        } else {
            vals[0] = S[0]; vecs[0] = 1.0;
        }
        double lmax = 0;
        for (int k = 0; k < m; ++k) lmax = fmax(lmax, fabs(vals[k]));
        if (!(lmax > 0.0)) continue;
        int claimed[3] = {0, 0, 0};
        for (int k = 0; k < m; ++k) {
            if (!(fabs(vals[k]) / lmax < eps_zero)) continue;
            int best_a = -1;
            double best_w = -1.0;
            for (int ii = 0; ii < m; ++ii) {
                const double wv = fabs(vecs[ii + 3 * k]);
                const int a = fl[ii];
                if (claimed[a]) continue;
                if (wv > best_w + 1e-12) { best_w = wv; best_a = a; }
            }
            if (best_a < 0) continue;
            const int64_t g = 6 * n + off + best_a;
            if (fixed[g]) continue;
            claimed[best_a] = 1;
            fixed[g] = 1;
        }
    }
}

__global__ void reset_status_kernel(int64_t *status) {
    status[0] = INT64_MAX;
    status[1] = 0;
}

}  // namespace

// ----------------------------------------------------------------- shell recovery
// Force / stress / strain recovery of isotropic PSHELL shells from a displacement
// vector (recover_shell_stresses!, src/solver/stress_recovery.jl:505-1268), one
// element per thread: a post-processing pass, so the element's operators are
// rebuilt in registers and the 24 x 4 / 12 x 12 couplings are never stored (only
// their products with the element's displacements).
// CQUAD4 output (44 doubles per element): N[3] M[3] Q[2] stress(-h/2)[3]
// stress(+h/2)[3] strain(-h/2)[3] strain(+h/2)[3] corner N[4][3] corner M[4][3];
// CTRIA3 output: the first 20.

// bilinear extrapolation of 2 x 2 Gauss rows G[3g + c] to (xi, eta)
// (interp_2x2_gauss_sigma, src/experimental/plate_kernels.jl:53-70)
__device__ inline void rec_interp(const double *G, double xi, double eta, double *o) {
    const double pt = 1.0 / sqrt(3.0);
    const double lx1 = (pt - xi) / (2.0 * pt), lx2 = (xi + pt) / (2.0 * pt);
    const double ly1 = (pt - eta) / (2.0 * pt), ly2 = (eta + pt) / (2.0 * pt);
    const double w1 = lx1 * ly1, w2 = lx2 * ly1, w3 = lx2 * ly2, w4 = lx1 * ly2;
    for (int c = 0; c < 3; ++c) o[c] = w1 * G[c] + w2 * G[3 + c] + w3 * G[6 + c] + w4 * G[9 + c];
}

// u <- Tg u for a warped cell, Tg the finite-warp displacement map (identity plus
// the entries of _quad4_finite_warp_fill!, src/FEMKernels.jl:880-945), applied
// entry by entry from a copy of the physical displacements
__device__ void rec_warp_map(const double *lc, const double *zl, double *u) {
    double u0[24];
    for (int q = 0; q < 24; ++q) u0[q] = u[q];
    const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; ++i) {
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(cr[i], cs_[i], dNr, dNs);
        jac4(dNr, dNs, lc, J11, J12, J21, J22);
        const double dj = J11 * J22 - J12 * J21;
        if (fabs(dj) < 1e-30) continue;
        const double i11 = J22 / dj, i12 = -J12 / dj, i21 = -J21 / dj, i22 = J11 / dj;
        const double zr = dNr[0] * zl[0] + dNr[1] * zl[1] + dNr[2] * zl[2] + dNr[3] * zl[3];
        const double zs = dNs[0] * zl[0] + dNs[1] * zl[1] + dNs[2] * zl[2] + dNs[3] * zl[3];
        const double gx = i11 * zr + i12 * zs, gy = i21 * zr + i22 * zs;
        double *urx = u + 6 * i + 3, *ury = u + 6 * i + 4;
        *urx += gx * u0[6 * i + 5];
        *ury += gy * u0[6 * i + 5];
        const int a = (i % 2 == 0) ? 0 : 1, b = (i % 2 == 0) ? 2 : 3;
        const double dx = lc[2 * b] - lc[2 * a], dy = lc[2 * b + 1] - lc[2 * a + 1];
        const double dl2 = dx * dx + dy * dy;
        if (dl2 > 1e-30) {
            const int js[2] = {a, b};
            const double cu[2] = {dy / dl2, -dy / dl2}, cv[2] = {-dx / dl2, dx / dl2};
            for (int t = 0; t < 2; ++t) {
                const int j = js[t];
                *urx -= (gx * cu[t]) * u0[6 * j] + (gx * cv[t]) * u0[6 * j + 1];
                *ury -= (gy * cu[t]) * u0[6 * j] + (gy * cv[t]) * u0[6 * j + 1];
            }
        }
        for (int j = 0; j < 4; ++j) {
            const double dNdx = i11 * dNr[j] + i12 * dNs[j], dNdy = i21 * dNr[j] + i22 * dNs[j];
            u[6 * i + 0] += (zl[i] * dNdx) * u0[6 * j + 2];
            u[6 * i + 1] += (zl[i] * dNdy) * u0[6 * j + 2];
        }
    }
}

// 3-vector C x (C column-major 3 x 3)
__device__ inline void rec_mv3(const double *C, const double *x, double *y) {
    for (int l = 0; l < 3; ++l) y[l] = C[l] * x[0] + C[l + 3] * x[1] + C[l + 6] * x[2];
}

// CQUAD4: stress_strain_quad4 default path (src/FEMKernels.jl:6759-6975) with the
// SOL 101/105/106 section corner moments and MacNeal area-mean shear
// (stress_recovery.jl:873-946, quad4_macneal_mean_shear_resultant :6623-6645,
// quad4_condensed_section_fields :7140-7209) when mean_shear, else the MITC4
// centre shear (:6515-6599) blended with the equilibrium shear of the compatible
// corner moments (stress_recovery.jl:126-168, 948-952).  Returns a status code.
__device__ int q4_recover(const double *p, const double *pr, const double *ug, int mean_shear, double *out) {
    Q4Prep P;
    q4_setup(P, p, pr, 0.0);
    const double *lc = P.lc;
    double u[24];
    for (int k = 0; k < 4; ++k)           // element frame (stress_recovery.jl:609-617)
        for (int hf = 0; hf < 2; ++hf)
            for (int i = 0; i < 3; ++i) {
                const double *s = ug + 6 * k + 3 * hf;
                u[6 * k + 3 * hf + i] = P.R[3 * i] * s[0] + P.R[3 * i + 1] * s[1] + P.R[3 * i + 2] * s[2];
            }
    if (P.warped) rec_warp_map(lc, P.zl, u);
    const double h = pr[JFEM_SP_H], E = pr[JFEM_SP_E], nu = pr[JFEM_SP_NU], br = pr[JFEM_SP_BEND_RATIO];
    const double tst = pr[JFEM_SP_TS_T];
    // stress_strain_quad4's plane-stress law D and membrane matrix Cm = D h;
    // section matrices A = h C, D = br h^3/12 C of the mean-shear data
    const double cmem = E / (1 - nu * nu);
    const double Dm3[9] = {cmem * 1.0, cmem * nu, cmem * 0.0, cmem * nu, cmem * 1.0, cmem * 0.0,
                           cmem * 0.0, cmem * 0.0, cmem * ((1 - nu) / 2)};
    const double pc = E / fmax(1.0 - nu * nu, 1e-30);
    const double Csec[9] = {pc * 1.0, pc * nu, pc * 0.0, pc * nu, pc * 1.0, pc * 0.0, pc * 0.0, pc * 0.0,
                            pc * ((1.0 - nu) / 2.0)};
    double Cm[9], Asec[9], Dsec[9], Cbc[9];
    for (int q = 0; q < 9; ++q) {
        Cm[q] = Dm3[q] * h;
        Asec[q] = h * Csec[q];
        Dsec[q] = ((br * (h * h * h)) / 12) * Csec[q];
        Cbc[q] = (br * Dm3[q]) * ((h * h * h) / 12.0);
    }
    const double dNr_c[4] = {-0.25, 0.25, 0.25, -0.25}, dNs_c[4] = {-0.25, -0.25, 0.25, 0.25};
    double J11c, J12c, J21c, J22c;
    jac4(dNr_c, dNs_c, lc, J11c, J12c, J21c, J22c);
    const double detJc = J11c * J22c - J12c * J21c;
    const double iJ11c = J22c / detJc, iJ12c = -J12c / detJc, iJ21c = -J21c / detJc, iJ22c = J11c / detJc;
    const double pt = 1.0 / sqrt(3.0);
    const double gr[4] = {-pt, pt, pt, -pt}, gs[4] = {-pt, -pt, pt, pt};
    // per Gauss point: Bm u, Bb u, the internal-mode B generators, the weight
    double em[4][3], kb[4][3], bi[4][4], bb[4][4], wg[4];
    double Kabu[4] = {0, 0, 0, 0}, Kbb[16], Kabsu[4] = {0, 0, 0, 0}, Kbbs[16];
    for (int q = 0; q < 16; ++q) { Kbb[q] = 0.0; Kbbs[q] = 0.0; }
    for (int g = 0; g < 4; ++g) {
        const double r = gr[g], s = gs[g];
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(r, s, dNr, dNs);
        jac4(dNr, dNs, lc, J11, J12, J21, J22);
        const double detJ = J11 * J22 - J12 * J21;
        if (!jac_ok(detJ, J11, J12, J21, J22)) return JFEM_ERR_DEGENERATE;
        const double iJ11 = J22 / detJ, iJ12 = -J12 / detJ, iJ21 = -J21 / detJ, iJ22 = J11 / detJ;
        double e0 = 0, e1 = 0, e2 = 0, k0 = 0, k1 = 0, k2 = 0;
        for (int k = 0; k < 4; ++k) {
            const double dx = iJ11 * dNr[k] + iJ12 * dNs[k], dy = iJ21 * dNr[k] + iJ22 * dNs[k];
            const double *uk = u + 6 * k;
            e0 += dx * uk[0]; e1 += dy * uk[1]; e2 += dy * uk[0] + dx * uk[1];
            k0 += dx * uk[4]; k1 += -dy * uk[3]; k2 += dy * uk[4] - dx * uk[3];
        }
        em[g][0] = e0; em[g][1] = e1; em[g][2] = e2;
        kb[g][0] = k0; kb[g][1] = k1; kb[g][2] = k2;
        wg[g] = fabs(detJ);
        // membrane modes (patch-consistent Jacobian), bending modes (qm6 Jacobian)
        const double ratio = (iJ11 * iJ22 - iJ12 * iJ21) / (iJ11c * iJ22c - iJ12c * iJ21c);
        bi[g][0] = (ratio * iJ11c) * (-2.0 * r); bi[g][1] = (ratio * iJ21c) * (-2.0 * r);
        bi[g][2] = (ratio * iJ12c) * (-2.0 * s); bi[g][3] = (ratio * iJ22c) * (-2.0 * s);
        const double gq = detJc / detJ;
        bb[g][0] = (iJ11c * (-2.0 * r)) * gq; bb[g][1] = (iJ21c * (-2.0 * r)) * gq;
        bb[g][2] = (iJ12c * (-2.0 * s)) * gq; bb[g][3] = (iJ22c * (-2.0 * s)) * gq;
        // columns m of Bi (membrane) and Bib (bending) as 3-vectors
        double Bm_[4][3] = {{bi[g][0], 0.0, bi[g][1]}, {0.0, bi[g][1], bi[g][0]},
                            {bi[g][2], 0.0, bi[g][3]}, {0.0, bi[g][3], bi[g][2]}};
        double Bb_[4][3] = {{0.0, -bb[g][1], -bb[g][0]}, {0.0, -bb[g][3], -bb[g][2]},
                            {bb[g][0], 0.0, bb[g][1]}, {bb[g][2], 0.0, bb[g][3]}};
        double tm[4][3], tb[4][3];
        for (int m = 0; m < 4; ++m) { rec_mv3(Cm, Bm_[m], tm[m]); rec_mv3(Asec, Bb_[m], tb[m]); }
        const double w = wg[g];
        for (int m = 0; m < 4; ++m) {
            Kabu[m] += (e0 * tm[m][0] + e1 * tm[m][1] + e2 * tm[m][2]) * w;
            Kabsu[m] += (k0 * tb[m][0] + k1 * tb[m][1] + k2 * tb[m][2]) * w;
            for (int i = 0; i < 4; ++i) {
                Kbb[i + 4 * m] += (Bm_[i][0] * tm[m][0] + Bm_[i][1] * tm[m][1] + Bm_[i][2] * tm[m][2]) * w;
                Kbbs[i + 4 * m] += (Bb_[i][0] * tb[m][0] + Bb_[i][1] * tb[m][1] + Bb_[i][2] * tb[m][2]) * w;
            }
        }
    }
    // alpha = -K_bb^{-1} K_ab' u; bending enrichment ab = -K_bbs^{-1} K_abs' u
    double Ci[16], alpha[4] = {0, 0, 0, 0}, ab[4] = {0, 0, 0, 0};
    if (!inv4_dev(Kbb, Ci)) return JFEM_ERR_DEGENERATE;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) alpha[i] -= Ci[i + 4 * k] * Kabu[k];
    double mk = 0.0;
    for (int q = 0; q < 16; ++q) mk = fmax(mk, fabs(Kbbs[q]));
    if (mk > 0) {
        if (!inv4_dev(Kbbs, Ci)) return JFEM_ERR_DEGENERATE;
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) ab[i] -= Ci[i + 4 * k] * Kabsu[k];
    }
    double Ngp[12], Mgp[12], eavg[3] = {0, 0, 0}, kavg[3] = {0, 0, 0}, area = 0.0;
    for (int g = 0; g < 4; ++g) {
        const double eps[3] = {em[g][0] + (bi[g][0] * alpha[0] + bi[g][2] * alpha[2]),
                               em[g][1] + (bi[g][1] * alpha[1] + bi[g][3] * alpha[3]),
                               em[g][2] + (((bi[g][1] * alpha[0] + bi[g][0] * alpha[1]) + bi[g][3] * alpha[2]) +
                                           bi[g][2] * alpha[3])};
        rec_mv3(Cm, eps, Ngp + 3 * g);
        if (mean_shear) {
            const double kap[3] = {kb[g][0] + (bb[g][0] * ab[2] + bb[g][2] * ab[3]),
                                   kb[g][1] + (-bb[g][1] * ab[0] - bb[g][3] * ab[1]),
                                   kb[g][2] + (((-bb[g][0] * ab[0] - bb[g][2] * ab[1]) + bb[g][1] * ab[2]) +
                                               bb[g][3] * ab[3])};
            rec_mv3(Dsec, kap, Mgp + 3 * g);
        } else {
            double t[3];
            rec_mv3(Cbc, kb[g], t);
            for (int l = 0; l < 3; ++l) Mgp[3 * g + l] = -t[l];
        }
        for (int l = 0; l < 3; ++l) { eavg[l] += eps[l] * wg[g]; kavg[l] += kb[g][l] * wg[g]; }
        area += wg[g];
    }
    for (int l = 0; l < 3; ++l) { eavg[l] /= area; kavg[l] /= area; }
    double *N = out, *Mo = out + 3, *Q = out + 6, *sz1 = out + 8, *sz2 = out + 11, *ez1 = out + 14, *ez2 = out + 17;
    double *Nc = out + 20, *Mc = out + 32;
    rec_mv3(Cm, eavg, N);
    {
        double dk[3];
        rec_mv3(Dm3, kavg, dk);
        for (int l = 0; l < 3; ++l) Mo[l] = ((-br) * dk[l]) * ((h * h * h) / 12.0);
    }
    const double z1 = -h / 2.0, z2 = h / 2.0;
    for (int l = 0; l < 3; ++l) { ez1[l] = eavg[l] + z1 * kavg[l]; ez2[l] = eavg[l] + z2 * kavg[l]; }
    rec_mv3(Dm3, ez1, sz1);
    rec_mv3(Dm3, ez2, sz2);
    const double cr[4] = {-1.0, 1.0, 1.0, -1.0}, cs_[4] = {-1.0, -1.0, 1.0, 1.0};
    const bool bending = br > 1e-12;
    for (int k = 0; k < 4; ++k) {
        double t[3];
        rec_interp(Ngp, cr[k], cs_[k], Nc + 3 * k);
        rec_interp(Mgp, cr[k], cs_[k], t);
        for (int l = 0; l < 3; ++l) Mc[3 * k + l] = mean_shear ? (bending ? -t[l] : 0.0) : t[l];
    }
    Q[0] = 0.0; Q[1] = 0.0;
    if (!mean_shear) {
        if (bending) {
            // MITC4 centre shear: covariant shear at the edge tying points, phi2 = min(1, 10 h^2 / (4 |J_c|))
            const double tr[4] = {0.0, 0.0, -1.0, 1.0}, tsy[4] = {-1.0, 1.0, 0.0, 0.0};
            double g1 = 0.0, g2 = 0.0;
            for (int t = 0; t < 4; ++t) {
                double dNr[4], dNs[4], J11, J12, J21, J22;
                shape_derivs(tr[t], tsy[t], dNr, dNs);
                jac4(dNr, dNs, lc, J11, J12, J21, J22);
                const double Nt[4] = {0.25 * (1.0 - tr[t]) * (1.0 - tsy[t]), 0.25 * (1.0 + tr[t]) * (1.0 - tsy[t]),
                                      0.25 * (1.0 + tr[t]) * (1.0 + tsy[t]), 0.25 * (1.0 - tr[t]) * (1.0 + tsy[t])};
                double gam = 0.0;
                for (int k = 0; k < 4; ++k) {
                    const double *uk = u + 6 * k;
                    gam += t < 2 ? (dNr[k] * uk[2] + (-J12 * Nt[k]) * uk[3] + (J11 * Nt[k]) * uk[4])
                                 : (dNs[k] * uk[2] + (-J22 * Nt[k]) * uk[3] + (J21 * Nt[k]) * uk[4]);
                }
                if (t < 2) g1 += 0.5 * gam; else g2 += 0.5 * gam;
            }
            const double Gs = E / (2 * (1 + nu));
            const double phi2 = fmin(1.0, 10.0 * (h * h) / fmax(4.0 * fmax(fabs(detJc), 1e-30), 1e-30));
            const double kk = ((5.0 / 6.0) * Gs) * h;
            const double t11 = kk * iJ11c, t12 = kk * iJ12c, t21 = kk * iJ21c, t22 = kk * iJ22c;
            const double q1 = (phi2 * (iJ11c * t11 + iJ21c * t21)) * g1 + (phi2 * (iJ11c * t12 + iJ21c * t22)) * g2;
            const double q2 = (phi2 * (iJ12c * t11 + iJ22c * t21)) * g1 + (phi2 * (iJ12c * t12 + iJ22c * t22)) * g2;
            Q[0] = J11c * q1 + J21c * q2;
            Q[1] = J12c * q1 + J22c * q2;
            // equilibrium shear (dMx/dx, dMy/dy) of the corner moments, smaller magnitude kept
            if (fabs(detJc) >= 1e-12) {
                const double i11 = J22c / detJc, i12 = -J12c / detJc, i21 = -J21c / detJc, i22 = J11c / detJc;
                double eq[2];
                for (int comp = 0; comp < 2; ++comp) {
                    double dr = 0.0, ds = 0.0;
                    for (int k = 0; k < 4; ++k) { dr += dNr_c[k] * Mc[3 * k + comp]; ds += dNs_c[k] * Mc[3 * k + comp]; }
                    eq[comp] = comp == 0 ? i11 * dr + i12 * ds : i21 * dr + i22 * ds;
                }
                for (int i = 0; i < 2; ++i)
                    if (isfinite(eq[i]) && fabs(eq[i]) < fabs(Q[i])) Q[i] = eq[i];
            }
        }
        for (int k = 0; k < 4; ++k) Mc[3 * k + 2] = Mo[2];
        return 0;
    }
    for (int k = 0; k < 4; ++k) Mc[3 * k + 2] = Mo[2];
    if (bending && tst > 0 && h > 0 && P.do_shear) {
        // MacNeal plate operator of the section (Cb = D, Cs), rows for rx / ry of
        // each corner against the plate DOFs (w, rx, ry)
        if (q4_prep_shear(P) != 0) return JFEM_ERR_DEGENERATE;
        double qx = 0.0, qy = 0.0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int c = 0; c < 3; ++c) {
                    auto kp = [&](int a, int b) {
                        double rab = 0.0, rba = 0.0;
                        for (int k = 0; k < 4; ++k) rab += P.Dm[k + 4 * a] * P.Xm[k + 4 * b];
                        for (int k = 0; k < 4; ++k) rba += P.Dm[k + 4 * b] * P.Xm[k + 4 * a];
                        return 0.5 * (rab + rba);
                    };
                    qx += kp(3 * i + 2, 3 * j + c) * u[6 * j + c + 2];
                    qy -= kp(3 * i + 1, 3 * j + c) * u[6 * j + c + 2];
                }
        const double ar = 4 * fabs(detJc);
        Q[0] = qx / ar; Q[1] = qy / ar;
    }
    return 0;
}

// CTRIA3: stress_strain_tria3 with the MITC3 centroid shear (src/FEMKernels.jl:9859-9937),
// frame shell_element_frame_fast n = 3 (src/solver/helpers.jl:1486-1490)
__device__ int t3_recover(const double *p, const double *pr, const double *ug, double *out) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6;
    double v1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]}, e13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    double v2[3], v3[3];
    normalize3(v1);
    cross3(v1, e13, v3);
    normalize3(v3);
    cross3(v3, v1, v2);
    double c[3], lc[6];
    for (int k = 0; k < 3; ++k) c[k] = ((p1[k] + p2[k]) + p3[k]) / 3.0;
    for (int k = 0; k < 3; ++k) {
        const double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = dot3(d, v1); lc[2 * k + 1] = dot3(d, v2);
    }
    const double *R[3] = {v1, v2, v3};
    double u[18];
    for (int k = 0; k < 3; ++k)
        for (int hf = 0; hf < 2; ++hf)
            for (int i = 0; i < 3; ++i) {
                const double *s = ug + 6 * k + 3 * hf;
                u[6 * k + 3 * hf + i] = R[i][0] * s[0] + R[i][1] * s[1] + R[i][2] * s[2];
            }
    for (int q = 0; q < 20; ++q) out[q] = 0.0;
    const double h = pr[JFEM_SP_H], E = pr[JFEM_SP_E], nu = pr[JFEM_SP_NU], br = pr[JFEM_SP_BEND_RATIO];
    const double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    const double A = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
    if (A < 1e-12) return 0;
    const double b[3] = {(y2 - y3) / (2 * A), (y3 - y1) / (2 * A), (y1 - y2) / (2 * A)};
    const double cc[3] = {(x3 - x2) / (2 * A), (x1 - x3) / (2 * A), (x2 - x1) / (2 * A)};
    const double cm = E / (1 - nu * nu);
    const double D[9] = {cm * 1.0, cm * nu, cm * 0.0, cm * nu, cm * 1.0, cm * 0.0, cm * 0.0, cm * 0.0, cm * ((1 - nu) / 2)};
    double eps[3] = {0, 0, 0}, kap[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        const double *ui = u + 6 * i;
        eps[0] += b[i] * ui[0]; eps[1] += cc[i] * ui[1]; eps[2] += cc[i] * ui[0] + b[i] * ui[1];
        kap[0] += b[i] * ui[4]; kap[1] += -cc[i] * ui[3]; kap[2] += -b[i] * ui[3] + cc[i] * ui[4];
    }
    double de[3], dk[3];
    rec_mv3(D, eps, de);
    rec_mv3(D, kap, dk);
    for (int l = 0; l < 3; ++l) { out[l] = de[l] * h; out[3 + l] = ((-br) * dk[l]) * ((h * h * h) / 12.0); }
    if (br > 1e-12) {
        const double Jm11 = x2 - x1, Jm12 = y2 - y1, Jm21 = x3 - x1, Jm22 = y3 - y1;
        const double detJ = Jm11 * Jm22 - Jm12 * Jm21;
        if (fabs(detJ) >= 1e-30) {
            // covariant shear at A (1/2, 0) along r, B (0, 1/2) along s, and C (1/2, 1/2)
            auto gcov = [&](double r, double s, bool along_r) {
                const double Nn[3] = {1.0 - r - s, r, s};
                const double dwdr[3] = {-1.0, 1.0, 0.0}, dwds[3] = {-1.0, 0.0, 1.0};
                double g = 0.0;
                for (int i = 0; i < 3; ++i) {
                    const double *ui = u + 6 * i;
                    g += along_r ? (dwdr[i] * ui[2] + (-Nn[i] * Jm12) * ui[3] + (Nn[i] * Jm11) * ui[4])
                                 : (dwds[i] * ui[2] + (-Nn[i] * Jm22) * ui[3] + (Nn[i] * Jm21) * ui[4]);
                }
                return g;
            };
            const double a = gcov(0.5, 0.0, true), bb = gcov(0.0, 0.5, false);
            const double cv = (gcov(0.5, 0.5, true) - a) - (gcov(0.5, 0.5, false) - bb);
            const double g1 = a + cv / 3, g2 = bb - cv / 3;
            const double i11 = Jm22 / detJ, i12 = -Jm12 / detJ, i21 = -Jm21 / detJ, i22 = Jm11 / detJ;
            const double f = ((5.0 / 6.0) * (E / (2 * (1 + nu)))) * h;
            out[6] = f * (i11 * g1 + i12 * g2);
            out[7] = f * (i21 * g1 + i22 * g2);
        }
    }
    const double z1 = -h / 2.0, z2 = h / 2.0;
    double e1[3], e2[3];
    for (int l = 0; l < 3; ++l) { e1[l] = eps[l] + z1 * kap[l]; e2[l] = eps[l] + z2 * kap[l]; }
    rec_mv3(D, e1, out + 8);
    rec_mv3(D, e2, out + 11);
    for (int l = 0; l < 3; ++l) { out[14 + l] = e1[l]; out[17 + l] = e2[l]; }
    return 0;
}

// one thread per shell: CQUAD4s [0, nq), then CTRIA3s; the failing shell's index in
// that order is reported
__global__ void __launch_bounds__(64)
shell_recover_kernel(int64_t nq, int64_t nt, const double *__restrict__ x, const double *__restrict__ y,
                     const double *__restrict__ z, const int32_t *__restrict__ qconn, const double *__restrict__ qprop,
                     const int32_t *__restrict__ tconn, const double *__restrict__ tprop, const double *__restrict__ u,
                     int mean_shear, double *__restrict__ qout, double *__restrict__ tout, int64_t *status) {
    const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (e >= nq + nt) return;
    const bool quad = e < nq;
    const int64_t el = quad ? e : e - nq, n = quad ? nq : nt;
    const int nn = quad ? 4 : 3;
    const int32_t *conn = quad ? qconn : tconn;
    const double *prop = quad ? qprop : tprop;
    double p[12], pr[8], ug[24];
    for (int k = 0; k < nn; ++k) {
        const int64_t nd = conn[k * n + el];
        p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd];
        for (int c = 0; c < 6; ++c) ug[6 * k + c] = u[6 * nd + c];
    }
    for (int c = 0; c < 8; ++c) pr[c] = prop[c * n + el];
    const int st = quad ? q4_recover(p, pr, ug, mean_shear, qout + 44 * el) : t3_recover(p, pr, ug, tout + 20 * el);
    if (st != 0) report(status, e, st);
}


// ------------------------------------------------------------------------------
// SOL 105 differential stiffness of a flat isotropic CQUAD4 under the membrane
// state of the preload displacements: geometric_stiffness_quad4_reference_kdjj_iso
// (src/FEMKernels.jl:8810-8942) with the assembly's flatness gate, preload map,
// warp congruence and element-frame transform (src/solver/assembly.jl:6575-6594,
// 8063-8075, 8264-8320).  One thread per element.
// ------------------------------------------------------------------------------

// bilinear gradients at (r, s) in the diagonal-bisector frame (ce, se) and |detJ|;
// false when |detJ| < 1e-14 (grad_p, src/FEMKernels.jl:8857-8876)
__device__ bool kg_grad(const double *X, const double *Y, double ce, double se, double r, double s, double *dNx,
                        double *dNy, double &adet) {
    const double dNr[4] = {-(1.0 - s) * 0.25, (1.0 - s) * 0.25, (1.0 + s) * 0.25, -(1.0 + s) * 0.25};
    const double dNs[4] = {-(1.0 - r) * 0.25, -(1.0 + r) * 0.25, (1.0 + r) * 0.25, (1.0 - r) * 0.25};
    const double J11 = dNr[0] * X[0] + dNr[1] * X[1] + dNr[2] * X[2] + dNr[3] * X[3];
    const double J12 = dNr[0] * Y[0] + dNr[1] * Y[1] + dNr[2] * Y[2] + dNr[3] * Y[3];
    const double J21 = dNs[0] * X[0] + dNs[1] * X[1] + dNs[2] * X[2] + dNs[3] * X[3];
    const double J22 = dNs[0] * Y[0] + dNs[1] * Y[1] + dNs[2] * Y[2] + dNs[3] * Y[3];
    const double detJ = J11 * J22 - J12 * J21;
    if (fabs(detJ) < 1e-14) return false;
    const double i11 = J22 / detJ, i12 = -J12 / detJ, i21 = -J21 / detJ, i22 = J11 / detJ;
    for (int k = 0; k < 4; ++k) {
        const double gx = i11 * dNr[k] + i12 * dNs[k], gy = i21 * dNr[k] + i22 * dNs[k];
        dNx[k] = ce * gx + se * gy;
        dNy[k] = -se * gx + ce * gy;
    }
    adet = fabs(detJ);
    return true;
}

// element-frame operator on local coordinates lc and element-frame displacements
// ue; Kg (24 x 24 column-major) is overwritten.  Only u, v, w rows are written.
__device__ void q4_kg_local(const double *lc, const double *ue, double E, double nu, double h, bool assumed,
                            double *Kg) {
    for (int q = 0; q < 576; ++q) Kg[q] = 0.0;
    if (h < 1e-30) return;
    const double X[4] = {lc[0], lc[2], lc[4], lc[6]}, Y[4] = {lc[1], lc[3], lc[5], lc[7]};
    // an exact rectangle keeps the centre-gradient transverse form
    const double e12x = X[1] - X[0], e12y = Y[1] - Y[0], e14x = X[3] - X[0], e14y = Y[3] - Y[0];
    const bool exact_rectangle = (e12x != 0.0 || e12y != 0.0) && (e14x != 0.0 || e14y != 0.0) &&
                                 X[2] - X[3] == e12x && Y[2] - Y[3] == e12y && X[2] - X[1] == e14x &&
                                 Y[2] - Y[1] == e14y && e12x * e14x + e12y * e14y == 0.0;
    assumed = assumed && !exact_rectangle;
    const double d13x = X[2] - X[0], d13y = Y[2] - Y[0], l13 = hypot(d13x, d13y);
    const double d42x = X[1] - X[3], d42y = Y[1] - Y[3], l42 = hypot(d42x, d42y);
    if (l13 < 1e-12 || l42 < 1e-12) return;
    const double bx = d13x / l13 + d42x / l42, by = d13y / l13 + d42y / l42, lb = hypot(bx, by);
    if (lb < 1e-12) return;
    const double ce = bx / lb, se = by / lb;
    double up[4], vp[4];
    for (int k = 0; k < 4; ++k) {
        up[k] = ce * ue[6 * k] + se * ue[6 * k + 1];
        vp[k] = -se * ue[6 * k] + ce * ue[6 * k + 1];
    }
    const double d11 = E / (1.0 - nu * nu), d12 = nu * d11, d33 = E / (2.0 * (1.0 + nu));
    double dNx0[4], dNy0[4], a0;
    if (!kg_grad(X, Y, ce, se, 0.0, 0.0, dNx0, dNy0, a0)) return;
    double gam0 = 0.0;                      // centre-sampled membrane shear
    for (int k = 0; k < 4; ++k) gam0 += dNy0[k] * up[k] + dNx0[k] * vp[k];
    const double gp = 1.0 / sqrt(3.0);
    // tying gradients: w,x at (0, -+gp), w,y at (-+gp, 0)
    double tx[4][4], ty[4][4], ta;
    if (assumed) {
        const double tr[4] = {0.0, 0.0, -gp, gp}, ts[4] = {-gp, gp, 0.0, 0.0};
        for (int t = 0; t < 4; ++t)
            if (!kg_grad(X, Y, ce, se, tr[t], ts[t], tx[t], ty[t], ta)) return;
    }
    const double gr[4] = {-gp, gp, gp, -gp}, gs[4] = {-gp, -gp, gp, gp};
    for (int q = 0; q < 4; ++q) {
        double dNx[4], dNy[4], adet;
        if (!kg_grad(X, Y, ce, se, gr[q], gs[q], dNx, dNy, adet)) continue;
        const double w = h * adet;
        double epx = 0.0, epy = 0.0;
        for (int k = 0; k < 4; ++k) { epx += dNx[k] * up[k]; epy += dNy[k] * vp[k]; }
        const double sxx = d11 * epx + d12 * epy, syy = d12 * epx + d11 * epy, sxy = d33 * gam0;
        const double *dWx = assumed ? tx[q <= 1 ? 0 : 1] : dNx;
        const double *dWy = assumed ? ty[(q == 0 || q == 3) ? 2 : 3] : dNy;
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) {
                const int r0 = 6 * i, c0 = 6 * j;
                const double sym_xy = 0.5 * (dNx[i] * dNy[j] + dNy[i] * dNx[j]);
                const double kuu = w * (syy * dNy[i] * dNy[j] + sxy * sym_xy);
                const double kvv = w * (sxx * dNx[i] * dNx[j] + sxy * sym_xy);
                const double kuv = -w * sxy * 0.5 * (dNx[i] * dNx[j] + dNy[i] * dNy[j]);
                Kg[r0 + 24 * c0] += ce * ce * kuu - 2.0 * ce * se * kuv + se * se * kvv;
                Kg[r0 + 24 * (c0 + 1)] += ce * se * kuu + (ce * ce - se * se) * kuv - se * ce * kvv;
                Kg[(r0 + 1) + 24 * c0] += se * ce * kuu + (ce * ce - se * se) * kuv - ce * se * kvv;
                Kg[(r0 + 1) + 24 * (c0 + 1)] += se * se * kuu + 2.0 * se * ce * kuv + ce * ce * kvv;
                Kg[(r0 + 2) + 24 * (c0 + 2)] +=
                    assumed ? w * (sxx * dWx[i] * dWx[j] + syy * dWy[i] * dWy[j] + sxy * (dWx[i] * dWy[j] + dWy[i] * dWx[j]))
                            : w * (sxx * dNx[i] * dNx[j] + syy * dNy[i] * dNy[j] +
                                   sxy * (dNx0[i] * dNy0[j] + dNy0[i] * dNx0[j]));
            }
    }
}

// the finite-warp displacement map Tg (24 x 24 column-major; identity for a flat
// cell), _quad4_finite_warp_fill! (src/FEMKernels.jl:880-945)
__device__ void kg_warp_map(const double *lc, const double *zl, double *Tg) {
    for (int q = 0; q < 576; ++q) Tg[q] = (q % 25 == 0) ? 1.0 : 0.0;
    const double cr[4] = {-1, 1, 1, -1}, cs_[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; ++i) {
        double dNr[4], dNs[4], J11, J12, J21, J22;
        shape_derivs(cr[i], cs_[i], dNr, dNs);
        jac4(dNr, dNs, lc, J11, J12, J21, J22);
        const double dj = J11 * J22 - J12 * J21;
        if (fabs(dj) < 1e-30) continue;
        const double i11 = J22 / dj, i12 = -J12 / dj, i21 = -J21 / dj, i22 = J11 / dj;
        const double zr = dNr[0] * zl[0] + dNr[1] * zl[1] + dNr[2] * zl[2] + dNr[3] * zl[3];
        const double zs = dNs[0] * zl[0] + dNs[1] * zl[1] + dNs[2] * zl[2] + dNs[3] * zl[3];
        const double gx = i11 * zr + i12 * zs, gy = i21 * zr + i22 * zs;
        const int rxi = 6 * i + 3, ryi = 6 * i + 4;
        Tg[rxi + 24 * (6 * i + 5)] += gx;
        Tg[ryi + 24 * (6 * i + 5)] += gy;
        const int a = (i % 2 == 0) ? 0 : 1, b = (i % 2 == 0) ? 2 : 3;
        const double dx = lc[2 * b] - lc[2 * a], dy = lc[2 * b + 1] - lc[2 * a + 1];
        const double dl2 = dx * dx + dy * dy;
        if (dl2 > 1e-30) {
            const int js[2] = {a, b};
            const double cu[2] = {dy / dl2, -dy / dl2}, cv[2] = {-dx / dl2, dx / dl2};
            for (int t = 0; t < 2; ++t) {
                const int j = js[t];
                Tg[rxi + 24 * (6 * j + 0)] -= gx * cu[t];
                Tg[rxi + 24 * (6 * j + 1)] -= gx * cv[t];
                Tg[ryi + 24 * (6 * j + 0)] -= gy * cu[t];
                Tg[ryi + 24 * (6 * j + 1)] -= gy * cv[t];
            }
        }
        for (int j = 0; j < 4; ++j) {
            const double dNdx = i11 * dNr[j] + i12 * dNs[j], dNdy = i21 * dNr[j] + i22 * dNs[j];
            Tg[(6 * i + 0) + 24 * (6 * j + 2)] += zl[i] * dNdx;
            Tg[(6 * i + 1) + 24 * (6 * j + 2)] += zl[i] * dNdy;
        }
    }
}

// out = R' blk R for every 3 x 3 block of the N x N element-frame matrix Kl (R rows
// = v1, v2, v3: the frame axes in global coordinates)
template <int N>
__device__ void kg_rotate_blocks(const double *Kl, const double *R, double *out) {
    for (int B = 0; B < N / 3; ++B)
        for (int A = 0; A < N / 3; ++A) {
            double t[9];
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    double s = 0;
                    for (int l = 0; l < 3; ++l) s += Kl[(3 * A + i) + N * (3 * B + l)] * R[l * 3 + j];
                    t[i + 3 * j] = s;
                }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    double s = 0;
                    for (int l = 0; l < 3; ++l) s += R[l * 3 + i] * t[l + 3 * j];
                    out[(3 * A + i) + N * (3 * B + j)] = s;
                }
        }
}

// global 24 x 24 Kg of one CQUAD4 (column-major into out); returns a status code
__device__ int q4_kg(const double *p, const double *pr, const double *ug, double *out) {
    {   // flat-element gate: corner deviation below 1e-2 of the longer diagonal
        const double d13[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
        const double d24[3] = {p[9] - p[3], p[10] - p[4], p[11] - p[5]};
        double n[3];
        cross3(d13, d24, n);
        const double nl = sqrt(dot3(n, n));
        if (nl > 1e-12) {
            const double v3g[3] = {n[0] / nl, n[1] / nl, n[2] / nl};
            double cg[3], dev = 0.0;
            for (int k = 0; k < 3; ++k) cg[k] = (((p[k] + p[3 + k]) + p[6 + k]) + p[9 + k]) / 4.0;
            for (int k = 0; k < 4; ++k) {
                const double d[3] = {p[3 * k] - cg[0], p[3 * k + 1] - cg[1], p[3 * k + 2] - cg[2]};
                dev = fmax(dev, fabs(dot3(d, v3g)));
            }
            const double ld = fmax(sqrt(dot3(d13, d13)), sqrt(dot3(d24, d24)));
            if (!(dev < 1e-2 * fmax(ld, 1e-12))) return JFEM_ERR_WARPED;
        }
    }
    Q4Prep P;
    q4_setup(P, p, pr, 0.0);
    double ul[24], uk[24], Tg[576], Kl[576];
    for (int k = 0; k < 4; ++k)
        for (int hf = 0; hf < 2; ++hf)
            for (int i = 0; i < 3; ++i) {
                const double *s = ug + 6 * k + 3 * hf;
                ul[6 * k + 3 * hf + i] = P.R[3 * i] * s[0] + P.R[3 * i + 1] * s[1] + P.R[3 * i + 2] * s[2];
            }
    kg_warp_map(P.lc, P.zl, Tg);            // preload map (identity on a flat cell)
    for (int i = 0; i < 24; ++i) {
        double s = 0.0;
        for (int m = 0; m < 24; ++m) s += Tg[i + 24 * m] * ul[m];
        uk[i] = s;
    }
    q4_kg_local(P.lc, uk, pr[JFEM_SP_E], pr[JFEM_SP_NU], pr[JFEM_SP_H], true, Kl);
    if (P.warped) {                         // sym(W' Kg W); Tg reused for W' (Kg W)
        double W[576];
        for (int j = 0; j < 24; ++j)
            for (int i = 0; i < 24; ++i) {
                double v = 0.0;
                for (int m = 0; m < 24; ++m) v += Kl[i + 24 * m] * Tg[m + 24 * j];
                W[i + 24 * j] = v;
            }
        for (int j = 0; j < 24; ++j)
            for (int i = 0; i < 24; ++i) {
                double v = 0.0;
                for (int l = 0; l < 24; ++l) v += Tg[l + 24 * i] * W[l + 24 * j];
                Kl[i + 24 * j] = v;
            }
        for (int j = 0; j < 24; ++j)
            for (int i = 0; i < j; ++i) {
                const double s = 0.5 * (Kl[i + 24 * j] + Kl[j + 24 * i]);
                Kl[i + 24 * j] = s;
                Kl[j + 24 * i] = s;
            }
    }
    kg_rotate_blocks<24>(Kl, P.R, out);
    return 0;
}

// principal stresses of (sxx, syy, sxy) and their directions (c, s), the second
// direction (-s1, c1): principal_stress_2d_components (src/FEMKernels.jl:8090-8101)
__device__ void kg_principal(double sxx, double syy, double sxy, double *lam, double *dc, double *ds) {
    const double mean_s = 0.5 * (sxx + syy), half_d = 0.5 * (sxx - syy);
    const double radius = sqrt(half_d * half_d + sxy * sxy);
    if (radius <= 1e-30) {
        lam[0] = mean_s; dc[0] = 1.0; ds[0] = 0.0;
        lam[1] = mean_s; dc[1] = 0.0; ds[1] = 1.0;
        return;
    }
    const double theta = 0.5 * atan2(2.0 * sxy, sxx - syy);
    const double c = cos(theta), s = sin(theta);
    lam[0] = mean_s + radius; dc[0] = c; ds[0] = s;
    lam[1] = mean_s - radius; dc[1] = -s; ds[1] = c;
}

// element-frame CTRIA3 Kg (18 x 18, column-major) for membrane stresses sig on a
// triangle of local corners lc: per node pair, the in-plane DOFs take the principal
// directions of each stress component alone and w takes those of the whole stress
// (geometric_stiffness_tria3 in principal-transverse mode, src/FEMKernels.jl:
// 11142-11288 with 8205-8338); drill rows stay zero
__device__ void t3_kg_local(const double *lc, const double *sig, double h, double *Kl) {
    for (int q = 0; q < 324; ++q) Kl[q] = 0.0;
    if (h < 1e-30) return;
    const double x1 = lc[0], y1 = lc[1], x2 = lc[2], y2 = lc[3], x3 = lc[4], y3 = lc[5];
    const double A = 0.5 * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
    if (A < 1e-12) return;
    const double bv[3] = {(y2 - y3) / (2 * A), (y3 - y1) / (2 * A), (y1 - y2) / (2 * A)};
    const double cv[3] = {(x3 - x2) / (2 * A), (x1 - x3) / (2 * A), (x2 - x1) / (2 * A)};
    const double sc = h * A;
    // four stress states: sxx, syy, sxy alone (in-plane), the whole stress (w)
    const double st[4][3] = {{sig[0], 0.0, 0.0}, {0.0, sig[1], 0.0}, {0.0, 0.0, sig[2]}, {sig[0], sig[1], sig[2]}};
    double lam[4][2], dc[4][2], ds[4][2];
    for (int k = 0; k < 4; ++k) kg_principal(st[k][0], st[k][1], st[k][2], lam[k], dc[k], ds[k]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int r0 = 6 * i, c0 = 6 * j;
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 2; ++d) {
                    if (fabs(lam[k][d]) <= 1e-30) continue;
                    const double c = dc[k][d], s = ds[k][d];
                    const double gi = c * bv[i] + s * cv[i], gj = c * bv[j] + s * cv[j];
                    const double val = sc * lam[k][d] * gi * gj;
                    if (k < 3) {
                        const double p11 = s * s, p22 = 1.0 * c * c, p12 = 1.0 * -c * s;
                        Kl[r0 + 18 * c0] += val * p11;
                        Kl[r0 + 18 * (c0 + 1)] += val * p12;
                        Kl[r0 + 1 + 18 * c0] += val * p12;
                        Kl[r0 + 1 + 18 * (c0 + 1)] += val * p22;
                    } else {
                        Kl[r0 + 2 + 18 * (c0 + 2)] += 1.0 * val;
                    }
                }
        }
}

// global 18 x 18 Kg of one CTRIA3 (column-major into out) under the membrane
// resultants that stress_strain_tria3 gives for ug; returns a status code
__device__ int t3_kg(const double *p, const double *pr, const double *ug, double *out) {
    const double *p1 = p, *p2 = p + 3, *p3 = p + 6;
    double v1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]}, e13[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    double v2[3], v3[3];
    normalize3(v1);
    cross3(v1, e13, v3);
    normalize3(v3);
    cross3(v3, v1, v2);
    double c[3], lc[6];
    for (int k = 0; k < 3; ++k) c[k] = ((p1[k] + p2[k]) + p3[k]) / 3.0;
    for (int k = 0; k < 3; ++k) {
        const double d[3] = {p[3 * k] - c[0], p[3 * k + 1] - c[1], p[3 * k + 2] - c[2]};
        lc[2 * k] = dot3(d, v1); lc[2 * k + 1] = dot3(d, v2);
    }
    double rec[20];
    const int st = t3_recover(p, pr, ug, rec);
    if (st != 0) return st;
    const double h = pr[JFEM_SP_H];
    const double sig[3] = {rec[0] / h, rec[1] / h, rec[2] / h};
    double Kl[324];
    t3_kg_local(lc, sig, h, Kl);
    const double R[9] = {v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
    kg_rotate_blocks<18>(Kl, R, out);
    return 0;
}

// c = a + sigma b over the values of matrices sharing one pattern (K + sigma Kg)
__global__ void csr_shift_kernel(int64_t n, const double *__restrict__ a, double sigma, const double *__restrict__ b,
                                 double *__restrict__ c) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = a[i] + sigma * b[i];
}

// the scale 1/sqrt(K_ii) per row (1 where the diagonal is absent or not positive)
__global__ void diag_scale_kernel(int64_t n, const int32_t *__restrict__ diagpos, const double *__restrict__ a,
                                  double *__restrict__ sd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t p = diagpos[i];
    const double d = p >= 0 ? a[p] : 0.0;
    sd[i] = d > 0.0 ? 1.0 / sqrt(d) : 1.0;
}

// c = s_i (a + sigma b) s_j over CSR rows, s the diagonal scale of a
__global__ void csr_shift_scaled_kernel(int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                        const double *__restrict__ a, double sigma, const double *__restrict__ b,
                                        const double *__restrict__ sd, double *__restrict__ c) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double si = sd[i];
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) c[p] = (si * (a[p] + sigma * b[p])) * sd[col[p]];
}

// one thread per CQUAD4: its global Kg at Ke + 576 e (the layout the assembly sums)
__global__ void __launch_bounds__(64)
q4_kg_kernel(int64_t nq, const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
             const int32_t *__restrict__ qconn, const double *__restrict__ qprop, const double *__restrict__ u,
             double *__restrict__ Ke, int64_t *status) {
    const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (e >= nq) return;
    double p[12], pr[8], ug[24];
    for (int k = 0; k < 4; ++k) {
        const int64_t nd = qconn[k * nq + e];
        p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd];
        for (int c = 0; c < 6; ++c) ug[6 * k + c] = u[6 * nd + c];
    }
    for (int c = 0; c < 8; ++c) pr[c] = qprop[c * nq + e];
    const int st = q4_kg(p, pr, ug, Ke + 576 * e);
    if (st != 0) report(status, e, st);
}

// one thread per CTRIA3: its global Kg at Ke + 324 e (Ke: the CTRIA3 part of the
// triplet layout); failures reported with the index nq + e
__global__ void __launch_bounds__(64)
t3_kg_kernel(int64_t nq, int64_t nt, const double *__restrict__ x, const double *__restrict__ y,
             const double *__restrict__ z, const int32_t *__restrict__ tconn, const double *__restrict__ tprop,
             const double *__restrict__ u, double *__restrict__ Ke, int64_t *status) {
    const int64_t e = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (e >= nt) return;
    double p[9], pr[8], ug[18];
    for (int k = 0; k < 3; ++k) {
        const int64_t nd = tconn[k * nt + e];
        p[3 * k] = x[nd]; p[3 * k + 1] = y[nd]; p[3 * k + 2] = z[nd];
        for (int c = 0; c < 6; ++c) ug[6 * k + c] = u[6 * nd + c];
    }
    for (int c = 0; c < 8; ++c) pr[c] = tprop[c * nt + e];
    const int st = t3_kg(p, pr, ug, Ke + 324 * e);
    if (st != 0) report(status, nq + e, st);
}


namespace jfem {

hipError_t launch_reset_status(hipStream_t s, int64_t *status) {
    clear_stale_hip_error();                           // (the first launch of an element or assembly call)
    hipLaunchKernelGGL(reset_status_kernel, dim3(1), dim3(1), 0, s, status);
    return hipGetLastError();
}

// element matrices of the kinds in kinds_mask (1 CQUAD4, 2 CTRIA3, 4 CTETRA10) for V
// section sets (variant v's records at qp / tp / tetp + v * (records of a set),
// its matrices at Ke + v * Ke_stride); within one set the kinds follow each other
// when several are generated (triplet order)
hipError_t launch_elements(hipStream_t s, const jfem_mesh *m, const double *qp, const double *tp, const double *tetp,
                           double *Ke, int64_t Ke_stride, int V, int kinds_mask, int64_t *status) {
    const int64_t total = m->nq + m->nt + m->ntet;
    int64_t off = 0;
    if ((kinds_mask & 1) && m->nq) {
        const ElemVar ev{8 * m->nq, Ke_stride, total, 0};
        hipLaunchKernelGGL(q4_prep_global_kernel, dim3((unsigned)((m->nq + 63) / 64), V), dim3(64), 0, s, m->nq, m->d_x,
                           m->d_y, m->d_z, m->d_q, qp, m->k6rot, Ke, ev);
        dim3 grid((unsigned)((m->nq + kElemsPerBlock - 1) / kElemsPerBlock), V);
        hipLaunchKernelGGL(q4_elem_kernel, grid, dim3(kWave * kElemsPerBlock), 0, s, m->nq, Ke, status, ev);
        off += m->nq * 576;
    }
    if ((kinds_mask & 2) && m->nt) {
        const ElemVar ev{8 * m->nt, Ke_stride, total, m->nq};
        dim3 grid((unsigned)((m->nt + kElemsPerBlock - 1) / kElemsPerBlock), V);
        hipLaunchKernelGGL(t3_kernel<false>, grid, dim3(kWave * kElemsPerBlock), 0, s, m->nt, m->d_x, m->d_y, m->d_z,
                           m->d_t, tp, m->k6rot, T3LocalIn{}, Ke + off, status, ev);
        off += m->nt * 324;
    }
    if ((kinds_mask & 4) && m->ntet) {
        const ElemVar ev{2 * m->ntet, Ke_stride, total, m->nq + m->nt};
        dim3 grid((unsigned)((m->ntet + kElemsPerBlock - 1) / kElemsPerBlock), V);
        hipLaunchKernelGGL(tet10_kernel<false>, grid, dim3(kWave * kElemsPerBlock), 0, s, m->ntet, m->d_x, m->d_y,
                           m->d_z, m->d_tet, nullptr, tetp, Ke + off, status, ev);
    }
    return hipGetLastError();
}

hipError_t launch_t3_local(hipStream_t s, int64_t n, const double *lc, const double *Dm, const double *Db,
                           const double *Ds, const double *k6, double *Ke, int64_t *status) {
    if (n == 0) return hipSuccess;
    dim3 grid((unsigned)((n + kElemsPerBlock - 1) / kElemsPerBlock));
    hipLaunchKernelGGL(t3_kernel<true>, grid, dim3(kWave * kElemsPerBlock), 0, s, n, nullptr, nullptr, nullptr, nullptr,
                       nullptr, 0.0, T3LocalIn{lc, Dm, Db, Ds, k6}, Ke, status, ElemVar{});
    return hipGetLastError();
}

hipError_t launch_shell_recover(hipStream_t s, const jfem_mesh *m, const double *u, int mean_shear, double *qout,
                                double *tout, int64_t *status) {
    const int64_t n = m->nq + m->nt;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(shell_recover_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, m->nq, m->nt, m->d_x,
                       m->d_y, m->d_z, m->d_q, m->d_qp, m->d_t, m->d_tp, u, mean_shear, qout, tout, status);
    return hipGetLastError();
}

hipError_t launch_csr_shift(hipStream_t s, int64_t n, const double *a, double sigma, const double *b, double *c) {
    if (n == 0) return hipSuccess;
    const int bs = 256;
    hipLaunchKernelGGL(csr_shift_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, n, a, sigma, b, c);
    return hipGetLastError();
}

hipError_t launch_csr_shift_scaled(hipStream_t s, const jfem_pattern *P, const int32_t *d_diagpos, double *d_sd,
                                   const double *a, double sigma, const double *b, double *c) {
    const int64_t n = P->ndof;
    if (n == 0) return hipSuccess;
    const int bs = 256;
    hipLaunchKernelGGL(diag_scale_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, n, d_diagpos, a, d_sd);
    hipLaunchKernelGGL(csr_shift_scaled_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, n, P->d_rowptr,
                       P->d_col, a, sigma, b, d_sd, c);
    return hipGetLastError();
}

hipError_t launch_shell_kg(hipStream_t s, const jfem_mesh *m, const double *u, double *Ke, int64_t *status) {
    if (m->nq)
        hipLaunchKernelGGL(q4_kg_kernel, dim3((unsigned)((m->nq + 63) / 64)), dim3(64), 0, s, m->nq, m->d_x, m->d_y,
                           m->d_z, m->d_q, m->d_qp, u, Ke, status);
    if (m->nt)
        hipLaunchKernelGGL(t3_kg_kernel, dim3((unsigned)((m->nt + 63) / 64)), dim3(64), 0, s, m->nq, m->nt, m->d_x,
                           m->d_y, m->d_z, m->d_t, m->d_tp, u, Ke + 576 * m->nq, status);
    return hipGetLastError();
}

hipError_t launch_tet10_direct(hipStream_t s, int64_t n, const double *X, const double *prop, double *Ke,
                               int64_t *status) {
    if (n == 0) return hipSuccess;
    dim3 grid((unsigned)((n + kElemsPerBlock - 1) / kElemsPerBlock));
    hipLaunchKernelGGL(tet10_kernel<true>, grid, dim3(kWave * kElemsPerBlock), 0, s, n, nullptr, nullptr, nullptr,
                       nullptr, X, prop, Ke, status, ElemVar{});
    return hipGetLastError();
}

hipError_t launch_q4_local(hipStream_t s, int64_t n, const double *lc, const double *Cm, const double *Cb,
                           const double *Cs, const double *h, const double *k6, const double *ds,
                           const int32_t *flags, double *Ke, int64_t *status) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(q4_prep_local_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, n, lc, Cm, Cb, Cs, h, k6,
                       ds, flags, Ke);
    dim3 grid((unsigned)((n + kElemsPerBlock - 1) / kElemsPerBlock));
    hipLaunchKernelGGL(q4_elem_kernel, grid, dim3(kWave * kElemsPerBlock), 0, s, n, Ke, status, ElemVar{});
    return hipGetLastError();
}

hipError_t launch_assemble(hipStream_t s, const jfem_pattern *P, const double *Ke, double *val, int V,
                           int64_t Ke_stride, int64_t val_stride) {
    if (P->nnz == 0) return hipSuccess;
    const int bs = 256;
    dim3 grid((unsigned)((P->nnz + bs - 1) / bs), V);
    hipLaunchKernelGGL(assemble_kernel, grid, dim3(bs), 0, s, P->nnz, P->d_seg_ptr, P->d_seg_src, Ke, val, Ke_stride,
                       val_stride);
    return hipGetLastError();
}

hipError_t launch_bc_screen(hipStream_t s, const jfem_pattern *P, const double *val, const uint8_t *spc, double thresh,
                            uint8_t *fixed0, uint8_t *flag) {
    const int64_t nn = P->n_nodes;
    if (nn == 0) return hipSuccess;
    const int bs = 128;
    hipLaunchKernelGGL(bc_screen_kernel, dim3((unsigned)((nn + bs - 1) / bs)), dim3(bs), 0, s, nn, P->d_nodeblk, val, spc,
                       thresh, fixed0, flag);
    return hipGetLastError();
}

hipError_t launch_bc_decide(hipStream_t s, const jfem_pattern *P, const double *val, const uint8_t *flag,
                            double eps_zero, uint8_t *fixed) {
    const int64_t nn = P->n_nodes;
    if (nn == 0) return hipSuccess;
    const int bs = 128;
    hipLaunchKernelGGL(bc_decide_kernel, dim3((unsigned)((nn + bs - 1) / bs)), dim3(bs), 0, s, nn, P->d_nodeblk, val,
                       flag, eps_zero, fixed);
    return hipGetLastError();
}

hipError_t launch_gather_nodeblk(hipStream_t s, const jfem_pattern *P, const double *val, double *out) {
    int64_t n = P->n_nodes * 36;
    if (n == 0) return hipSuccess;
    const int bs = 256;
    hipLaunchKernelGGL(gather_nodeblk_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, n, P->d_nodeblk,
                       val, out);
    return hipGetLastError();
}

}  // namespace jfem
