// Load vectors on the device: FORCE / MOMENT cards on grid points, uniform
// PLOAD4 cards on shell elements and GRAV cards (lumped shell mass), one load
// set (SID) in card order.
//
// Reference: _resolve_loads! (src/solver/loads.jl:510-587) accumulates, per
// subcase, every PLOAD4 on a CQUAD4/CTRIA3 straight into F (the element's
// nodes in connectivity order, cards in input order), and the FORCE / MOMENT
// cards into a per-grid 6-vector that is added to F after all distributed loads
// (:1338-1340).  Per PLOAD4 card (:541-587):
//   n = (X3 - X1) x (X4 - X2) (quad) or (X2 - X1) x (X3 - X1) (tria),
//   area = 0.5 norm(n), skipped unless area > 1e-30,
//   d = N ./ norm(N) when a direction N with norm > 1e-30 is given, else normalize(n),
//   each node gets d .* ((area * P * scale) / nn) on its three translations;
// per FORCE card (:525-530): F[g, 1:3] += (Dir * Mag) * scale (MOMENT: F[g, 4:6]).
// norm and normalize are Julia's for a 3-vector: the plain sum of squares below
// the BLAS cutoff (scaled only when it would overflow or underflow), and the
// vector times the correctly rounded reciprocal of its norm.
//
// Layout: card data as given; per grid point the PLOAD4 cards touching it (card
// order), its FORCE records and its MOMENT records (card order), CSR by grid
// point, built once per load set on the host.  Two launches per load vector:
// one thread per PLOAD4 card (nodal share, 3 doubles + a flag), then one thread
// per grid point summing in the reference's order -- no atomics, so the result
// is bitwise reproducible and independent of the launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "jfem_internal.h"

struct jfem_loadset {
    int device = 0;
    const jfem_mesh *mesh = nullptr;
    int64_t n_nodes = 0, npl = 0, nf = 0, nm = 0;
    int32_t *d_pl_elem = nullptr;   // npl
    double *d_pl_p = nullptr;       // npl
    double *d_pl_dir = nullptr;     // 3 npl (null: element normals)
    double *d_plf = nullptr;        // 4 npl: nodal share (3) and 1.0 if the card applies
    int64_t *d_pl_ptr = nullptr;    // n_nodes + 1
    int32_t *d_pl_card = nullptr;   // PLOAD4 cards per grid point, card order
    int64_t *d_f_ptr = nullptr;     // n_nodes + 1
    double *d_f_rec = nullptr;      // 4 per FORCE record (mag, dir), per grid point in card order
    int64_t *d_m_ptr = nullptr;
    double *d_m_rec = nullptr;
    double *d_F = nullptr;          // 6 n_nodes: staging of the host variant
    // GRAV on shells: per card (A, N1, N2, N3), per shell its areal mass
    // (rho T + NSM; 0: no mass) and lumped nodal share, per grid point its shells
    int64_t ng = 0, nsh = 0;
    double *d_grav = nullptr;       // 4 ng
    double *d_areal = nullptr;      // nsh
    double *d_gsh = nullptr;        // nsh
    int64_t *d_sh_ptr = nullptr;    // n_nodes + 1
    int32_t *d_sh_elem = nullptr;   // shells per grid point, element order
};

namespace {

constexpr int kLoadBS = 256;

// Julia's norm of a 3-vector (LinearAlgebra.generic_norm2)
__device__ double jl_norm3(const double *v) {
#pragma clang fp contract(off)
    const double a0 = fabs(v[0]), a1 = fabs(v[1]), a2 = fabs(v[2]);
    double mx = a0;
    if (a1 > mx || isnan(a1)) mx = a1;
    if (a2 > mx || isnan(a2)) mx = a2;
    if (mx == 0.0 || isinf(mx) || isnan(mx)) return mx;
    const double t = 3.0 * mx * mx;
    if (isfinite(t) && mx * mx != 0.0) {
        double s = v[0] * v[0];
        s = s + v[1] * v[1];
        s = s + v[2] * v[2];
        return sqrt(s);
    }
    double s = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double q = fabs(v[i]) / mx;
        s = s + q * q;
    }
    return mx * sqrt(s);
}

// Julia's normalize (LinearAlgebra.__normalize!): times inv(norm), or in two
// steps when the reciprocal would overflow
__device__ void jl_normalize3(double *v) {
#pragma clang fp contract(off)
    const double nrm = jl_norm3(v);
    const double delta = 1.0 / DBL_MAX;                  // inv(prevfloat(typemax(Float64)))
    if (nrm >= delta) {
        const double inv = 1.0 / nrm;
        for (int i = 0; i < 3; ++i) v[i] = v[i] * inv;
    } else {
        const double ed = DBL_EPSILON / delta;
        for (int i = 0; i < 3; ++i) v[i] = v[i] * ed;
        const double inv = 1.0 / (nrm * ed);
        for (int i = 0; i < 3; ++i) v[i] = v[i] * inv;
    }
}

__global__ void __launch_bounds__(kLoadBS) pload4_kernel(int64_t npl, const int32_t *__restrict__ elem,
                                                         const double *__restrict__ p, const double *__restrict__ dir,
                                                         double scale, const double *__restrict__ x,
                                                         const double *__restrict__ y, const double *__restrict__ z,
                                                         const int32_t *__restrict__ qc, int64_t nq,
                                                         const int32_t *__restrict__ tc, int64_t nt,
                                                         double *__restrict__ plf) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npl) return;
    const int64_t e = elem[i];
    const bool quad = e < nq;
    const int nn = quad ? 4 : 3;
    double X[4][3];
    for (int k = 0; k < nn; ++k) {
        const int32_t g = quad ? qc[e + nq * k] : tc[(e - nq) + nt * k];
        X[k][0] = x[g]; X[k][1] = y[g]; X[k][2] = z[g];
    }
    double a[3], b[3], n[3];
    for (int c = 0; c < 3; ++c) {
        a[c] = quad ? X[2][c] - X[0][c] : X[1][c] - X[0][c];
        b[c] = quad ? X[3][c] - X[1][c] : X[2][c] - X[0][c];
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    const double area = 0.5 * jl_norm3(n);
    double *out = plf + 4 * i;
    if (!(area > 1e-30)) {
        out[0] = out[1] = out[2] = 0.0;
        out[3] = 0.0;
        return;
    }
    double d[3];
    bool given = false;
    if (dir) {
        const double nd[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
        const double nrm = jl_norm3(nd);
        if (nrm > 1e-30) {
            for (int c = 0; c < 3; ++c) d[c] = nd[c] / nrm;
            given = true;
        }
    }
    if (!given) {
        for (int c = 0; c < 3; ++c) d[c] = n[c];
        jl_normalize3(d);
    }
    const double tf = area * p[i] * scale;
    const double share = tf / (double)nn;
    for (int c = 0; c < 3; ++c) out[c] = d[c] * share;
    out[3] = 1.0;
}

// lumped gravity share of each shell with mass: (areal * area) / nn, area half
// the norm of the diagonal (CQUAD4) or edge (CTRIA3) cross product
// (src/solver/loads.jl:758-766)
__global__ void __launch_bounds__(kLoadBS) grav_share_kernel(int64_t nsh, const double *__restrict__ areal,
                                                             const double *__restrict__ x, const double *__restrict__ y,
                                                             const double *__restrict__ z, const int32_t *__restrict__ qc,
                                                             int64_t nq, const int32_t *__restrict__ tc, int64_t nt,
                                                             double *__restrict__ gsh) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nsh) return;
    const bool quad = e < nq;
    const int nn = quad ? 4 : 3;
    double X[4][3];
    for (int k = 0; k < nn; ++k) {
        const int32_t g = quad ? qc[e + nq * k] : tc[(e - nq) + nt * k];
        X[k][0] = x[g]; X[k][1] = y[g]; X[k][2] = z[g];
    }
    double a[3], b[3], n[3];
    for (int c = 0; c < 3; ++c) {
        a[c] = quad ? X[2][c] - X[0][c] : X[1][c] - X[0][c];
        b[c] = quad ? X[3][c] - X[1][c] : X[2][c] - X[0][c];
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    const double area = 0.5 * jl_norm3(n);
    const double total_mass = areal[e] * area;
    gsh[e] = total_mass / (double)nn;
}

__global__ void __launch_bounds__(kLoadBS) load_node_kernel(int64_t n_nodes, const int64_t *__restrict__ pl_ptr,
                                                            const int32_t *__restrict__ pl_card,
                                                            const double *__restrict__ plf,
                                                            const int64_t *__restrict__ f_ptr,
                                                            const double *__restrict__ f_rec,
                                                            const int64_t *__restrict__ m_ptr,
                                                            const double *__restrict__ m_rec, double scale,
                                                            int64_t ng, const double *__restrict__ grav,
                                                            const int64_t *__restrict__ sh_ptr,
                                                            const int32_t *__restrict__ sh_elem,
                                                            const double *__restrict__ areal,
                                                            const double *__restrict__ gsh,
                                                            double *__restrict__ F) {
#pragma clang fp contract(off)
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_nodes) return;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (pl_ptr)
        for (int64_t q = pl_ptr[g]; q < pl_ptr[g + 1]; ++q) {
            const double *s = plf + 4 * (int64_t)pl_card[q];
            if (s[3] == 0.0) continue;
            for (int c = 0; c < 3; ++c) v[c] = v[c] + s[c];
        }
    // GRAV cards in order, each over the grid point's shells in element order
    // (the shell loop of src/solver/loads.jl:730-772)
    for (int64_t i = 0; i < ng; ++i) {
        const double *gr = grav + 4 * i;
        double gv[3];
        for (int c = 0; c < 3; ++c) gv[c] = (gr[0] * scale) * gr[1 + c];
        for (int64_t q = sh_ptr[g]; q < sh_ptr[g + 1]; ++q) {
            const int32_t e = sh_elem[q];
            if (areal[e] == 0.0) continue;
            const double sh = gsh[e];
            for (int c = 0; c < 3; ++c) v[c] = v[c] + sh * gv[c];
        }
    }
    const int64_t f0 = f_ptr ? f_ptr[g] : 0, f1 = f_ptr ? f_ptr[g + 1] : 0;
    const int64_t m0 = m_ptr ? m_ptr[g] : 0, m1 = m_ptr ? m_ptr[g + 1] : 0;
    if (f1 > f0 || m1 > m0) {
        double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t q = f0; q < f1; ++q) {
            const double *s = f_rec + 4 * q;
            for (int c = 0; c < 3; ++c) r[c] = r[c] + s[1 + c] * s[0] * scale;
        }
        for (int64_t q = m0; q < m1; ++q) {
            const double *s = m_rec + 4 * q;
            for (int c = 0; c < 3; ++c) r[3 + c] = r[3 + c] + s[1 + c] * s[0] * scale;
        }
        for (int c = 0; c < 6; ++c) v[c] = v[c] + r[c];
    }
    for (int c = 0; c < 6; ++c) F[6 * g + c] = v[c];
}

template <class T>
hipError_t upload_n(T **dst, const T *src, size_t n, hipStream_t s) {
    *dst = nullptr;
    if (n == 0 || !src) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, n * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}

// per grid point, in card order: the indices i of the cards whose node list holds it
std::vector<int64_t> node_ptr(int64_t n_nodes, const std::vector<int32_t> &node_of, std::vector<int32_t> &order) {
    std::vector<int64_t> ptr(n_nodes + 1, 0);
    for (int32_t g : node_of) ++ptr[g + 1];
    for (int64_t g = 0; g < n_nodes; ++g) ptr[g + 1] += ptr[g];
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    order.assign(node_of.size(), 0);
    for (size_t i = 0; i < node_of.size(); ++i) order[pos[node_of[i]]++] = (int32_t)i;
    return ptr;
}

}  // namespace

extern "C" int jfem_loadset_free(jfem_loadset *ls) {
    if (!ls) return JFEM_OK;
    (void)hipSetDevice(ls->device);
    for (void *p : {(void *)ls->d_pl_elem, (void *)ls->d_pl_p, (void *)ls->d_pl_dir, (void *)ls->d_plf,
                    (void *)ls->d_pl_ptr, (void *)ls->d_pl_card, (void *)ls->d_f_ptr, (void *)ls->d_f_rec,
                    (void *)ls->d_m_ptr, (void *)ls->d_m_rec, (void *)ls->d_F, (void *)ls->d_grav,
                    (void *)ls->d_areal, (void *)ls->d_gsh, (void *)ls->d_sh_ptr, (void *)ls->d_sh_elem})
        if (p) (void)hipFree(p);
    delete ls;
    return JFEM_OK;
}

extern "C" int jfem_loadset_create(jfem_ctx *ctx, const jfem_mesh *mesh, int64_t n_pload4, const int32_t *pload4_elem,
                                   const double *pload4_p, const double *pload4_dir3, int64_t n_force,
                                   const int32_t *force_node, const double *force_mag, const double *force_dir3,
                                   int64_t n_moment, const int32_t *moment_node, const double *moment_mag,
                                   const double *moment_dir3, jfem_loadset **out) {
    using jfem::fail;
    if (!out) return fail(ctx, JFEM_ERR_BAD_ARG, "null output handle");
    *out = nullptr;
    if (!mesh) return fail(ctx, JFEM_ERR_BAD_ARG, "null mesh");
    if (n_pload4 < 0 || n_force < 0 || n_moment < 0) return fail(ctx, JFEM_ERR_BAD_ARG, "negative card count");
    if ((n_pload4 && (!pload4_elem || !pload4_p)) || (n_force && (!force_node || !force_mag || !force_dir3)) ||
        (n_moment && (!moment_node || !moment_mag || !moment_dir3)))
        return fail(ctx, JFEM_ERR_BAD_ARG, "missing card arrays");
    const int64_t nsh = mesh->nq + mesh->nt, nn = mesh->n_nodes;
    for (int64_t i = 0; i < n_pload4; ++i)
        if (pload4_elem[i] < 0 || pload4_elem[i] >= nsh)
            return fail(ctx, JFEM_ERR_BAD_ARG, "PLOAD4 card references a shell element out of range");
    for (int64_t i = 0; i < n_force; ++i)
        if (force_node[i] < 0 || force_node[i] >= nn)
            return fail(ctx, JFEM_ERR_BAD_ARG, "FORCE card references a grid point out of range");
    for (int64_t i = 0; i < n_moment; ++i)
        if (moment_node[i] < 0 || moment_node[i] >= nn)
            return fail(ctx, JFEM_ERR_BAD_ARG, "MOMENT card references a grid point out of range");
    if (n_pload4 > INT32_MAX || n_force > INT32_MAX || n_moment > INT32_MAX)
        return fail(ctx, JFEM_ERR_BAD_ARG, "too many cards");
    // PLOAD4 cards per grid point: (node, card) pairs in card order, nodes in
    // connectivity order
    std::vector<int32_t> pl_node, pl_cardof;
    for (int64_t i = 0; i < n_pload4; ++i) {
        const int64_t e = pload4_elem[i];
        const bool quad = e < mesh->nq;
        const int k_n = quad ? 4 : 3;
        for (int k = 0; k < k_n; ++k) {
            pl_node.push_back(quad ? mesh->q[e + mesh->nq * k] : mesh->t[(e - mesh->nq) + mesh->nt * k]);
            pl_cardof.push_back((int32_t)i);
        }
    }
    std::vector<int32_t> ord;
    std::vector<int64_t> pl_ptr = node_ptr(nn, pl_node, ord);
    std::vector<int32_t> pl_card(ord.size());
    for (size_t q = 0; q < ord.size(); ++q) pl_card[q] = pl_cardof[ord[q]];
    auto records = [&](int64_t n, const int32_t *node, const double *mag, const double *dir3,
                        std::vector<int64_t> &ptr, std::vector<double> &rec) {
        std::vector<int32_t> nodes(node, node + n), o;
        ptr = node_ptr(nn, nodes, o);
        rec.resize(4 * (size_t)n);
        for (size_t q = 0; q < o.size(); ++q) {
            const int32_t i = o[q];
            rec[4 * q] = mag[i];
            for (int c = 0; c < 3; ++c) rec[4 * q + 1 + c] = dir3[3 * (int64_t)i + c];
        }
    };
    std::vector<int64_t> f_ptr, m_ptr;
    std::vector<double> f_rec, m_rec;
    records(n_force, force_node, force_mag, force_dir3, f_ptr, f_rec);
    records(n_moment, moment_node, moment_mag, moment_dir3, m_ptr, m_rec);

    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    jfem_loadset *ls = new jfem_loadset();
    ls->device = ctx->device;
    ls->mesh = mesh;
    ls->n_nodes = nn; ls->npl = n_pload4; ls->nf = n_force; ls->nm = n_moment;
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = upload_n(&ls->d_pl_elem, pload4_elem, (size_t)n_pload4, s);
    if (e == hipSuccess) e = upload_n(&ls->d_pl_p, pload4_p, (size_t)n_pload4, s);
    if (e == hipSuccess) e = upload_n(&ls->d_pl_dir, pload4_dir3, pload4_dir3 ? 3 * (size_t)n_pload4 : 0, s);
    if (e == hipSuccess && n_pload4) e = hipMalloc((void **)&ls->d_plf, 4 * (size_t)n_pload4 * sizeof(double));
    if (e == hipSuccess && n_pload4) e = upload_n(&ls->d_pl_ptr, pl_ptr.data(), pl_ptr.size(), s);
    if (e == hipSuccess) e = upload_n(&ls->d_pl_card, pl_card.data(), pl_card.size(), s);
    if (e == hipSuccess && n_force) e = upload_n(&ls->d_f_ptr, f_ptr.data(), f_ptr.size(), s);
    if (e == hipSuccess) e = upload_n(&ls->d_f_rec, f_rec.data(), f_rec.size(), s);
    if (e == hipSuccess && n_moment) e = upload_n(&ls->d_m_ptr, m_ptr.data(), m_ptr.size(), s);
    if (e == hipSuccess) e = upload_n(&ls->d_m_rec, m_rec.data(), m_rec.size(), s);
    if (e == hipSuccess) e = hipMalloc((void **)&ls->d_F, 6 * (size_t)nn * sizeof(double));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        jfem_loadset_free(ls);
        return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    }
    *out = ls;
    return JFEM_OK;
}

extern "C" int jfem_loadset_set_gravity(jfem_ctx *ctx, jfem_loadset *ls, int64_t n_grav, const double *grav_a,
                                        const double *grav_n3, const double *shell_areal) {
    using jfem::fail;
    if (!ls) return fail(ctx, JFEM_ERR_BAD_ARG, "null load set");
    if (ls->device != ctx->device) return fail(ctx, JFEM_ERR_BAD_ARG, "load set lives on another device");
    if (n_grav < 0) return fail(ctx, JFEM_ERR_BAD_ARG, "negative card count");
    const jfem_mesh *m = ls->mesh;
    const int64_t nsh = m->nq + m->nt, nn = m->n_nodes;
    if (n_grav > 0 && (!grav_a || !grav_n3 || (nsh > 0 && !shell_areal)))
        return fail(ctx, JFEM_ERR_BAD_ARG, "missing GRAV arrays");
    for (int64_t i = 0; i < nsh && n_grav > 0; ++i)
        if (!std::isfinite(shell_areal[i])) return fail(ctx, JFEM_ERR_BAD_ARG, "non-finite shell mass");
    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    for (void *p : {(void *)ls->d_grav, (void *)ls->d_areal, (void *)ls->d_gsh, (void *)ls->d_sh_ptr,
                    (void *)ls->d_sh_elem})
        if (p) (void)hipFree(p);
    ls->d_grav = ls->d_areal = ls->d_gsh = nullptr;
    ls->d_sh_ptr = nullptr;
    ls->d_sh_elem = nullptr;
    ls->ng = 0;
    ls->nsh = 0;
    if (n_grav == 0 || nsh == 0) return JFEM_OK;
    std::vector<double> gr(4 * (size_t)n_grav);
    for (int64_t i = 0; i < n_grav; ++i) {
        gr[4 * i] = grav_a[i];
        for (int c = 0; c < 3; ++c) gr[4 * i + 1 + c] = grav_n3[3 * i + c];
    }
    // shells per grid point in element order (CQUAD4s, then CTRIA3s)
    std::vector<int32_t> node_of, elem_of;
    for (int64_t e = 0; e < nsh; ++e) {
        const bool quad = e < m->nq;
        for (int k = 0; k < (quad ? 4 : 3); ++k) {
            node_of.push_back(quad ? m->q[e + m->nq * k] : m->t[(e - m->nq) + m->nt * k]);
            elem_of.push_back((int32_t)e);
        }
    }
    std::vector<int32_t> ord;
    std::vector<int64_t> sh_ptr = node_ptr(nn, node_of, ord);
    std::vector<int32_t> sh_elem(ord.size());
    for (size_t q = 0; q < ord.size(); ++q) sh_elem[q] = elem_of[ord[q]];
    hipStream_t s = ctx->stream;
    hipError_t e = upload_n(&ls->d_grav, gr.data(), gr.size(), s);
    if (e == hipSuccess) e = upload_n(&ls->d_areal, shell_areal, (size_t)nsh, s);
    if (e == hipSuccess) e = hipMalloc((void **)&ls->d_gsh, (size_t)nsh * sizeof(double));
    if (e == hipSuccess) e = upload_n(&ls->d_sh_ptr, sh_ptr.data(), sh_ptr.size(), s);
    if (e == hipSuccess) e = upload_n(&ls->d_sh_elem, sh_elem.data(), sh_elem.size(), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    ls->ng = n_grav;
    ls->nsh = nsh;
    return JFEM_OK;
}

extern "C" int jfem_load_vector_dev(jfem_ctx *ctx, const jfem_loadset *ls, double scale, double *F_dev) {
    using jfem::fail;
    if (!ls || !F_dev) return fail(ctx, JFEM_ERR_BAD_ARG, "null load set or output");
    if (ls->device != ctx->device) return fail(ctx, JFEM_ERR_BAD_ARG, "load set lives on another device");
    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const jfem_mesh *m = ls->mesh;
    if (ls->npl > 0)
        hipLaunchKernelGGL(pload4_kernel, dim3((unsigned)((ls->npl + kLoadBS - 1) / kLoadBS)), dim3(kLoadBS), 0, s,
                           ls->npl, ls->d_pl_elem, ls->d_pl_p, ls->d_pl_dir, scale, m->d_x, m->d_y, m->d_z, m->d_q,
                           m->nq, m->d_t, m->nt, ls->d_plf);
    if (ls->ng > 0 && ls->nsh > 0)
        hipLaunchKernelGGL(grav_share_kernel, dim3((unsigned)((ls->nsh + kLoadBS - 1) / kLoadBS)), dim3(kLoadBS), 0, s,
                           ls->nsh, ls->d_areal, m->d_x, m->d_y, m->d_z, m->d_q, m->nq, m->d_t, m->nt, ls->d_gsh);
    hipLaunchKernelGGL(load_node_kernel, dim3((unsigned)((ls->n_nodes + kLoadBS - 1) / kLoadBS)), dim3(kLoadBS), 0, s,
                       ls->n_nodes, ls->npl > 0 ? ls->d_pl_ptr : nullptr, ls->d_pl_card, ls->d_plf,
                       ls->nf > 0 ? ls->d_f_ptr : nullptr, ls->d_f_rec, ls->nm > 0 ? ls->d_m_ptr : nullptr,
                       ls->d_m_rec, scale, ls->nsh > 0 ? ls->ng : 0, ls->d_grav, ls->d_sh_ptr, ls->d_sh_elem,
                       ls->d_areal, ls->d_gsh, F_dev);
    JFEM_CHECK_HIP(ctx, hipGetLastError());
    return JFEM_OK;
}

extern "C" int jfem_load_vector(jfem_ctx *ctx, const jfem_loadset *ls, double scale, double *F_host) {
    using jfem::fail;
    if (!ls || !F_host) return fail(ctx, JFEM_ERR_BAD_ARG, "null load set or output");
    if (int st = jfem_load_vector_dev(ctx, ls, scale, ls->d_F)) return st;
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(F_host, ls->d_F, 6 * (size_t)ls->n_nodes * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFEM_OK;
}
