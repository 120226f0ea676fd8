// C ABI: contexts, meshes, assembly pattern, element matrices, assembly and the
// boundary-condition partition.  Factor/solve entry points live in factor.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>

#include "jfem_internal.h"

namespace jfem {

int fail(jfem_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

void tl_begin(jfem_ctx *ctx) {
    ctx->tl_n = 0;
    tl_mark(ctx, -1);
}

void tl_mark(jfem_ctx *ctx, int fam) {
    if (ctx->tl_n == ctx->tl_ev.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        ctx->tl_ev.push_back(e);
        ctx->tl_fam.push_back(-1);
    }
    (void)hipEventRecord(ctx->tl_ev[ctx->tl_n], ctx->stream);
    ctx->tl_fam[ctx->tl_n] = fam;
    ++ctx->tl_n;
}

void tl_flush(jfem_ctx *ctx) {
    if (ctx->tl_n == 0) return;
    (void)hipEventSynchronize(ctx->tl_ev[ctx->tl_n - 1]);
    for (size_t i = 1; i < ctx->tl_n; ++i) {
        const int f = ctx->tl_fam[i];
        if (f < 0 || f >= FAM_COUNT) continue;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ctx->tl_ev[i - 1], ctx->tl_ev[i]);
        ctx->ms[f] += ms;
        ctx->launches[f] += 1;
    }
    ctx->tl_n = 0;
}

void timer_start(jfem_ctx *ctx) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
}

void timer_stop(jfem_ctx *ctx, int family, int64_t launches) {
    if (!ctx->timing) return;
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    (void)hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->ms[family] += ms;
    ctx->launches[family] += launches;
}

template <class T>
static hipError_t upload(T **dst, const T *src, size_t n, hipStream_t s) {
    *dst = nullptr;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, n * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}

}  // namespace jfem

using namespace jfem;

extern "C" {

int jfem_ctx_create(int device, jfem_ctx **out) {
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return JFEM_ERR_HIP;
    if (device < 0 || device >= n) return JFEM_ERR_BAD_ARG;
    jfem_ctx *c = new jfem_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess ||
        hipMalloc((void **)&c->d_status, 2 * sizeof(int64_t)) != hipSuccess ||
        hipHostMalloc((void **)&c->h_red, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return JFEM_ERR_HIP;
    }
    *out = c;
    return JFEM_OK;
}

int jfem_ctx_destroy(jfem_ctx *ctx) {
    if (!ctx) return JFEM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->bc_dev) (void)hipFree(ctx->bc_dev);
    if (ctx->bc_host) (void)hipHostFree(ctx->bc_host);
    if (ctx->h_red) (void)hipHostFree(ctx->h_red);
    (void)hipFree(ctx->d_status);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipEventDestroy(ctx->ev2);
    (void)hipEventDestroy(ctx->ev3);
    for (hipEvent_t e : ctx->tl_ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return JFEM_OK;
}

int jfem_ctx_synchronize(jfem_ctx *ctx) {
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFEM_OK;
}

const char *jfem_last_error(jfem_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int jfem_device_memory(jfem_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes) {
    size_t f = 0, t = 0;
    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    JFEM_CHECK_HIP(ctx, hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return JFEM_OK;
}

int jfem_ctx_kernel_time(jfem_ctx *ctx, int family, double *ms, int64_t *launches) {
    if (family < 0 || family >= FAM_COUNT) return JFEM_ERR_BAD_ARG;
    *ms = ctx->ms[family];
    *launches = ctx->launches[family];
    return JFEM_OK;
}

int jfem_ctx_reset_kernel_time(jfem_ctx *ctx) {
    for (int i = 0; i < FAM_COUNT; ++i) { ctx->ms[i] = 0; ctx->launches[i] = 0; }
    return JFEM_OK;
}

int jfem_ctx_enable_timing(jfem_ctx *ctx, int on) {
    ctx->timing = on != 0;
    ctx->timing_level = on;
    return JFEM_OK;
}

int jfem_dev_alloc(jfem_ctx *ctx, int64_t bytes, void **p) {
    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    JFEM_CHECK_HIP(ctx, hipMalloc(p, (size_t)std::max<int64_t>(bytes, 1)));
    return JFEM_OK;
}
int jfem_dev_free(jfem_ctx *ctx, void *p) {
    JFEM_CHECK_HIP(ctx, hipFree(p));
    return JFEM_OK;
}
int jfem_memcpy_h2d(jfem_ctx *ctx, void *d, const void *h, int64_t bytes) {
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(d, h, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFEM_OK;
}
int jfem_memcpy_d2h(jfem_ctx *ctx, void *h, const void *d, int64_t bytes) {
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(h, d, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFEM_OK;
}

// ---------------------------------------------------------------------------- mesh
int jfem_mesh_create(jfem_ctx *ctx, int64_t n_nodes, const double *x, const double *y, const double *z,
                     int64_t nq, const int32_t *qc, const double *qp, int64_t nt, const int32_t *tc,
                     const double *tp, int64_t ntet, const int32_t *ec, const double *ep, double k6rot,
                     jfem_mesh **out) {
    *out = nullptr;
    if (n_nodes <= 0 || nq < 0 || nt < 0 || ntet < 0) return fail(ctx, JFEM_ERR_BAD_ARG, "bad mesh sizes");
    if (n_nodes > (int64_t)INT32_MAX / 6) return fail(ctx, JFEM_ERR_BAD_ARG, "too many grid points");
    auto check_conn = [&](int64_t n, int nn, const int32_t *c) {
        for (int64_t i = 0; i < n * nn; ++i)
            if (c[i] < 0 || c[i] >= n_nodes) return false;
        return true;
    };
    if ((nq && !check_conn(nq, 4, qc)) || (nt && !check_conn(nt, 3, tc)) || (ntet && !check_conn(ntet, 10, ec)))
        return fail(ctx, JFEM_ERR_BAD_ARG, "connectivity references a grid point out of range");
    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    jfem_mesh *m = new jfem_mesh();
    m->device = ctx->device;
    m->n_nodes = n_nodes; m->nq = nq; m->nt = nt; m->ntet = ntet; m->k6rot = k6rot;
    m->xyz.resize(3 * n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) { m->xyz[3 * i] = x[i]; m->xyz[3 * i + 1] = y[i]; m->xyz[3 * i + 2] = z[i]; }
    m->q.assign(qc, qc + 4 * nq);
    m->t.assign(tc, tc + 3 * nt);
    m->tet.assign(ec, ec + 10 * ntet);
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = upload(&m->d_x, x, n_nodes, s);
    if (e == hipSuccess) e = upload(&m->d_y, y, n_nodes, s);
    if (e == hipSuccess) e = upload(&m->d_z, z, n_nodes, s);
    if (e == hipSuccess) e = upload(&m->d_q, qc, 4 * nq, s);
    if (e == hipSuccess) e = upload(&m->d_qp, qp, 8 * nq, s);
    if (e == hipSuccess) e = upload(&m->d_t, tc, 3 * nt, s);
    if (e == hipSuccess) e = upload(&m->d_tp, tp, 8 * nt, s);
    if (e == hipSuccess) e = upload(&m->d_tet, ec, 10 * ntet, s);
    if (e == hipSuccess) e = upload(&m->d_tetp, ep, 2 * ntet, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        jfem_mesh_free(m);
        return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    }
    *out = m;
    return JFEM_OK;
}

int jfem_mesh_update_props(jfem_ctx *ctx, jfem_mesh *m, const double *qp, const double *tp, const double *ep) {
    hipStream_t s = ctx->stream;
    if (qp && m->nq) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(m->d_qp, qp, 8 * m->nq * sizeof(double), hipMemcpyHostToDevice, s));
    if (tp && m->nt) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(m->d_tp, tp, 8 * m->nt * sizeof(double), hipMemcpyHostToDevice, s));
    if (ep && m->ntet) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(m->d_tetp, ep, 2 * m->ntet * sizeof(double), hipMemcpyHostToDevice, s));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return JFEM_OK;
}

int jfem_mesh_free(jfem_mesh *m) {
    if (!m) return JFEM_OK;
    (void)hipSetDevice(m->device);
    for (void *p : {(void *)m->d_x, (void *)m->d_y, (void *)m->d_z, (void *)m->d_q, (void *)m->d_t, (void *)m->d_tet,
                    (void *)m->d_qp, (void *)m->d_tp, (void *)m->d_tetp})
        if (p) (void)hipFree(p);
    delete m;
    return JFEM_OK;
}

// ------------------------------------------------------------------- element output
static int run_elements(jfem_ctx *ctx, const jfem_mesh *m, double *d_Ke, int64_t *bad, int kinds_mask) {
    hipStream_t s = ctx->stream;
    JFEM_CHECK_HIP(ctx, launch_reset_status(s, ctx->d_status));
    timer_start(ctx);
    JFEM_CHECK_HIP(ctx, launch_elements(s, m, m->d_qp, m->d_tp, m->d_tetp, d_Ke, 0, 1, kinds_mask, ctx->d_status));
    timer_stop(ctx, FAM_ELEM, 1);
    int64_t st[2];
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, s));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (st[1] != 0) {
        if (bad) *bad = st[0];
        return fail(ctx, (int)st[1], "element " + std::to_string(st[0]) + " rejected (status " + std::to_string(st[1]) + ")");
    }
    if (bad) *bad = -1;
    return JFEM_OK;
}

int jfem_element_matrices(jfem_ctx *ctx, const jfem_mesh *m, int kind, double *Ke_out, int64_t *bad) {
    int64_t n = kind == JFEM_KIND_CQUAD4 ? m->nq : (kind == JFEM_KIND_CTRIA3 ? m->nt : m->ntet);
    int64_t per = kind == JFEM_KIND_CQUAD4 ? 576 : (kind == JFEM_KIND_CTRIA3 ? 324 : 900);
    if (kind < 0 || kind > 2) return fail(ctx, JFEM_ERR_BAD_ARG, "bad kind");
    if (bad) *bad = -1;
    if (n == 0) return JFEM_OK;
    double *d = nullptr;
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d, n * per * sizeof(double)));
    int st = run_elements(ctx, m, d, bad, 1 << kind);
    if (st == JFEM_OK) {
        hipError_t e = hipMemcpyAsync(Ke_out, d, n * per * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(d);
    return st;
}

int jfem_q4_local_batch(jfem_ctx *ctx, int64_t n, const double *lc, const double *Cm, const double *Cb,
                        const double *Cs, const double *h, const double *k6, const double *ds,
                        const int32_t *flags, double *Ke_out, int64_t *bad) {
    if (bad) *bad = -1;
    if (n <= 0) return JFEM_OK;
    hipStream_t s = ctx->stream;
    double *d_lc, *d_Cm, *d_Cb, *d_Cs, *d_h, *d_k6, *d_ds, *d_Ke;
    int32_t *d_f;
    JFEM_CHECK_HIP(ctx, upload(&d_lc, lc, 8 * n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_Cm, Cm, 9 * n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_Cb, Cb, 9 * n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_Cs, Cs, 4 * n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_h, h, n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_k6, k6, n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_ds, ds, n, s));
    JFEM_CHECK_HIP(ctx, upload(&d_f, flags, n, s));
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_Ke, 576 * n * sizeof(double)));
    JFEM_CHECK_HIP(ctx, launch_reset_status(s, ctx->d_status));
    JFEM_CHECK_HIP(ctx, launch_q4_local(s, n, d_lc, d_Cm, d_Cb, d_Cs, d_h, d_k6, d_ds, d_f, d_Ke, ctx->d_status));
    int64_t st[2];
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, s));
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(Ke_out, d_Ke, 576 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    for (void *p : {(void *)d_lc, (void *)d_Cm, (void *)d_Cb, (void *)d_Cs, (void *)d_h, (void *)d_k6, (void *)d_ds,
                    (void *)d_f, (void *)d_Ke})
        (void)hipFree(p);
    if (st[1] != 0) {
        if (bad) *bad = st[0];
        return fail(ctx, (int)st[1], "element " + std::to_string(st[0]) + " rejected");
    }
    return JFEM_OK;
}

// element kernels on caller-given per-element inputs: upload, launch, download,
// status check (the inputs are n x per_k doubles each)
static int run_direct(jfem_ctx *ctx, int64_t n, int nd, std::initializer_list<std::pair<const double *, int>> in,
                      double *Ke_out, int64_t *bad,
                      const std::function<hipError_t(const std::vector<double *> &, double *)> &launch) {
    if (bad) *bad = -1;
    if (n <= 0) return JFEM_OK;
    hipStream_t s = ctx->stream;
    std::vector<double *> d;
    double *d_Ke = nullptr;
    auto release = [&]() {
        for (double *p : d) (void)hipFree(p);
        if (d_Ke) (void)hipFree(d_Ke);
    };
    hipError_t e = hipSuccess;
    for (const auto &a : in) {
        double *p = nullptr;
        if (e == hipSuccess) e = upload(&p, a.first, (size_t)a.second * n, s);
        d.push_back(p);
    }
    if (e == hipSuccess) e = hipMalloc((void **)&d_Ke, (size_t)nd * nd * n * sizeof(double));
    if (e == hipSuccess) e = launch_reset_status(s, ctx->d_status);
    if (e == hipSuccess) e = launch(d, d_Ke);
    int64_t st[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(Ke_out, d_Ke, (size_t)nd * nd * n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    release();
    if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    if (st[1] != 0) {
        if (bad) *bad = st[0];
        return fail(ctx, (int)st[1], "element " + std::to_string(st[0]) + " rejected");
    }
    return JFEM_OK;
}

int jfem_t3_local_batch(jfem_ctx *ctx, int64_t n, const double *lc, const double *Dm, const double *Db,
                        const double *Ds, const double *h, const double *G_ref, const double *bend_ratio,
                        const double *k6rot, double *Ke_out, int64_t *bad) {
    (void)h; (void)G_ref; (void)bend_ratio;      // not read by the default branch
    return run_direct(ctx, n, 18, {{lc, 6}, {Dm, 9}, {Db, 9}, {Ds, 4}, {k6rot, 1}}, Ke_out, bad,
                      [&](const std::vector<double *> &d, double *Ke) {
                          return launch_t3_local(ctx->stream, n, d[0], d[1], d[2], d[3], d[4], Ke, ctx->d_status);
                      });
}

int jfem_tet10_batch(jfem_ctx *ctx, int64_t n, const double *X, const double *E, const double *nu, double *Ke_out,
                     int64_t *bad) {
    if (bad) *bad = -1;
    if (n <= 0) return JFEM_OK;
    std::vector<double> prop(2 * (size_t)n);
    for (int64_t e = 0; e < n; ++e) { prop[e] = E[e]; prop[n + e] = nu[e]; }
    return run_direct(ctx, n, 30, {{X, 30}, {prop.data(), 2}}, Ke_out, bad,
                      [&](const std::vector<double *> &d, double *Ke) {
                          return launch_tet10_direct(ctx->stream, n, d[0], d[1], Ke, ctx->d_status);
                      });
}

// ------------------------------------------------------------------------- shell recovery
int jfem_shell_recover(jfem_ctx *ctx, const jfem_mesh *m, const double *u, int mean_shear, double *q4_out,
                       double *t3_out, int64_t *bad) {
    if (bad) *bad = -1;
    if (!ctx || !m || !u) return fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if ((m->nq > 0 && !q4_out) || (m->nt > 0 && !t3_out)) return fail(ctx, JFEM_ERR_BAD_ARG, "null output");
    if (m->nq + m->nt == 0) return JFEM_OK;
    hipStream_t s = ctx->stream;
    const int64_t ndof = 6 * m->n_nodes;
    double *d_u = nullptr, *d_q = nullptr, *d_t = nullptr;
    hipError_t e = upload(&d_u, u, (size_t)ndof, s);
    if (e == hipSuccess && m->nq) e = hipMalloc((void **)&d_q, (size_t)m->nq * 44 * sizeof(double));
    if (e == hipSuccess && m->nt) e = hipMalloc((void **)&d_t, (size_t)m->nt * 20 * sizeof(double));
    if (e == hipSuccess) e = launch_reset_status(s, ctx->d_status);
    if (e == hipSuccess) e = launch_shell_recover(s, m, d_u, mean_shear, d_q, d_t, ctx->d_status);
    int64_t st[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && m->nq)
        e = hipMemcpyAsync(q4_out, d_q, (size_t)m->nq * 44 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && m->nt)
        e = hipMemcpyAsync(t3_out, d_t, (size_t)m->nt * 20 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (double *p : {d_u, d_q, d_t})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    if (st[1] != 0) {
        if (bad) *bad = st[0];
        return fail(ctx, (int)st[1], "element " + std::to_string(st[0]) + " rejected");
    }
    return JFEM_OK;
}

// ------------------------------------------------------------------ geometric stiffness
namespace {
// Kg of every shell into Ke (device, nq x 576 then nt x 324) for host displacements u
int run_kg(jfem_ctx *ctx, const jfem_mesh *m, const double *u, double *Ke, int64_t *bad) {
    if (bad) *bad = -1;
    if (m->ntet) return fail(ctx, JFEM_ERR_BAD_ARG, "geometric stiffness covers shell meshes only");
    hipStream_t s = ctx->stream;
    double *d_u = nullptr;
    hipError_t e = upload(&d_u, u, (size_t)(6 * m->n_nodes), s);
    if (e == hipSuccess) e = launch_reset_status(s, ctx->d_status);
    if (e == hipSuccess) e = launch_shell_kg(s, m, d_u, Ke, ctx->d_status);
    int64_t st[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_u) (void)hipFree(d_u);
    if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    if (st[1] != 0) {
        if (bad) *bad = st[0];
        return fail(ctx, (int)st[1], "element " + std::to_string(st[0]) + " rejected");
    }
    return JFEM_OK;
}
}  // namespace

int jfem_shell_kg_batch(jfem_ctx *ctx, const jfem_mesh *m, const double *u, double *kg_out, int64_t *bad) {
    if (bad) *bad = -1;
    const size_t count = (size_t)m->nq * 576 + (size_t)m->nt * 324;
    if (!ctx || !m || !u || (count && !kg_out)) return fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (m->ntet) return fail(ctx, JFEM_ERR_BAD_ARG, "geometric stiffness covers shell meshes only");
    if (count == 0) return JFEM_OK;
    double *d_kg = nullptr;
    if (hipMalloc((void **)&d_kg, count * sizeof(double)) != hipSuccess)
        return fail(ctx, JFEM_ERR_OOM, "element matrices");
    int st = run_kg(ctx, m, u, d_kg, bad);
    if (st == JFEM_OK) {
        hipError_t e = hipMemcpyAsync(kg_out, d_kg, count * sizeof(double), hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    }
    (void)hipFree(d_kg);
    return st;
}

int jfem_assemble_kg(jfem_ctx *ctx, const jfem_mesh *m, const jfem_pattern *P, const double *u, jfem_matrix **out,
                     int64_t *bad) {
    if (bad) *bad = -1;
    if (!out) return fail(ctx, JFEM_ERR_BAD_ARG, "null output");
    *out = nullptr;
    if (!ctx || !m || !P || !u) return fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (P->nq != m->nq || P->nt != m->nt || P->ntet != m->ntet) return fail(ctx, JFEM_ERR_BAD_ARG, "mesh/pattern mismatch");
    jfem_matrix *G = new jfem_matrix();
    G->device = ctx->device;
    G->P = P;
    hipError_t e = hipMalloc((void **)&G->d_val, std::max<int64_t>(P->nnz, 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&G->d_Ke, std::max<int64_t>(P->ntrip, 1) * sizeof(double));
    if (e != hipSuccess) {
        jfem_matrix_free(G);
        return fail(ctx, JFEM_ERR_OOM, "geometric stiffness buffers");
    }
    int st = run_kg(ctx, m, u, G->d_Ke, bad);
    if (st == JFEM_OK) {
        hipError_t e2 = launch_assemble(ctx->stream, P, G->d_Ke, G->d_val, 1, 0, 0);
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(ctx->stream);
        if (e2 != hipSuccess) st = fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e2));
    }
    if (st != JFEM_OK) {
        jfem_matrix_free(G);
        return st;
    }
    *out = G;
    return JFEM_OK;
}

int jfem_matrix_shift(jfem_ctx *ctx, const jfem_matrix *K, double sigma, const jfem_matrix *G, jfem_matrix **out) {
    if (!ctx || !K || !G || !out) return fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (K->P != G->P || (*out && (*out)->P != K->P)) return fail(ctx, JFEM_ERR_BAD_ARG, "matrices of different patterns");
    if (!std::isfinite(sigma)) return fail(ctx, JFEM_ERR_BAD_ARG, "shift is not finite");
    jfem_matrix *A = *out;
    if (!A) {
        A = new jfem_matrix();
        A->device = ctx->device;
        A->P = K->P;
        if (hipMalloc((void **)&A->d_val, std::max<int64_t>(K->P->nnz, 1) * sizeof(double)) != hipSuccess) {
            jfem_matrix_free(A);
            return fail(ctx, JFEM_ERR_OOM, "shifted matrix");
        }
    }
    hipError_t e = launch_csr_shift(ctx->stream, K->P->nnz, K->d_val, sigma, G->d_val, A->d_val);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        if (!*out) jfem_matrix_free(A);
        return fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    }
    *out = A;
    return JFEM_OK;
}

int jfem_matrix_shift_scaled(jfem_ctx *ctx, const jfem_matrix *K, double sigma, const jfem_matrix *G,
                             jfem_matrix **out) {
    if (!ctx || !K || !G || !out) return fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (K->P != G->P || (*out && (*out)->P != K->P)) return fail(ctx, JFEM_ERR_BAD_ARG, "matrices of different patterns");
    if (!std::isfinite(sigma)) return fail(ctx, JFEM_ERR_BAD_ARG, "shift is not finite");
    const jfem_pattern *P = K->P;
    jfem_matrix *A = *out;
    if (!A) {
        A = new jfem_matrix();
        A->device = ctx->device;
        A->P = P;
        if (hipMalloc((void **)&A->d_val, std::max<int64_t>(P->nnz, 1) * sizeof(double)) != hipSuccess) {
            jfem_matrix_free(A);
            return fail(ctx, JFEM_ERR_OOM, "shifted matrix");
        }
    }
    int32_t *d_dp = nullptr;
    double *d_sd = nullptr;
    hipError_t e = upload(&d_dp, P->diagpos.data(), P->diagpos.size(), ctx->stream);
    if (e == hipSuccess) e = hipMalloc((void **)&d_sd, std::max<int64_t>(P->ndof, 1) * sizeof(double));
    if (e == hipSuccess) e = launch_csr_shift_scaled(ctx->stream, P, d_dp, d_sd, K->d_val, sigma, G->d_val, A->d_val);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_dp) (void)hipFree(d_dp);
    if (d_sd) (void)hipFree(d_sd);
    if (e != hipSuccess) {
        if (!*out) jfem_matrix_free(A);
        return fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    }
    *out = A;
    return JFEM_OK;
}

// ------------------------------------------------------------------------- pattern
int jfem_pattern_create(jfem_ctx *ctx, const jfem_mesh *m, jfem_pattern **out, int64_t *nnz_out) {
    *out = nullptr;
    const int64_t nn = m->n_nodes, ndof = 6 * nn;
    const int64_t ntrip = m->nq * 576 + m->nt * 324 + m->ntet * 900;
    if (ntrip >= (int64_t)UINT32_MAX) return fail(ctx, JFEM_ERR_BAD_ARG, "more than 2^32 element entries");
    // node pairs with coupling type: 2 = all six DOFs (shell), 1 = translations only (solid)
    std::vector<uint64_t> key;
    key.reserve((size_t)(m->nq * 16 + m->nt * 9 + m->ntet * 100));
    auto add_elem = [&](const std::vector<int32_t> &conn, int64_t n, int k, int64_t e, int type) {
        for (int a = 0; a < k; ++a)
            for (int b = 0; b < k; ++b) {
                uint64_t na = (uint64_t)conn[a * n + e], nb = (uint64_t)conn[b * n + e];
                key.push_back((na << 33) | (nb << 2) | (uint64_t)type);
            }
    };
    for (int64_t e = 0; e < m->nq; ++e) add_elem(m->q, m->nq, 4, e, 2);
    for (int64_t e = 0; e < m->nt; ++e) add_elem(m->t, m->nt, 3, e, 2);
    for (int64_t e = 0; e < m->ntet; ++e) add_elem(m->tet, m->ntet, 10, e, 1);
    std::sort(key.begin(), key.end());
    // keep max type per (a,b)
    std::vector<int64_t> nptr(nn + 1, 0);
    std::vector<int32_t> nadj;
    std::vector<uint8_t> ntype;
    nadj.reserve(key.size() / 4);
    for (size_t i = 0; i < key.size(); ++i) {
        uint64_t a = key[i] >> 33, b = (key[i] >> 2) & 0x7FFFFFFFull;
        uint8_t t = (uint8_t)(key[i] & 3);
        if (i + 1 < key.size() && (key[i + 1] >> 2) == (key[i] >> 2)) continue;  // later entry has larger type
        nadj.push_back((int32_t)b);
        ntype.push_back(t);
        nptr[a + 1]++;
    }
    key.clear();
    key.shrink_to_fit();
    for (int64_t a = 0; a < nn; ++a) nptr[a + 1] += nptr[a];
    jfem_pattern *P = new jfem_pattern();
    P->device = ctx->device;
    P->ndof = ndof; P->n_nodes = nn; P->ntrip = ntrip; P->nq = m->nq; P->nt = m->nt; P->ntet = m->ntet;
    P->rowptr.assign(ndof + 1, 0);
    for (int64_t a = 0; a < nn; ++a)
        for (int i = 0; i < 6; ++i) {
            int64_t cnt = 0;
            for (int64_t p = nptr[a]; p < nptr[a + 1]; ++p) cnt += ntype[p] == 2 ? 6 : (i < 3 ? 3 : 0);
            P->rowptr[6 * a + i + 1] = cnt;
        }
    for (int64_t r = 0; r < ndof; ++r) P->rowptr[r + 1] += P->rowptr[r];
    P->nnz = P->rowptr[ndof];
    if (P->nnz >= (int64_t)INT32_MAX) { delete P; return fail(ctx, JFEM_ERR_BAD_ARG, "nnz exceeds 2^31"); }
    P->col.resize(P->nnz);
    for (int64_t a = 0; a < nn; ++a)
        for (int i = 0; i < 6; ++i) {
            int64_t w = P->rowptr[6 * a + i];
            for (int64_t p = nptr[a]; p < nptr[a + 1]; ++p) {
                int jm = ntype[p] == 2 ? 6 : (i < 3 ? 3 : 0);
                for (int j = 0; j < jm; ++j) P->col[w++] = 6 * nadj[p] + j;
            }
        }
    auto find = [&](int64_t r, int64_t c) -> int64_t {
        const int32_t *b = P->col.data() + P->rowptr[r], *e = P->col.data() + P->rowptr[r + 1];
        const int32_t *it = std::lower_bound(b, e, (int32_t)c);
        return (it != e && *it == c) ? (int64_t)(it - P->col.data()) : -1;
    };
    // The columns of row 6a+i are the node blocks of a's neighbours in order, each
    // 6 wide (shell coupling) or 3 wide (solid coupling, translation rows only):
    // the entry (6a+i, 6b+j) sits at rowptr[6a+i] + (width of the blocks before
    // b's) + j.  Those prefix widths per adjacency entry, for translation and for
    // rotation rows:
    std::vector<int32_t> off_t(nadj.size()), off_r(nadj.size());
    for (int64_t a = 0; a < nn; ++a) {
        int32_t wt = 0, wr = 0;
        for (int64_t p = nptr[a]; p < nptr[a + 1]; ++p) {
            off_t[p] = wt; off_r[p] = wr;
            wt += ntype[p] == 2 ? 6 : 3;
            wr += ntype[p] == 2 ? 6 : 0;
        }
    }
    auto adj_pos = [&](int64_t a, int32_t b) -> int64_t {   // b's entry in a's sorted adjacency
        const int32_t *lo = nadj.data() + nptr[a], *hi = nadj.data() + nptr[a + 1];
        return (int64_t)(std::lower_bound(lo, hi, b) - nadj.data());
    };
    // slot of every triplet, in triplet order: one adjacency search per node pair
    // of an element, then the DOF offsets within the node blocks
    std::vector<int32_t> slot(ntrip);
    int64_t t = 0;
    auto emit = [&](const std::vector<int32_t> &conn, int64_t n, int k, int dpn, int64_t e) {
        int64_t pos[10][10];
        for (int bb = 0; bb < k; ++bb)
            for (int aa = 0; aa < k; ++aa) pos[aa][bb] = adj_pos(conn[aa * n + e], conn[bb * n + e]);
        const int ndl = k * dpn;
        for (int cc = 0; cc < ndl; ++cc)
            for (int rr = 0; rr < ndl; ++rr) {
                const int aa = rr / dpn, i = rr % dpn, bb = cc / dpn, j = cc % dpn;
                const int64_t p = pos[aa][bb];
                const int64_t r = 6 * (int64_t)conn[aa * n + e] + i;
                slot[t++] = (int32_t)(P->rowptr[r] + (i < 3 ? off_t[p] : off_r[p]) + j);
            }
    };
    for (int64_t e = 0; e < m->nq; ++e) emit(m->q, m->nq, 4, 6, e);
    for (int64_t e = 0; e < m->nt; ++e) emit(m->t, m->nt, 3, 6, e);
    for (int64_t e = 0; e < m->ntet; ++e) emit(m->tet, m->ntet, 10, 3, e);
    std::vector<int64_t> seg_ptr(P->nnz + 1, 0);
    for (int64_t i = 0; i < ntrip; ++i) seg_ptr[slot[i] + 1]++;
    for (int64_t s = 0; s < P->nnz; ++s) seg_ptr[s + 1] += seg_ptr[s];
    std::vector<int32_t> seg_src(ntrip);
    {
        std::vector<int64_t> fill(seg_ptr.begin(), seg_ptr.end() - 1);
        for (int64_t i = 0; i < ntrip; ++i) seg_src[fill[slot[i]]++] = (int32_t)(uint32_t)i;
    }
    P->nodeblk.assign(nn * 36, -1);
    P->diagpos.assign(ndof, -1);
    for (int64_t a = 0; a < nn; ++a)
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) P->nodeblk[36 * a + 6 * i + j] = find(6 * a + i, 6 * a + j);
    for (int64_t r = 0; r < ndof; ++r) P->diagpos[r] = (int32_t)find(r, r);
    hipStream_t s = ctx->stream;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = upload(&P->d_rowptr, P->rowptr.data(), ndof + 1, s);
    if (e == hipSuccess) e = upload(&P->d_col, P->col.data(), P->nnz, s);
    if (e == hipSuccess) e = upload(&P->d_seg_ptr, seg_ptr.data(), P->nnz + 1, s);
    if (e == hipSuccess) e = upload(&P->d_seg_src, seg_src.data(), ntrip, s);
    if (e == hipSuccess) e = upload(&P->d_nodeblk, P->nodeblk.data(), nn * 36, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        jfem_pattern_free(P);
        return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    }
    if (nnz_out) *nnz_out = P->nnz;
    *out = P;
    return JFEM_OK;
}

int jfem_pattern_free(jfem_pattern *P) {
    if (!P) return JFEM_OK;
    (void)hipSetDevice(P->device);
    for (void *p : {(void *)P->d_rowptr, (void *)P->d_col, (void *)P->d_seg_ptr, (void *)P->d_seg_src, (void *)P->d_nodeblk})
        if (p) (void)hipFree(p);
    delete P;
    return JFEM_OK;
}

// ------------------------------------------------------------------------ assembly
int jfem_reassemble(jfem_ctx *ctx, const jfem_mesh *m, jfem_matrix *K, int64_t *bad) {
    const jfem_pattern *P = K->P;
    if (P->nq != m->nq || P->nt != m->nt || P->ntet != m->ntet) return fail(ctx, JFEM_ERR_BAD_ARG, "mesh/pattern mismatch");
    int st = run_elements(ctx, m, K->d_Ke, bad, 7);
    if (st != JFEM_OK) return st;
    timer_start(ctx);
    JFEM_CHECK_HIP(ctx, launch_assemble(ctx->stream, P, K->d_Ke, K->d_val, 1, 0, 0));
    timer_stop(ctx, FAM_ASSEMBLY, 1);
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JFEM_OK;
}

int jfem_assemble(jfem_ctx *ctx, const jfem_mesh *m, const jfem_pattern *P, jfem_matrix **out, int64_t *bad) {
    *out = nullptr;
    jfem_matrix *K = new jfem_matrix();
    K->device = ctx->device;
    K->P = P;
    hipError_t e = hipMalloc((void **)&K->d_val, std::max<int64_t>(P->nnz, 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&K->d_Ke, std::max<int64_t>(P->ntrip, 1) * sizeof(double));
    if (e != hipSuccess) {
        jfem_matrix_free(K);
        return fail(ctx, JFEM_ERR_OOM, "assembly buffers");
    }
    int st = jfem_reassemble(ctx, m, K, bad);
    if (st != JFEM_OK) {
        jfem_matrix_free(K);
        return st;
    }
    *out = K;
    return JFEM_OK;
}

int jfem_matrix_download(jfem_ctx *ctx, const jfem_matrix *K, int64_t *row_ptr, int32_t *col, double *val) {
    const jfem_pattern *P = K->P;
    if (row_ptr) std::memcpy(row_ptr, P->rowptr.data(), (P->ndof + 1) * sizeof(int64_t));
    if (col) std::memcpy(col, P->col.data(), P->nnz * sizeof(int32_t));
    if (val) {
        JFEM_CHECK_HIP(ctx, hipMemcpyAsync(val, K->d_val, P->nnz * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return JFEM_OK;
}

int jfem_matrix_free(jfem_matrix *K) {
    if (!K || K->view) return JFEM_OK;     // variants are freed with their batch
    (void)hipSetDevice(K->device);
    if (K->d_val) (void)hipFree(K->d_val);
    if (K->d_Ke) (void)hipFree(K->d_Ke);
    delete K;
    return JFEM_OK;
}

int64_t jfem_matrix_device_bytes(const jfem_matrix *K) {
    return K ? (K->P->nnz + K->P->ntrip) * (int64_t)sizeof(double) : 0;
}

int64_t jfem_mesh_device_bytes(const jfem_mesh *m) {
    if (!m) return 0;
    return 3 * m->n_nodes * (int64_t)sizeof(double) + (4 * m->nq + 3 * m->nt + 10 * m->ntet) * (int64_t)sizeof(int32_t) +
           (8 * m->nq + 8 * m->nt + 2 * m->ntet) * (int64_t)sizeof(double);
}

int64_t jfem_pattern_device_bytes(const jfem_pattern *P) {
    if (!P) return 0;
    return (P->ndof + 1) * (int64_t)sizeof(int64_t) + P->nnz * (int64_t)sizeof(int32_t) +
           (P->nnz + 1) * (int64_t)sizeof(int64_t) + P->ntrip * (int64_t)sizeof(int32_t) +
           36 * P->n_nodes * (int64_t)sizeof(int64_t);
}

int64_t jfem_ctx_device_bytes(const jfem_ctx *ctx) {
    if (!ctx) return 0;
    return (int64_t)(ctx->scratch_bytes + ctx->bc_dev_bytes) + 2 * (int64_t)sizeof(int64_t);
}

// ------------------------------------------------------------------ design variants
// V section sets of one topology, generated and assembled in one launch sequence:
// one element launch per kind and one assembly launch cover every variant
int jfem_batch_create(jfem_ctx *ctx, const jfem_mesh *m, const jfem_pattern *P, int64_t nvar,
                      const double *quad_props, const double *tria_props, const double *tet_props, jfem_batch **out) {
    *out = nullptr;
    if (nvar < 1 || nvar > 65535) return fail(ctx, JFEM_ERR_BAD_ARG, "variant count must be in 1..65535");
    if (P->nq != m->nq || P->nt != m->nt || P->ntet != m->ntet) return fail(ctx, JFEM_ERR_BAD_ARG, "mesh/pattern mismatch");
    JFEM_CHECK_HIP(ctx, hipSetDevice(ctx->device));
    jfem_batch *B = new jfem_batch();
    B->device = ctx->device;
    B->mesh = m;
    B->P = P;
    B->V = nvar;
    B->Ke_stride = jfem_stride(std::max<int64_t>(P->ntrip, 1), 8);
    B->val_stride = jfem_stride(std::max<int64_t>(P->nnz, 1), 8);
    hipStream_t s = ctx->stream;
    hipError_t e = upload(&B->d_qp, quad_props, (size_t)(8 * m->nq * nvar), s);
    if (e == hipSuccess) e = upload(&B->d_tp, tria_props, (size_t)(8 * m->nt * nvar), s);
    if (e == hipSuccess) e = upload(&B->d_tetp, tet_props, (size_t)(2 * m->ntet * nvar), s);
    if (e == hipSuccess) e = hipMalloc((void **)&B->d_Ke, (size_t)(B->Ke_stride * nvar) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&B->d_val, (size_t)(B->val_stride * nvar) * sizeof(double));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        jfem_batch_free(B);
        return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    }
    for (int64_t v = 0; v < nvar; ++v) {
        jfem_matrix *K = new jfem_matrix();
        K->device = ctx->device;
        K->P = P;
        K->d_val = B->d_val + v * B->val_stride;
        K->d_Ke = B->d_Ke + v * B->Ke_stride;
        K->view = true;
        B->mviews.push_back(K);
    }
    *out = B;
    return JFEM_OK;
}

int jfem_batch_update_props(jfem_ctx *ctx, jfem_batch *B, const double *quad_props, const double *tria_props,
                            const double *tet_props) {
    const jfem_mesh *m = B->mesh;
    hipStream_t s = ctx->stream;
    if (m->nq) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(B->d_qp, quad_props, 8 * m->nq * B->V * sizeof(double), hipMemcpyHostToDevice, s));
    if (m->nt) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(B->d_tp, tria_props, 8 * m->nt * B->V * sizeof(double), hipMemcpyHostToDevice, s));
    if (m->ntet) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(B->d_tetp, tet_props, 2 * m->ntet * B->V * sizeof(double), hipMemcpyHostToDevice, s));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    return JFEM_OK;
}

int jfem_batch_assemble(jfem_ctx *ctx, jfem_batch *B, int64_t *bad_variant, int64_t *bad_element) {
    if (bad_variant) *bad_variant = -1;
    if (bad_element) *bad_element = -1;
    const jfem_mesh *m = B->mesh;
    hipStream_t s = ctx->stream;
    JFEM_CHECK_HIP(ctx, launch_reset_status(s, ctx->d_status));
    timer_start(ctx);
    JFEM_CHECK_HIP(ctx, launch_elements(s, m, B->d_qp, B->d_tp, B->d_tetp, B->d_Ke, B->Ke_stride, (int)B->V, 7,
                                        ctx->d_status));
    timer_stop(ctx, FAM_ELEM, 1);
    timer_start(ctx);
    JFEM_CHECK_HIP(ctx, launch_assemble(s, B->P, B->d_Ke, B->d_val, (int)B->V, B->Ke_stride, B->val_stride));
    timer_stop(ctx, FAM_ASSEMBLY, 1);
    int64_t st[2];
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, s));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(s));
    if (st[1] != 0) {
        const int64_t total = std::max<int64_t>(1, m->nq + m->nt + m->ntet);
        if (bad_variant) *bad_variant = st[0] / total;
        if (bad_element) *bad_element = st[0] % total;
        return fail(ctx, (int)st[1], "variant " + std::to_string(st[0] / total) + ": element " +
                                         std::to_string(st[0] % total) + " rejected");
    }
    return JFEM_OK;
}

int jfem_batch_matrix(const jfem_batch *B, int64_t v, jfem_matrix **view) {
    *view = nullptr;
    if (!B || v < 0 || v >= B->V) return JFEM_ERR_BAD_ARG;
    *view = B->mviews[(size_t)v];
    return JFEM_OK;
}

int jfem_batch_free(jfem_batch *B) {
    if (!B) return JFEM_OK;
    (void)hipSetDevice(B->device);
    batch_free_factor(B);
    for (jfem_matrix *K : B->mviews) delete K;
    for (void *p : {(void *)B->d_qp, (void *)B->d_tp, (void *)B->d_tetp, (void *)B->d_Ke, (void *)B->d_val})
        if (p) (void)hipFree(p);
    delete B;
    return JFEM_OK;
}

// ------------------------------------------------------------------ BC partition
// src/solver/boundary_conditions.jl:1480-1607 (SPC), 1028-1138 (AUTOSPC with grid
// point processing on: only exact-zero / negative diagonals), 745-845 (GPST).
}  // extern "C"

namespace {

hipError_t grow_bc_dev(jfem_ctx *ctx, size_t bytes) {
    if (ctx->bc_dev_bytes >= bytes) return hipSuccess;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    if (ctx->bc_dev) (void)hipFree(ctx->bc_dev);
    ctx->bc_dev = nullptr;
    ctx->bc_dev_bytes = 0;
    e = hipMalloc(&ctx->bc_dev, bytes);
    if (e == hipSuccess) ctx->bc_dev_bytes = bytes;
    return e;
}

hipError_t grow_bc_host(jfem_ctx *ctx, size_t bytes) {
    if (ctx->bc_host_bytes >= bytes) return hipSuccess;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    if (ctx->bc_host) (void)hipHostFree(ctx->bc_host);
    ctx->bc_host = nullptr;
    ctx->bc_host_bytes = 0;
    e = hipHostMalloc(&ctx->bc_host, bytes, hipHostMallocDefault);
    if (e == hipSuccess) ctx->bc_host_bytes = bytes;
    return e;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

// SPC + AUTOSPC (zero / negative diagonal) + grid-point singularity, with the
// decision rules of src/solver/boundary_conditions.jl:1480-1607.  The device
// applies the diagonal rule to every DOF and screens every grid point; the
// eigen test (and closest-DOF claim) runs on the host for the flagged grid
// points only, from their 6 x 6 blocks.
int jfem_bc_partition(jfem_ctx *ctx, const jfem_matrix *K, const uint8_t *spc, double eps_zero, uint8_t *fixed,
                      int64_t *n_free) {
    const jfem_pattern *P = K->P;
    const int64_t nn = P->n_nodes, ndof = P->ndof;
    if (n_free) *n_free = 0;
    if (ndof == 0) return JFEM_OK;
    const size_t o_spc = 0, o_fix = align256(ndof), o_flag = o_fix + align256(ndof);
    JFEM_CHECK_HIP(ctx, grow_bc_dev(ctx, o_flag + align256(nn)));
    JFEM_CHECK_HIP(ctx, grow_bc_host(ctx, align256(ndof)));
    char *dv = (char *)ctx->bc_dev;
    uint8_t *hfix = (uint8_t *)ctx->bc_host;
    if (spc) JFEM_CHECK_HIP(ctx, hipMemcpyAsync(dv + o_spc, spc, ndof, hipMemcpyHostToDevice, ctx->stream));
    else JFEM_CHECK_HIP(ctx, hipMemsetAsync(dv + o_spc, 0, ndof, ctx->stream));
    const double thresh = 1e3 * eps_zero;
    JFEM_CHECK_HIP(ctx, launch_bc_screen(ctx->stream, P, K->d_val, (const uint8_t *)(dv + o_spc), thresh,
                                         (uint8_t *)(dv + o_fix), (uint8_t *)(dv + o_flag)));
    // grid-point singularity of the flagged groups, decided on the device
    JFEM_CHECK_HIP(ctx, launch_bc_decide(ctx->stream, P, K->d_val, (const uint8_t *)(dv + o_flag), eps_zero,
                                         (uint8_t *)(dv + o_fix)));
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(hfix, dv + o_fix, ndof, hipMemcpyDeviceToHost, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(fixed, hfix, ndof);
    int64_t cnt = 0;
    for (int64_t i = 0; i < ndof; ++i) cnt += fixed[i] ? 0 : 1;
    if (n_free) *n_free = cnt;
    return JFEM_OK;
}

}  // extern "C"
