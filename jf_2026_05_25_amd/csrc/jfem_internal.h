// Internal types shared by the HIP translation units of libjfem_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/jfem_hip.h"

#define JFEM_CHECK_HIP(ctx, call)                                                   \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            return jfem::fail(ctx, e_ == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, \
                              std::string(#call) + ": " + hipGetErrorString(e_));   \
        }                                                                           \
    } while (0)

namespace jfem {

enum Family { FAM_ELEM = 0, FAM_ASSEMBLY = 1, FAM_FACTOR = 2, FAM_SOLVE = 3, FAM_SPMV = 4, FAM_UPDATE = 5,
              FAM_PANEL = 6, FAM_UNUSED7 = 7, FAM_FLOW_FWD = 8, FAM_FLOW_BWD = 9, FAM_SCHUR = 10,
              FAM_INV = 11, FAM_COUNT = 12 };

}  // namespace jfem

struct jfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    int timing_level = 0;          // 2 = per-launch timing inside the factorization
    hipEvent_t ev2 = nullptr, ev3 = nullptr;
    double ms[jfem::FAM_COUNT] = {0};
    int64_t launches[jfem::FAM_COUNT] = {0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // timing level 2: a timeline of events, one recorded after each kernel of a
    // launch sequence and resolved once the sequence has been issued, so that no
    // host wait (and no launch latency) falls inside a measured interval
    std::vector<hipEvent_t> tl_ev;
    std::vector<int> tl_fam;
    size_t tl_n = 0;
    // pinned host array (16 doubles) the refinement's norm reductions write into
    double *h_red = nullptr;
    // device status word shared by kernels that report element / pivot failures
    int64_t *d_status = nullptr;   // [0] = first bad index (INT64_MAX if none), [1] = code
    // scratch for solves
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // boundary-condition screening buffers (device + pinned host staging)
    void *bc_dev = nullptr;
    size_t bc_dev_bytes = 0;
    void *bc_host = nullptr;
    size_t bc_host_bytes = 0;
};

struct jfem_mesh {
    int device = 0;
    int64_t n_nodes = 0, nq = 0, nt = 0, ntet = 0;
    double k6rot = 100.0;
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    int32_t *d_q = nullptr, *d_t = nullptr, *d_tet = nullptr;   // column-major conn
    double *d_qp = nullptr, *d_tp = nullptr, *d_tetp = nullptr; // column-major props
    // host copies needed by the symbolic phases
    std::vector<double> xyz;        // n_nodes x 3, row-major (x,y,z per node)
    std::vector<int32_t> q, t, tet; // column-major conn
};

struct jfem_pattern {
    int device = 0;
    int64_t ndof = 0, nnz = 0, ntrip = 0, n_nodes = 0;
    int64_t nq = 0, nt = 0, ntet = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    int64_t *d_rowptr = nullptr;
    int32_t *d_col = nullptr;
    int64_t *d_seg_ptr = nullptr;   // nnz + 1
    int32_t *d_seg_src = nullptr;   // ntrip: index into the element-matrix buffer
    int64_t *d_nodeblk = nullptr;   // n_nodes * 36: CSR position of node diagonal 6x6 block (-1 = absent)
    std::vector<int64_t> nodeblk;
    std::vector<int32_t> diagpos;   // ndof: CSR position of (i,i) or -1
};

struct jfem_matrix {
    int device = 0;
    const jfem_pattern *P = nullptr;
    double *d_val = nullptr;
    double *d_Ke = nullptr;         // element matrices (ntrip doubles), element-major
    bool view = false;              // a variant of a jfem_batch: storage owned by the batch
};

struct jfem_factor;
struct jfem_symbolic;

// V design variants of one topology (same mesh and pattern, own section records):
// sections, element matrices, K values and, after jfem_batch_factor, the numeric
// factors of all variants, each stored variant after variant at the strides below
struct jfem_batch {
    int device = 0;
    const jfem_mesh *mesh = nullptr;
    const jfem_pattern *P = nullptr;
    int64_t V = 0;
    double *d_qp = nullptr, *d_tp = nullptr, *d_tetp = nullptr;   // V section sets
    double *d_Ke = nullptr, *d_val = nullptr;
    int64_t Ke_stride = 0, val_stride = 0;                         // doubles
    std::vector<jfem_matrix *> mviews;
    const jfem_symbolic *S = nullptr;                              // of the factors below
    int precision = 0;
    void *d_L = nullptr, *d_U = nullptr, *d_Linv = nullptr;
    int64_t L_stride = 0, U_stride = 0, inv_stride = 0;            // entries
    std::vector<jfem_factor *> fviews;
};

// per-variant strides rounded up to 256 bytes, so every variant's storage starts
// as aligned as a single allocation
// A HIP call whose status is not taken (a free on a release path) leaves its
// error in the calling thread's last-error slot, where the next launch check
// (hipGetLastError after a launch) would report it as its own: the entry points
// that check launches that way start from a cleared slot.
inline void clear_stale_hip_error() { (void)hipGetLastError(); }

inline int64_t jfem_stride(int64_t n, int64_t es) {
    const int64_t q = 256 / es;
    return ((n + q - 1) / q) * q;
}

namespace jfem {
int fail(jfem_ctx *ctx, int code, const std::string &msg);
// timeline of timing level 2: tl_begin starts a sequence; tl_mark records an event
// after the kernel just issued, charged to family fam (fam < 0: not charged);
// tl_flush waits for the last event and adds each interval to its family
void tl_begin(jfem_ctx *ctx);
void tl_mark(jfem_ctx *ctx, int fam);
void tl_flush(jfem_ctx *ctx);
void timer_start(jfem_ctx *ctx);
void timer_stop(jfem_ctx *ctx, int family, int64_t launches);

// element generation (elements.hip)
hipError_t launch_elements(hipStream_t s, const jfem_mesh *m, const double *qp, const double *tp, const double *tetp,
                           double *Ke, int64_t Ke_stride, int V, int kinds_mask, int64_t *status);
hipError_t launch_q4_local(hipStream_t s, int64_t n, const double *lc, const double *Cm,
                           const double *Cb, const double *Cs, const double *h,
                           const double *k6, const double *ds, const int32_t *flags,
                           double *Ke, int64_t *status);
hipError_t launch_t3_local(hipStream_t s, int64_t n, const double *lc, const double *Dm, const double *Db,
                           const double *Ds, const double *k6, double *Ke, int64_t *status);
// differential stiffness of a mesh's shells for device displacements u (6 per node):
// CQUAD4s at Ke (576 each), then CTRIA3s (324 each)
hipError_t launch_shell_kg(hipStream_t s, const jfem_mesh *m, const double *u, double *Ke, int64_t *status);
// c = a + sigma b; and its diagonally scaled form (scale factors of a into d_sd)
hipError_t launch_csr_shift(hipStream_t s, int64_t n, const double *a, double sigma, const double *b, double *c);
hipError_t launch_csr_shift_scaled(hipStream_t s, const jfem_pattern *P, const int32_t *d_diagpos, double *d_sd,
                                   const double *a, double sigma, const double *b, double *c);
// shell force / stress / strain recovery of a mesh's CQUAD4s (44 doubles each) and
// CTRIA3s (20 each) from device displacements u (6 per node)
hipError_t launch_shell_recover(hipStream_t s, const jfem_mesh *m, const double *u, int mean_shear, double *qout,
                                double *tout, int64_t *status);
hipError_t launch_tet10_direct(hipStream_t s, int64_t n, const double *X, const double *prop, double *Ke,
                               int64_t *status);
hipError_t launch_assemble(hipStream_t s, const jfem_pattern *P, const double *Ke, double *val, int V,
                           int64_t Ke_stride, int64_t val_stride);
void batch_free_factor(jfem_batch *B);   // factor.hip
hipError_t launch_bc_screen(hipStream_t s, const jfem_pattern *P, const double *val, const uint8_t *spc, double thresh,
                            uint8_t *fixed0, uint8_t *flag);
hipError_t launch_bc_decide(hipStream_t s, const jfem_pattern *P, const double *val, const uint8_t *flag,
                            double eps_zero, uint8_t *fixed);
hipError_t launch_gather_nodeblk(hipStream_t s, const jfem_pattern *P, const double *val, double *out);
hipError_t launch_reset_status(hipStream_t s, int64_t *status);
}  // namespace jfem
