// Block kernels of the buckling eigen-extraction: the products and projections
// that the subspace (inverse) iteration applies to its block of p vectors, all on
// the device in FULL DOF numbering (ndof x p, column-major):
//   Y = A X restricted to the free DOFs (fixed rows zero, fixed columns skipped),
//   C = X' Y (p x q, summed in a fixed order, copied to the host),
//   X = Y Q (or X += Y Q) for a small host matrix Q (p x q).
// The reference extracts EIGB roots with a shifted sparse factorization and an
// eigensolver on (K_ff, -Kg_ff) (_solve_eigb_buckling, src/solver/helpers.jl:
// 280-420; solve_buckling, src/solver/solve_case.jl:2803); here the factor and
// solves are the library's, and these kernels carry the O(ndof) work between
// them.  Every sum runs in an order fixed by the data layout alone (no atomics),
// so results are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "jfem_internal.h"

namespace {

constexpr int kBlkBS = 256;
constexpr int kMaxCols = 32;           // columns of one block
constexpr int64_t kGramRows = 4096;    // rows per partial of a projection

// Y[:, c] = A[free, free] X[:, c] for c < ncol; one thread per row, entries in
// row order; a fixed row gets zeros
__global__ void __launch_bounds__(kBlkBS) block_spmv_kernel(int64_t n, int ncol, const int64_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ col,
                                                            const double *__restrict__ val,
                                                            const uint8_t *__restrict__ fixed,
                                                            const double *__restrict__ X, double *__restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[kMaxCols];
    for (int c = 0; c < ncol; ++c) acc[c] = 0.0;
    if (!(fixed && fixed[i])) {
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int64_t j = col[p];
            if (fixed && fixed[j]) continue;
            const double a = val[p];
            for (int c = 0; c < ncol; ++c) acc[c] += a * X[j + n * c];
        }
    }
    for (int c = 0; c < ncol; ++c) Y[i + n * c] = acc[c];
}

// partial projections: part[b][a + p*c] = sum over rows of chunk b of X[:, a] Y[:, c]
// (rows in order); one workgroup per (chunk, pair group)
__global__ void __launch_bounds__(kBlkBS) gram_partial_kernel(int64_t n, int p, int q, const double *__restrict__ X,
                                                              const double *__restrict__ Y,
                                                              double *__restrict__ part) {
    const int64_t r0 = (int64_t)blockIdx.x * kGramRows, r1 = min<int64_t>(r0 + kGramRows, n);
    const int pairs = p * q;
    for (int e = blockIdx.y * kBlkBS + threadIdx.x; e < pairs; e += gridDim.y * kBlkBS) {
        const int a = e % p, c = e / p;
        const double *x = X + n * a, *y = Y + n * c;
        double s = 0.0;
        for (int64_t r = r0; r < r1; ++r) s += x[r] * y[r];
        part[(int64_t)blockIdx.x * pairs + e] = s;
    }
}

// C[e] = sum over chunks b (in order) of part[b][e]
__global__ void __launch_bounds__(kBlkBS) gram_sum_kernel(int64_t nchunk, int pairs, const double *__restrict__ part,
                                                          double *__restrict__ C) {
    const int e = blockIdx.x * kBlkBS + threadIdx.x;
    if (e >= pairs) return;
    double s = 0.0;
    for (int64_t b = 0; b < nchunk; ++b) s += part[b * pairs + e];
    C[e] = s;
}

// X[:, c] = (accumulate ? X[:, c] : 0) + sum_a Y[:, a] Q[a, c] (a in order, the sum
// formed first and added once)
__global__ void __launch_bounds__(kBlkBS) block_combine_kernel(int64_t n, int p, int q, const double *__restrict__ Y,
                                                               const double *__restrict__ Q, double *__restrict__ X,
                                                               int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double y[kMaxCols];
    for (int a = 0; a < p; ++a) y[a] = Y[i + n * a];
    for (int c = 0; c < q; ++c) {
        double s = 0.0;
        for (int a = 0; a < p; ++a) s += y[a] * Q[a + p * c];
        X[i + n * c] = accumulate ? X[i + n * c] + s : s;
    }
}

inline unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlkBS - 1) / kBlkBS); }

}  // namespace

int jfem_matrix_spmv(jfem_ctx *ctx, const jfem_matrix *A, const uint8_t *fixed, int ncol, const double *x_dev,
                     double *y_dev) {
    if (!ctx || !A) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (ncol < 1 || ncol > kMaxCols) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "block of 1 to 32 columns");
    if (!x_dev || !y_dev || x_dev == y_dev) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "x and y: distinct device blocks");
    const jfem_pattern *P = A->P;
    const int64_t n = P->ndof;
    hipStream_t s = ctx->stream;
    uint8_t *d_fixed = nullptr;
    if (fixed) {
        JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_fixed, std::max<int64_t>(n, 1)));
        hipError_t e = hipMemcpyAsync(d_fixed, fixed, (size_t)n, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            (void)hipFree(d_fixed);
            return jfem::fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
        }
    }
    hipLaunchKernelGGL(block_spmv_kernel, dim3(blocks(n)), dim3(kBlkBS), 0, s, n, ncol, P->d_rowptr, P->d_col, A->d_val,
                       d_fixed, x_dev, y_dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_fixed) (void)hipFree(d_fixed);
    if (e != hipSuccess) return jfem::fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    return JFEM_OK;
}

int jfem_block_gram(jfem_ctx *ctx, int64_t n, int p, const double *x_dev, int q, const double *y_dev, double *c_host) {
    if (!ctx || !x_dev || !y_dev || !c_host) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (n < 1 || p < 1 || q < 1 || p > kMaxCols || q > kMaxCols) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "block shape");
    hipStream_t s = ctx->stream;
    const int64_t nchunk = (n + kGramRows - 1) / kGramRows;
    const int pairs = p * q;
    double *d_part = nullptr, *d_C = nullptr;
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_part, (size_t)(nchunk * pairs + pairs) * sizeof(double)));
    d_C = d_part + nchunk * pairs;
    const unsigned gy = (unsigned)((pairs + kBlkBS - 1) / kBlkBS);
    hipLaunchKernelGGL(gram_partial_kernel, dim3((unsigned)nchunk, gy), dim3(kBlkBS), 0, s, n, p, q, x_dev, y_dev, d_part);
    hipLaunchKernelGGL(gram_sum_kernel, dim3(gy), dim3(kBlkBS), 0, s, nchunk, pairs, d_part, d_C);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(c_host, d_C, (size_t)pairs * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_part);
    if (e != hipSuccess) return jfem::fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    return JFEM_OK;
}

int jfem_block_combine(jfem_ctx *ctx, int64_t n, int p, const double *y_dev, int q, const double *Q_host,
                       double *x_dev, int accumulate) {
    if (!ctx || !y_dev || !Q_host || !x_dev) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (n < 1 || p < 1 || q < 1 || p > kMaxCols || q > kMaxCols) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "block shape");
    if (x_dev == y_dev) return jfem::fail(ctx, JFEM_ERR_BAD_ARG, "x and y: distinct device blocks");
    hipStream_t s = ctx->stream;
    double *d_Q = nullptr;
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_Q, (size_t)(p * q) * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d_Q, Q_host, (size_t)(p * q) * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(block_combine_kernel, dim3(blocks(n)), dim3(kBlkBS), 0, s, n, p, q, y_dev, d_Q, x_dev,
                           accumulate);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_Q);
    if (e != hipSuccess) return jfem::fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    return JFEM_OK;
}
