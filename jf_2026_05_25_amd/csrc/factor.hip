// Supernodal multifrontal Cholesky of K_ff on gfx950, its level-scheduled
// triangular solves, the FP64 residual SpMV and the iterative-refinement driver.
//
// Storage per supernode s (ncol n, front order m): L_s is m x n column-major
// (rows = the front's row list), the update-matrix storage U_s is (m-n) x (m-n).
// Factorization, level by level from the leaves:
//   scatter K entries into L (once, all supernodes), then for each 64-column
//   panel: factor the diagonal block + triangular solve of the panel rows
//   (panel kernels), and a rank-64 update of the trailing pivot columns (update
//   kernel, MFMA tiles); after the level's last panel the update matrix of each
//   front, with its full rank, added straight into the parent's front (Schur
//   kernel, one launch per sibling rank: deterministic).  All supernodes of a
//   level are batched into the same launches.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>

#include "jfem_internal.h"
#include "symbolic.h"

using namespace jfem;

struct jfem_symbolic {
    int device = 0;
    const jfem_pattern *P = nullptr;
    SymbolicHost H;
    int64_t *d_col0 = nullptr, *d_ncol = nullptr, *d_m = nullptr, *d_lofs = nullptr, *d_uofs = nullptr, *d_iofs = nullptr,
            *d_rofs = nullptr, *d_relofs = nullptr;
    int32_t *d_rows = nullptr, *d_relmap = nullptr, *d_child_ptr = nullptr, *d_child = nullptr, *d_perm = nullptr,
            *d_level_sn = nullptr;
    int32_t *d_amap_src = nullptr;
    int64_t *d_amap_dst = nullptr;
    PanelTask *d_ptasks = nullptr;
    FlowTask *d_flow_f = nullptr, *d_flow_b = nullptr;
    FlowChild *d_flow_gb = nullptr;
    UpdateTask *d_utasks = nullptr;
    PushTask *d_stasks = nullptr;
    InvTask *d_itasks = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;              // device bytes behind allocs
};

struct jfem_factor {
    int device = 0;
    const jfem_symbolic *S = nullptr;
    int precision = 64;
    void *d_L = nullptr;
    void *d_U = nullptr;
    void *d_Linv = nullptr;         // inverse diagonal blocks of the panels, X = L_11^{-1} of the multi-panel supernodes (solves)
    int64_t bytes = 0;
    bool view = false;              // a variant of a jfem_batch: storage owned by the batch
};

namespace {

constexpr int kBS = 256;

// Variants: a launch over V matrices that share one symbolic factorization runs
// variant blockIdx.y on storage offset by blockIdx.y times these strides
// (elements): K values, factor L, update matrices U, inverse blocks; failures are
// reported at index blockIdx.y * col + (column).  All zero for a single matrix.
struct VarStride {
    int64_t K = 0, L = 0, U = 0, inv = 0, col = 0;
};

template <class T>
__device__ inline T tsqrt(T x);
template <>
__device__ inline double tsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ inline float tsqrt<float>(float x) { return sqrtf(x); }

__device__ inline void report_min(int64_t *status, int64_t idx, int code) {
    unsigned long long old = atomicMin((unsigned long long *)&status[0], (unsigned long long)idx);
    if ((long long)old > idx) atomicExch((unsigned long long *)&status[1], (unsigned long long)code);
}

// ----------------------------------------------------------------- fill / scatter
// before a dataflow solve: Z, X, V and Tb (nv entries each), W (nw) and the
// partial sums P (np) to the pending mark, the two launch tickets to zero
template <class T>
__global__ void solve_marks_kernel(T *__restrict__ Z, T *__restrict__ X, T *__restrict__ V, T *__restrict__ Tb,
                                   int64_t nv, T *__restrict__ W, int64_t nw, T *__restrict__ P, int64_t np, T mark,
                                   int32_t *__restrict__ tickets) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < 2) tickets[i0] = 0;
    for (int64_t i = i0; i < nv; i += stride) {
        Z[i] = mark;
        X[i] = mark;
        V[i] = mark;
        Tb[i] = mark;
    }
    for (int64_t i = i0; i < nw; i += stride) W[i] = mark;
    for (int64_t i = i0; i < np; i += stride) P[i] = mark;
}

template <class T>
__global__ void fill_kernel(T *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = T(0);
}

// zero n entries of a 16-byte-aligned array with 16-byte stores (the factor's
// storage, about 1.4 GB in FP32 on the wing box: 4-byte stores leave it well
// below the write bandwidth)
template <class T>
__global__ void zero16_kernel(T *__restrict__ p, int64_t n) {
    const int64_t nv = (n * (int64_t)sizeof(T)) / 16;
    uint4 *q = (uint4 *)p;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) q[i] = make_uint4(0, 0, 0, 0);
    const int64_t done = nv * 16 / (int64_t)sizeof(T);
    for (int64_t i = done + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T(0);
}

template <class T>
__global__ void scatter_kernel(int64_t n, const int32_t *__restrict__ src, const int64_t *__restrict__ dst,
                               const double *__restrict__ val, T *__restrict__ L, VarStride vs) {
    val += blockIdx.y * vs.K;
    L += blockIdx.y * vs.L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        L[dst[i]] = (T)val[src[i]];
}

// ----------------------------------------------------------------- broadcasts
// value of v in lane `lane` (uniform lane index), for all lanes of the wave
template <class T>
__device__ inline T bcast(T v, int lane);
template <>
__device__ inline float bcast<float>(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <>
__device__ inline double bcast<double>(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// inverse of each panel's diagonal block, for the triangular solves (one
// wavefront per panel, after the factorization): lane c computes column c of
// L_pp^{-1} by forward substitution in FP64, L_pp read as broadcasts from LDS;
// stored column-major with leading dimension b, and for a supernode with two or
// more panels also as diagonal block p of its X = L_11^{-1} (leading dimension ns)
template <class T>
__global__ void __launch_bounds__(64) diag_inverse_kernel(const FlowTask *__restrict__ tasks, const T *__restrict__ L,
                                                          T *__restrict__ Linv, VarStride vs) {
    __shared__ __attribute__((aligned(16))) T D[kPanel][kPanel + 4];
    L += blockIdx.y * vs.L;
    Linv += blockIdx.y * vs.inv;
    const FlowTask t = tasks[blockIdx.x];
    const int b = t.nr, lane = threadIdx.x;
    const int64_t k = t.r0, ms = t.ms;
    const T *Ls = L + t.lofs;
    const int li = min(lane, b - 1);
#pragma unroll 8
    for (int j = 0; j < kPanel; ++j) {
        const T v = Ls[(k + li) + ms * (k + min(j, b - 1))];
        D[lane][j] = (lane < b && j <= lane) ? v : (lane == j ? T(1) : T(0));
    }
    __syncthreads();
    // substitution in FP64, rounded once: the inverse of the factor's own block
    double x[kPanel];
#pragma unroll
    for (int i = 0; i < kPanel; ++i) {
        double s = i == lane ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < i; ++j) s = fma(-(double)D[i][j], x[j], s);
        x[i] = s / (double)D[i][i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPanel; ++i) D[i][lane] = (T)x[i];
    __syncthreads();
    T *dst = Linv + t.iofs;
    if (lane < b)
        for (int c = 0; c < b; ++c) dst[lane + (int64_t)b * c] = D[lane][c];
    if (t.xofs >= 0 && lane < b) {
        T *xd = Linv + t.xofs + k + t.ns * k;
        for (int c = 0; c < b; ++c) xd[lane + t.ns * c] = D[lane][c];
    }
}

// ----------------------------------------------------------------- update kernel (reference)
// Scalar version of the trailing update, used only by jfem_selftest to check the
// MFMA kernel's operand/accumulator lane layouts on the device.
// C[r, c] -= sum_p A[r, p] A[c, p] over the panel columns p in [k, k+b) for the
// 64 x 64 tile at (r0, c0) of the trailing front (lower triangle only).
template <class T>
__global__ void __launch_bounds__(kBS) update_ref_kernel(const UpdateTask *__restrict__ tasks, const int64_t *__restrict__ ncol,
                                                     const int64_t *__restrict__ mrow, const int64_t *__restrict__ lofs,
                                                     const int64_t *__restrict__ uofs, T *__restrict__ L, T *__restrict__ U) {
    __shared__ T A[kPanel][kTile + 1];   // A[p][i]  (transposed tile of rows r0..)
    __shared__ T B[kPanel][kTile + 1];   // B[p][j]  (rows c0..)
    const UpdateTask t = tasks[blockIdx.x];
    const int s = t.s;
    const int64_t ns = ncol[s], ms = mrow[s], k = t.k;
    const int b = (int)min<int64_t>(kPanel, ns - k);
    const int64_t r0 = t.r0, c0 = t.c0;
    const int64_t rb = r0 < ns ? ns : ms, cb = c0 < ns ? ns : ms;
    const int nr = (int)(min<int64_t>(r0 + kTile, rb) - r0);
    const int nc = (int)(min<int64_t>(c0 + kTile, cb) - c0);
    const T *Ls = L + lofs[s];
    const int tid = threadIdx.x;
    for (int q = tid; q < kTile * b; q += blockDim.x) {
        int i = q % kTile, p = q / kTile;
        A[p][i] = i < nr ? Ls[(r0 + i) + ms * (k + p)] : T(0);
        B[p][i] = i < nc ? Ls[(c0 + i) + ms * (k + p)] : T(0);
    }
    __syncthreads();
    // 256 threads, each a 4x4 micro-tile of the 64x64 tile
    const int ti = (tid % 16) * 4, tj = (tid / 16) * 4;
    T acc[4][4];
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) acc[a][c] = T(0);
    for (int p = 0; p < b; ++p) {
        T av[4], bv[4];
        for (int a = 0; a < 4; ++a) { av[a] = A[p][ti + a]; bv[a] = B[p][tj + a]; }
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 4; ++c) acc[a][c] += av[a] * bv[c];
    }
    const bool inL = c0 < ns;
    T *Cs = inL ? (L + lofs[s]) : (U + uofs[s]);
    const int64_t ld = inL ? ms : ms - ns;
    const int64_t roff = inL ? 0 : ns, coff = inL ? 0 : ns;
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
            const int i = ti + a, j = tj + c;
            if (i < nr && j < nc && r0 + i >= c0 + j) Cs[(r0 + i - roff) + ld * (c0 + j - coff)] -= acc[a][c];
        }
}

// ----------------------------------------------------------------- update kernel
// C[r, c] -= sum_p A[r, p] A[c, p] over the panel columns p in [k, k+b) for the
// 64 x 64 tile at (r0, c0) of the trailing front (lower triangle only), on the
// matrix cores: 4 wavefronts, each a 32 x 32 quarter as 2 x 2 blocks of
// v_mfma_{f64,f32}_16x16x4.  The product is formed transposed, D[j][i] =
// sum_p A[c0+j][p] A[r0+i][p], so the accumulator's lane index runs along the
// column-major rows of C and the read-modify-write of C is coalesced.
// 16x16x4 operand layout (both types): lane l holds A[l%16][l/16] and
// B[l/16][l%16].  Accumulator layouts differ: the f32 instruction returns one
// group of 4 consecutive rows per 16-lane block, element v of lane l being
// D[4*(l/16)+v][l%16]; the f64 instruction returns 4 groups of one row,
// element v of lane l being D[4*v+(l/16)][l%16].  drow() gives that row.
template <class T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ inline int drow(int blk, int v) { return 4 * v + blk; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ inline int drow(int blk, int v) { return 4 * blk + v; }
};

// ----------------------------------------------------------------- panel factorization
// 64 x 64 Cholesky of a panel's diagonal block on a 256-thread workgroup, with
// the inverses of its four diagonal 16 x 16 blocks.  Thread (lane = row i, wave
// w) keeps entries [16w, 16w+16) of row i in a[16] (identity beyond the block
// width b).  Block step jb: wavefront jb alone eliminates its 16 columns over all
// 64 rows (pivots and multipliers broadcast with readlane: the sequential part
// stays inside one wavefront, with no barrier), publishes them in an LDS panel
// (double-buffered, rows padded to 16 bytes for 128-bit broadcast reads), and the
// later wavefronts apply the 16-wide update to their own columns: one workgroup
// barrier per block step.  Pivots use the hardware reciprocal square root
// (L_jj = d * rsq(d), column scaled by rsq(d)).
// Then wavefront w forms D_w^{-1}, the inverse of diagonal block w, by forward
// substitution in FP64 (lane c: column c; entries of D_w broadcast from the
// registers), rounded once to T: the tile solves of the panel rows use them.
template <class T>
__device__ inline T rsq_hw(T x);
template <>
__device__ inline float rsq_hw<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <>
__device__ inline double rsq_hw<double>(double x) { return rsqrt(x); }

template <class T>
struct PanelLds {
    static constexpr int kRow = 16 + 16 / (int)sizeof(T);          // padded panel row (elements)
    static constexpr int bytes = 2 * 64 * kRow * (int)sizeof(T) + 16 + 2 * 16 * (int)sizeof(T);
};

// A = L L' in place for the block (rows/columns >= b: identity).  lds: PanelLds<T>::bytes
// of 16-byte aligned shared memory.  Stores L's lower triangle at Lblk (leading
// dimension ld, rows/columns < b) and the lower triangles of D_w^{-1} at their
// positions in the b x b column-major block inverse at inv.  Returns the first
// column with a non-positive pivot, b if none (uniform).
// SG (signed): A = L J L' with J = diag(+-1) and L_jj = sqrt|d_j| > 0, J_j the sign
// of the pivot d_j (the unpivoted LDL' of an indefinite matrix, D = J L_jj^2):
// multipliers L_ij = J_j a_ij / L_jj, updates a_ic -= L_ij J_j L_cj; J of the
// block's columns is written to Jblk[0..b).  A pivot of 0 or not finite is bad.
template <class T, bool SG = false>
__device__ __forceinline__ int potrf_dinv(T (&a)[16], char *lds, int lane, int w, int b, T *Lblk, int64_t ld, T *inv,
                                          T *Jblk = nullptr) {
    constexpr int PS = PanelLds<T>::kRow;
    T (*P)[64][PS] = reinterpret_cast<T (*)[64][PS]>(lds);
    int *badw = reinterpret_cast<int *>(lds + 2 * 64 * PS * sizeof(T));
    T (*Jp)[16] = reinterpret_cast<T (*)[16]>(lds + 2 * 64 * PS * sizeof(T) + 16);
    int bad = b;
    const int nb = (b + 15) / 16;
    for (int jb = 0; jb < nb; ++jb) {
        if (w == jb) {
            T sg[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int j = 16 * jb + jj;
                const T d = bcast(a[jj], j);
                const bool ok = SG ? (d != T(0) && isfinite(d)) : d > T(0);
                bad = (!ok && j < b && bad == b) ? j : bad;
                const T sj = SG ? (d < T(0) ? T(-1) : T(1)) : T(1);
                const T dd = ok ? (SG ? fabs(d) : d) : T(1);
                const T r = rsq_hw(dd);
                sg[jj] = sj;
                a[jj] = lane > j ? (SG ? a[jj] * r * sj : a[jj] * r) : (lane == j ? dd * r : a[jj]);
#pragma unroll
                for (int cc = jj + 1; cc < 16; ++cc)
                    a[cc] -= (SG ? a[jj] * sj : a[jj]) * bcast(a[jj], 16 * jb + cc);
            }
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) P[jb & 1][lane][jj] = a[jj];
            if (SG && lane < 16) {
#pragma unroll
                for (int jj = 0; jj < 16; ++jj)
                    if (lane == jj) Jp[jb & 1][jj] = sg[jj];
                if (16 * jb + lane < b) {
                    T v = sg[0];
#pragma unroll
                    for (int jj = 1; jj < 16; ++jj) v = lane == jj ? sg[jj] : v;
                    Jblk[16 * jb + lane] = v;
                }
            }
        }
        __syncthreads();
        if (w > jb) {
            T pl[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) pl[jj] = SG ? P[jb & 1][lane][jj] * Jp[jb & 1][jj] : P[jb & 1][lane][jj];
#pragma unroll
            for (int cc = 0; cc < 16; ++cc) {
                const T *pc = P[jb & 1][16 * w + cc];
                T s = T(0);
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) s += pl[jj] * pc[jj];
                a[cc] -= s;
            }
        }
        // (the next block's owner writes the other buffer; this one is rewritten
        // two block steps later, after every wave has passed the barrier between)
    }
    if (lane == 0) badw[w] = w < nb ? bad : b;
    if (lane < b) {
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const int c = 16 * w + cc;
            if (c <= lane) Lblk[lane + ld * c] = a[cc];
        }
    }
    if (w < nb) {
        const int o = 16 * w, c = lane & 15;
        T dl = a[0];
#pragma unroll
        for (int r = 1; r < 16; ++r) dl = (lane - o == r) ? a[r] : dl;
        const double rdl = 1.0 / (double)dl;
        double y[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) y[r] = r == c ? 1.0 : 0.0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            y[r] *= bcast(rdl, o + r);
#pragma unroll
            for (int q = r + 1; q < 16; ++q) y[q] = fma(-(double)bcast(a[r], o + q), y[r], y[q]);
        }
        if (lane < 16 && o + c < b)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (r >= c && o + r < b) inv[(o + r) + (int64_t)b * (o + c)] = (T)y[r];
    }
    __syncthreads();
    int first = b;
#pragma unroll
    for (int q = 0; q < 4; ++q) first = min(first, badw[q]);
    return first;
}

// factor the diagonal blocks of the first panels of a level's supernodes (one
// workgroup per task; tasks are the level's step-0 diagonal panel tasks)
template <class T, bool SG = false>
__global__ void __launch_bounds__(kBS) diag_factor_kernel(const PanelTask *__restrict__ tasks, const int64_t *__restrict__ ncol,
                                                          const int64_t *__restrict__ mrow, const int64_t *__restrict__ lofs,
                                                          const int64_t *__restrict__ iofs, const int64_t *__restrict__ col0,
                                                          T *__restrict__ L, T *__restrict__ Linv,
                                                          int64_t *__restrict__ status, VarStride vs, T *__restrict__ J) {
    __shared__ __attribute__((aligned(16))) char lds[PanelLds<T>::bytes];
    L += blockIdx.y * vs.L;
    Linv += blockIdx.y * vs.inv;
    const PanelTask t = tasks[blockIdx.x];
    const int s = t.s;
    const int64_t ns = ncol[s], ms = mrow[s], k = t.k;
    const int b = (int)min<int64_t>(kPanel, ns - k);
    T *Ls = L + lofs[s];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int li = min(lane, b - 1);
    T a[16];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
        const int c = 16 * w + cc;
        const T v = Ls[(k + li) + ms * (k + min(c, b - 1))];
        a[cc] = (c <= lane && lane < b) ? v : (c == lane ? T(1) : T(0));
    }
    const int bad = potrf_dinv<T, SG>(a, lds, lane, w, b, Ls + k + ms * k, ms,
                                      Linv + iofs[s] + (int64_t)kPanel * kPanel * (k / kPanel),
                                      SG ? J + blockIdx.y * vs.col + col0[s] + k : nullptr);
    if (bad < b && tid == 0) report_min(status, blockIdx.y * vs.col + col0[s] + k + bad, JFEM_ERR_NOT_SPD);
}

// tile solve of panel task (s, k, r0): rows [max(r0, k+b), r0+kTile) of the panel
// get X <- X L_kk^{-T}, on the matrix cores, blocked by 16 columns.  Wave w takes
// tile rows [16w, 16w+16) and needs no other wave: with Y_j = X_j' (columns
// [16j, 16j+16) of its rows, transposed), Y_j = D_j^{-1} (A_j' - sum_{q<j} L_jq Y_q)
// where L_jq are 16-blocks of L_kk and D_j^{-1} the inverses of its diagonal
// blocks (written by the factorization of L_kk).  The transposed form puts the
// accumulator's lane index along the column-major rows, so loads and stores of
// the tile are coalesced.  Operand layout of the 16x16x4 MFMA as for the update.
template <class T>
__global__ void __launch_bounds__(kBS) tile_solve_kernel(const PanelTask *__restrict__ tasks, const int64_t *__restrict__ ncol,
                                                         const int64_t *__restrict__ mrow, const int64_t *__restrict__ lofs,
                                                         const int64_t *__restrict__ iofs, T *__restrict__ L,
                                                         const T *__restrict__ Linv, VarStride vs) {
    typedef typename Mfma<T>::v4 v4;
    __shared__ T Lb[6][16][17];          // L_jq (j > q) at index j(j-1)/2 + q
    __shared__ T Di[4][16][17];          // D_j^{-1} (identity beyond b)
    __shared__ T Ys[4][4][16][16];       // per wave: Y_j, and the staging of the right-hand factor
    L += blockIdx.y * vs.L;
    Linv += blockIdx.y * vs.inv;
    const PanelTask t = tasks[blockIdx.x];
    const int s = t.s;
    const int64_t ns = ncol[s], ms = mrow[s], k = t.k;
    const int b = (int)min<int64_t>(kPanel, ns - k);
    const int64_t rs = max<int64_t>(t.r0, k + b), re = min<int64_t>(t.r0 + kTile, ms);
    if (rs >= re) return;
    T *Ls = L + lofs[s];
    const T *Dg = Linv + iofs[s] + (int64_t)kPanel * kPanel * (k / kPanel);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int nbj = (b + 15) / 16;
    // this lane's row of the tile (accumulator column l16 of wave w)
    const int64_t row = t.r0 + 16 * w + l16;
    const bool rowok = row >= rs && row < re;
    const int64_t rowc = min<int64_t>(max<int64_t>(row, rs), re - 1);
    // the tile's blocks A_j' in accumulator layout: element v = A[row][16j + drow(l4, v)]
    T pre[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int col = 16 * j + Mfma<T>::drow(l4, v);
            pre[j][v] = Ls[rowc + ms * (k + min(col, b - 1))];
            if (!(rowok && col < b)) pre[j][v] = T(0);
        }
    for (int e = tid; e < 6 * 256; e += kBS) {
        const int blk = e >> 8, r = e & 15, c = (e >> 4) & 15;
        const int j = blk < 1 ? 1 : (blk < 3 ? 2 : 3), q = blk - j * (j - 1) / 2;
        const int rr = 16 * j + r, cc = 16 * q + c;
        Lb[blk][r][c] = (rr < b && cc < b) ? Ls[(k + rr) + ms * (k + cc)] : T(0);
    }
    for (int e = tid; e < 4 * 256; e += kBS) {
        const int j = e >> 8, r = e & 15, c = (e >> 4) & 15;
        const int rr = 16 * j + r, cc = 16 * j + c;
        Di[j][r][c] = (rr < b && cc < b) ? (r >= c ? Dg[rr + (int64_t)b * cc] : T(0)) : (r == c ? T(1) : T(0));
    }
    __syncthreads();
    for (int j = 0; j < nbj; ++j) {
        v4 acc = v4{pre[j][0], pre[j][1], pre[j][2], pre[j][3]};
        for (int q = 0; q < j; ++q) {
            const int blk = j * (j - 1) / 2 + q;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                acc = Mfma<T>::mac(-Lb[blk][l16][4 * kk + l4], Ys[w][q][4 * kk + l4][l16], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Ys[w][j][Mfma<T>::drow(l4, v)][l16] = acc[v];
        v4 y = v4{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) y = Mfma<T>::mac(Di[j][l16][4 * kk + l4], Ys[w][j][4 * kk + l4][l16], y);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            Ys[w][j][Mfma<T>::drow(l4, v)][l16] = y[v];
            const int col = 16 * j + Mfma<T>::drow(l4, v);
            if (rowok && col < b) Ls[row + ms * (k + col)] = y[v];
        }
    }
}

// factor_next != 0: the task whose tile holds the next panel's diagonal block
// (r0 == c0 == k + b < ns) also factors that block after its update (staged
// through LDS, blocked over the four wavefronts), so the next step's tile solves
// can start at once.
// SG (signed factorization): C -= W J W' with the panel's signs J on the column
// operand, and the next diagonal block factored with signs.
template <class T, bool SG = false>
__global__ void __launch_bounds__(kBS) update_kernel(const UpdateTask *__restrict__ tasks, const int64_t *__restrict__ ncol,
                                                     const int64_t *__restrict__ mrow, const int64_t *__restrict__ lofs,
                                                     const int64_t *__restrict__ uofs, const int64_t *__restrict__ iofs,
                                                     T *__restrict__ L, T *__restrict__ U, T *__restrict__ Linv,
                                                     int factor_next, const int64_t *__restrict__ col0,
                                                     int64_t *__restrict__ status, VarStride vs, T *__restrict__ J) {
    typedef typename Mfma<T>::v4 v4;
    L += blockIdx.y * vs.L;
    U += blockIdx.y * vs.U;
    Linv += blockIdx.y * vs.inv;
    __shared__ T A[kPanel][kTile + 1];   // A[p][i] = L[r0 + i, k + p]
    __shared__ __attribute__((aligned(16))) T B[kPanel][kTile + 1];   // B[p][j] = L[c0 + j, k + p]
    static_assert(PanelLds<T>::bytes <= (int)sizeof(T) * kPanel * (kTile + 1), "panel LDS fits in B");
    const UpdateTask t = tasks[blockIdx.x];
    const int s = t.s;
    const int64_t ns = ncol[s], ms = mrow[s], k = t.k;
    const int b = (int)min<int64_t>(kPanel, ns - k);
    const int b4 = (b + 3) & ~3;
    const int64_t r0 = t.r0, c0 = t.c0;
    const int64_t rb = r0 < ns ? ns : ms, cb = c0 < ns ? ns : ms;
    const int nr = (int)(min<int64_t>(r0 + kTile, rb) - r0);
    const int nc = (int)(min<int64_t>(c0 + kTile, cb) - c0);
    const T *Ls = L + lofs[s];
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wi = (wave & 1) * 32;      // rows of C (i) handled by this wave
    const int wj = (wave >> 1) * 32;     // columns of C (j)
    const int l16 = lane & 15, l4 = lane >> 4;
    const bool inL = c0 < ns;
    T *Cs = inL ? (L + lofs[s]) : (U + uofs[s]);
    const int64_t ld = inL ? ms : ms - ns;
    const int64_t off = inL ? 0 : ns;
    // Every load of the workgroup goes out before the first wait: the two panel
    // blocks (16 + 16 entries per thread) and this lane's 16 entries of C, at
    // addresses clamped into the blocks (out-of-block values are replaced by
    // zero afterwards), so one memory round trip covers them all.
    T av[16], bv[16], cpre[2][2][4];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = tid + kBS * u, i = q % kTile, p = min(q / kTile, b - 1);
        av[u] = Ls[(r0 + min(i, nr - 1)) + ms * (k + p)];
        bv[u] = Ls[(c0 + min(i, nc - 1)) + ms * (k + p)];
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = min(wi + 16 * x + l16, nr - 1), j = min(wj + 16 * y + Mfma<T>::drow(l4, v), nc - 1);
                cpre[x][y][v] = Cs[(r0 + i - off) + ld * (c0 + j - off)];
            }
    if (SG) {
        const T *Jk = J + blockIdx.y * vs.col + col0[s] + k;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int p = min((tid + kBS * u) / kTile, b - 1);
            bv[u] *= Jk[p];
        }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = tid + kBS * u, i = q % kTile, p = q / kTile;
        A[p][i] = (p < b && i < nr) ? av[u] : T(0);
        B[p][i] = (p < b && i < nc) ? bv[u] : T(0);
    }
    __syncthreads();
    v4 acc[2][2];
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) acc[x][y] = v4{T(0), T(0), T(0), T(0)};
    for (int p = 0; p < b4; p += 4) {
        // transposed product: MFMA "A" operand = B (columns j), "B" operand = A (rows i)
        T aj[2], bi[2];
        for (int y = 0; y < 2; ++y) aj[y] = B[p + l4][wj + 16 * y + l16];
        for (int x = 0; x < 2; ++x) bi[x] = A[p + l4][wi + 16 * x + l16];
        for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) acc[x][y] = Mfma<T>::mac(aj[y], bi[x], acc[x][y]);
    }
    const bool next = factor_next && r0 == c0 && c0 == k + b && c0 < ns;   // uniform
    if (next) __syncthreads();                 // all waves done reading A/B: A becomes the staging tile
    for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y)
            for (int v = 0; v < 4; ++v) {
                const int i = wi + 16 * x + l16;           // row of C within the tile
                const int j = wj + 16 * y + Mfma<T>::drow(l4, v);   // column of C within the tile
                if (i < nr && j < nc && r0 + i >= c0 + j) {
                    const T nv = cpre[x][y][v] - acc[x][y][v];
                    if (next)
                        A[j][i] = nv;                  // A[col][row] of the next diagonal block (factored below)
                    else
                        Cs[(r0 + i - off) + ld * (c0 + j - off)] = nv;
                }
            }
    if (!next) return;
    __syncthreads();
    const int bn = nr;                         // = min(kPanel, ns - c0)
    T a[16];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
        const int c = 16 * wave + cc;
        a[cc] = (c <= lane && lane < bn) ? A[c][lane] : (c == lane ? T(1) : T(0));
    }
    const int bad = potrf_dinv<T, SG>(a, reinterpret_cast<char *>(&B[0][0]), lane, wave, bn, L + lofs[s] + c0 + ms * c0,
                                      ms, Linv + iofs[s] + (int64_t)kPanel * c0,
                                      SG ? J + blockIdx.y * vs.col + col0[s] + c0 : nullptr);
    if (bad < bn && tid == 0) report_min(status, blockIdx.y * vs.col + col0[s] + c0 + bad, JFEM_ERR_NOT_SPD);
}

// Update matrix of a front, after its last panel step, pushed into the parent's
// front (PushTask): the tile (r0, c0) of U_s (rows and columns >= ncol) becomes
// U_s - L[r0.., 0:ncol] L[c0.., 0:ncol]' (lower triangle of the tile) with the
// whole rank ncol at once, and each entry is added to the parent's front at the
// relmap positions of its row and column (a pivot column of the parent lands in
// its L, the rest in its U): the update matrix is never stored, and no
// extend-add pass reads it back.  U_s itself holds what s's own children pushed.
// Chunks of kPanel columns are staged through LDS; the next chunk's loads are
// issued before the current chunk's products; the parent's entries are loaded
// with the tile's, so their latency hides behind the products.  MFMA operand
// layout and tile split over the four wavefronts as in update_kernel; the
// products accumulate over the chunks in order, and each parent entry receives
// t = U_s - product, rounded, then parent + t: results are reproducible bit for
// bit.  The launch holds at most one child of each parent, so no two tiles add
// into the same parent entry.
template <class T, bool SG = false>
__global__ void __launch_bounds__(kBS) schur_kernel(const PushTask *__restrict__ tasks, const int32_t *__restrict__ relmap,
                                                    T *__restrict__ L, T *__restrict__ U, VarStride vs,
                                                    const int64_t *__restrict__ col0, const T *__restrict__ J) {
    typedef typename Mfma<T>::v4 v4;
    L += blockIdx.y * vs.L;
    U += blockIdx.y * vs.U;
    __shared__ T A[kPanel][kTile + 1];   // A[p][i] = L[r0 + i, k + p]
    __shared__ T B[kPanel][kTile + 1];   // B[p][j] = L[c0 + j, k + p]
    const PushTask t = tasks[blockIdx.x];
    const int64_t ns = t.ns, ms = t.ms, nsp = t.nsp, msp = t.msp;
    const int64_t r0 = t.r0, c0 = t.c0;
    const int nr = (int)min<int64_t>(kTile, ms - r0), nc = (int)min<int64_t>(kTile, ms - c0);
    const T *Ls = L + t.lofs;
    const T *Cs = U + t.uofs;
    const int32_t *rel = relmap + t.relofs - ns;       // rel[front row >= ns] = parent front row
    T *Lp = L + t.lofsp, *Up = U + t.uofsp;
    const int64_t ld = ms - ns, ldp = msp - nsp;
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int l16 = lane & 15, l4 = lane >> 4;
    // this lane's rows (x) and columns (y, v) of the tile, clamped, and their parent positions
    int32_t prow[2], pcol[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x) prow[x] = rel[r0 + min(wi + 16 * x + l16, nr - 1)];
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int v = 0; v < 4; ++v) pcol[y][v] = rel[c0 + min(wj + 16 * y + Mfma<T>::drow(l4, v), nc - 1)];
    auto pidx = [&](int32_t pr, int32_t pc) -> T * {
        return pc < nsp ? Lp + pr + msp * (int64_t)pc : Up + (pr - nsp) + ldp * (int64_t)(pc - nsp);
    };
    T cpre[2][2][4], ppre[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = min(wi + 16 * x + l16, nr - 1), j = min(wj + 16 * y + Mfma<T>::drow(l4, v), nc - 1);
                cpre[x][y][v] = Cs[(r0 + i - ns) + ld * (c0 + j - ns)];
                // (a clamped entry above the diagonal reads the mirrored parent entry; it is not stored)
                const int32_t pr = max(prow[x], pcol[y][v]), pc = min(prow[x], pcol[y][v]);
                ppre[x][y][v] = *pidx(pr, pc);
            }
    T av[16], bv[16];
    const T *Js = SG ? J + blockIdx.y * vs.col + col0[t.s] : nullptr;
    auto load = [&](int64_t k) {
        const int b = (int)min<int64_t>(kPanel, ns - k);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = tid + kBS * u, i = q % kTile, p = min(q / kTile, b - 1);
            av[u] = Ls[(r0 + min(i, nr - 1)) + ms * (k + p)];
            bv[u] = Ls[(c0 + min(i, nc - 1)) + ms * (k + p)];
            if (SG) bv[u] *= Js[k + p];
        }
    };
    v4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = v4{T(0), T(0), T(0), T(0)};
    load(0);
    for (int64_t k = 0; k < ns; k += kPanel) {
        const int b = (int)min<int64_t>(kPanel, ns - k);
        const int b4 = (b + 3) & ~3;
        if (k) __syncthreads();                  // the previous chunk's products are done with A and B
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = tid + kBS * u, i = q % kTile, p = q / kTile;
            A[p][i] = (p < b && i < nr) ? av[u] : T(0);
            B[p][i] = (p < b && i < nc) ? bv[u] : T(0);
        }
        __syncthreads();
        if (k + kPanel < ns) load(k + kPanel);  // in flight during the products below
        for (int p = 0; p < b4; p += 4) {
            T aj[2], bi[2];
#pragma unroll
            for (int y = 0; y < 2; ++y) aj[y] = B[p + l4][wj + 16 * y + l16];
#pragma unroll
            for (int x = 0; x < 2; ++x) bi[x] = A[p + l4][wi + 16 * x + l16];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = Mfma<T>::mac(aj[y], bi[x], acc[x][y]);
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wi + 16 * x + l16, j = wj + 16 * y + Mfma<T>::drow(l4, v);
                if (i < nr && j < nc && r0 + i >= c0 + j) {
                    const T tv = cpre[x][y][v] - acc[x][y][v];
                    *pidx(prow[x], pcol[y][v]) = ppre[x][y][v] + tv;
                }
            }
}

// Off-diagonal blocks of X = L_11^{-1} for the supernodes with two or more
// panels, one workgroup per 64 x 64 output tile (InvTask, after the diagonal
// inverses).  C(r, c) = sum over the chunks k in [k0, k1) of
// sum_kk A[kk][r] B[kk][c], with A[kk][r] = Aop(r, kk) and B[kk][c] = Bop(kk, c)
// staged through LDS chunk by chunk (the next chunk's loads in flight during the
// products), the MFMA operand layout and split over the four wavefronts of
// schur_kernel, the chunks added in order:
//   PH 0: Aop = L[rows of panel i, columns of panel k], Bop = X[rows of panel k,
//         columns of panel j]; T = C stored transposed in the upper block
//         (rows of panel j, columns of panel i) of X;
//   PH 1: Aop = X[rows of panel i, columns of panel k], Bop = T of block (k, j)
//         (read back from the upper block); X[rows of panel i, columns of panel j] = -C.
// Panel i may be the last, partial one (nr < 64 rows), and so may chunk k = i in
// PH 1; entries beyond ns are masked (their indices clamped into the matrix).
template <class T, int PH>
__global__ void __launch_bounds__(kBS) inv_tile_kernel(const InvTask *__restrict__ tasks, const T *__restrict__ L,
                                                       T *__restrict__ Linv, VarStride vs) {
    typedef typename Mfma<T>::v4 v4;
    L += blockIdx.y * vs.L;
    Linv += blockIdx.y * vs.inv;
    __shared__ T A[kPanel][kTile + 1];
    __shared__ T B[kPanel][kTile + 1];
    const InvTask t = tasks[blockIdx.x];
    const int64_t ns = t.ns, ms = t.ms;
    T *X = Linv + t.xofs;
    const T *Ls = L + t.lofs;
    const int64_t r0 = (int64_t)kPanel * t.i, c0 = (int64_t)kPanel * t.j;
    const int nr = (int)min<int64_t>(kPanel, ns - r0);
    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    const int l16 = lane & 15, l4 = lane >> 4;
    T av[16], bv[16];
    // thread q of pass u: x = q % 64 (contiguous in memory), y = q / 64
    auto load = [&](int k) {
        const int64_t kc = (int64_t)kPanel * k;
        const int kb = (int)min<int64_t>(kPanel, ns - kc);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = tid + kBS * u, x = q % kTile, y = q / kTile;
            const int64_t ra = r0 + min(x, nr - 1), ka = kc + min(y, kb - 1);
            av[u] = PH == 0 ? Ls[ra + ms * ka] : X[ra + ns * ka];
            bv[u] = PH == 0 ? X[(kc + min(x, kb - 1)) + ns * (c0 + y)] : X[(c0 + x) + ns * ka];
        }
    };
    v4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = v4{T(0), T(0), T(0), T(0)};
    load(t.k0);
    for (int k = t.k0; k < t.k1; ++k) {
        const int kb = (int)min<int64_t>(kPanel, ns - (int64_t)kPanel * k);
        const int b4 = (kb + 3) & ~3;
        if (k > t.k0) __syncthreads();           // the previous chunk's products are done with A and B
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = tid + kBS * u, x = q % kTile, y = q / kTile;
            A[y][x] = (x < nr && y < kb) ? av[u] : T(0);
            if (PH == 0) B[x][y] = x < kb ? bv[u] : T(0);
            else B[y][x] = y < kb ? bv[u] : T(0);
        }
        __syncthreads();
        if (k + 1 < t.k1) load(k + 1);           // in flight during the products below
        for (int p = 0; p < b4; p += 4) {
            T aj[2], bi[2];
#pragma unroll
            for (int y = 0; y < 2; ++y) aj[y] = B[p + l4][wj + 16 * y + l16];
#pragma unroll
            for (int x = 0; x < 2; ++x) bi[x] = A[p + l4][wi + 16 * x + l16];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = Mfma<T>::mac(aj[y], bi[x], acc[x][y]);
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wi + 16 * x + l16, j = wj + 16 * y + Mfma<T>::drow(l4, v);
                if (i < nr) {
                    if (PH == 0) X[(c0 + j) + ns * (r0 + i)] = acc[x][y][v];
                    else X[(r0 + i) + ns * (c0 + j)] = -acc[x][y][v];
                }
            }
}

// ---------------------------------------------------------------- dataflow solves
// One launch per direction over a task list in dependency order (children before
// parents, panels in solve order); a task waits only on tasks listed before it.
// A workgroup does not take the task of its own index: its first lane draws a
// ticket from a per-launch counter (zeroed before the launch) and the workgroup
// takes the task of that ticket.  Tickets are drawn by running workgroups only,
// so every task listed before a running one is running or finished, whatever
// order the hardware dispatches in and whatever else is resident on the device
// (other contexts' solve launches included): the oldest unfinished task's
// predecessors are all finished, so it always progresses.
// Vectors exchanged between tasks of one launch (W, Z in the forward solve, X in
// the backward solve) are stored and loaded with device-scope accesses (sc1),
// coherent across the XCDs' L2 caches, without the L2 write-back and invalidate
// that plain stores and loads behind a fence would need (each of those walks the
// whole L2 of its XCD, and they serialise).  No counters: W, Z and X are filled
// with a signalling-NaN mark before the solve (arithmetic never produces one),
// each entry is written once, and a consumer polls the entries it needs until
// none holds the mark -- one memory round trip per dependency, and a producer
// never waits for its stores.
// Waits are bounded (kWaitTicks of the 100 MHz wall clock): on expiry the abort
// flag is raised, every waiting workgroup leaves, and the host reports an error.
// cnt: nsn per-supernode words (unused by the polls), then the abort flag, then
// the forward and backward launch tickets.
// Every sum has a fixed order -- children in child order, panels in solve order,
// the four waves' partial sums in wave order -- whatever the timing, so results
// are bitwise reproducible.
// Diagonal blocks are applied through their inverses (computed once per factor by
// diag_inverse_kernel), so a step is a matrix-vector product and not a
// 64-step recurrence.
constexpr long long kWaitTicks = 100000000;

template <class T>
__device__ __forceinline__ T ld_dev(const T *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <class T>
__device__ __forceinline__ void st_dev(T *p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <class T>
__host__ __device__ inline T pending_mark();
template <>
__host__ __device__ inline float pending_mark<float>() { return __builtin_bit_cast(float, 0x7FBADBADu); }
template <>
__host__ __device__ inline double pending_mark<double>() { return __builtin_bit_cast(double, 0x7FF4DEADBEEFDEADull); }
__device__ __forceinline__ bool is_pending(float v) { return __builtin_bit_cast(uint32_t, v) == 0x7FBADBADu; }
__device__ __forceinline__ bool is_pending(double v) { return __builtin_bit_cast(uint64_t, v) == 0x7FF4DEADBEEFDEADull; }

// every memory operation of the wavefront issued so far has completed
__device__ __forceinline__ void complete_stores() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0);
    asm volatile("" ::: "memory");
}

// the workgroup's task slot: one ticket from *ticket per workgroup, broadcast
// through LDS (the caller's first barrier)
__device__ __forceinline__ int64_t draw_ticket(int32_t *ticket, int32_t *slot) {
    if (threadIdx.x == 0) *slot = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return (int64_t)__builtin_amdgcn_readfirstlane(*slot);
}

// re-load the entries of a lane until each holds a published value, all
// pending entries per poll: one memory round trip per poll, not one per entry
template <class T, int N>
__device__ bool await_values(T (&v)[N], const T *const (&p)[N], const bool (&live)[N], int32_t *abort_flag) {
    bool pend = false;
#pragma unroll
    for (int g = 0; g < N; ++g) pend = pend || (live[g] && is_pending(v[g]));
    if (!pend) return true;
    const long long t0 = wall_clock64();
    int spins = 0;
    do {
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (int g = 0; g < N; ++g)
            if (live[g] && is_pending(v[g])) v[g] = ld_dev(p[g]);
        if ((++spins & 15) == 0) {
            if (ld_dev(abort_flag)) return false;
            if (wall_clock64() - t0 > kWaitTicks) {
                st_dev(abort_flag, 1);
                return false;
            }
        }
        pend = false;
#pragma unroll
        for (int g = 0; g < N; ++g) pend = pend || (live[g] && is_pending(v[g]));
    } while (pend);
    return true;
}

// thread (lane = row i, wave w) loads columns [16w, 16w+16) of its row of a
// rows x cols block (column-major, leading dimension ld); indices are clamped
// into the block, the caller masks what lies outside
template <class T>
__device__ __forceinline__ void load_cols16(T (&v)[16], const T *base, int64_t ld, int rows, int cols, int lane, int w) {
    const int64_t i = min(lane, rows - 1);
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = base[i + ld * min(16 * w + c, cols - 1)];
}

// ---- 64 x 64 block times 64 x NR vector block, spread over a workgroup ("g16")
// Wave w takes rows [16w, 16w+16) of the block: lane l has row i = 16w + (l & 15)
// and column group cg = l >> 4, i.e. columns [16cg, 16cg+16), whose 16 entries
// of row i it holds in m[16].  The vector block x (64 columns x NR right-hand
// sides) is held as xr[NR]: in the 16 lanes of column group cg, lane j of xr[g]
// has value q = 16g + j of the group's 16 x NR values, x[16cg + q / NR][q % NR].
// A product term m[c] * x[16cg + c][r] reads its x value by a DPP row broadcast
// (row_newbcast: lane k of each 16-lane row to the whole row), so the values a
// lane needs come from registers of its own row group and not from LDS.  The
// four column groups' partial sums are then added across lanes (two swaps), and
// every lane of a row ends with the same sums.
// Vectors hold nrhs interleaved right-hand sides (stride nrhs); the kernels work
// on NR >= nrhs of them (4, 8 or 16), the padding being zero.
//
// acc += m * x{lane K of the row}.  The DPP source must not have been written by
// a VALU instruction in the two preceding cycles, nor EXEC by a VALU instruction
// in the five preceding: the first product of a call carries s_nop 4 (NOP = 2),
// the first product of every other xr register s_nop 1 (NOP = 1).
template <int K, int NOP>
__device__ __forceinline__ void fmac_bcast(float &acc, float x, float m) {
    if constexpr (NOP == 2)
        asm volatile("s_nop 4\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(x), "v"(m), "i"(K));
    else if constexpr (NOP == 1)
        asm volatile("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(x), "v"(m), "i"(K));
    else
        asm volatile("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(x), "v"(m), "i"(K));
}
template <int K, int NOP>
__device__ __forceinline__ void fmac_bcast(double &acc, double x, double m) {
    if constexpr (NOP == 2)
        asm volatile("s_nop 4\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(x), "v"(m), "i"(K));
    else if constexpr (NOP == 1)
        asm volatile("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(x), "v"(m), "i"(K));
    else
        asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                     : "+v"(acc) : "v"(x), "v"(m), "i"(K));
}
// term Q = NR c + r: acc[r] += m[c] * xr[Q / 16]{lane Q % 16}; per acc[r] the
// columns are added in increasing order
template <class T, int NR, int... Q>
__device__ __forceinline__ void g16_mac_terms(T (&acc)[NR], const T (&m)[16], const T (&xr)[NR],
                                              std::integer_sequence<int, Q...>) {
    (fmac_bcast<Q % 16, Q == 0 ? 2 : (Q % 16 == 0 ? 1 : 0)>(acc[Q % NR], xr[Q / 16], m[Q / NR]), ...);
}
template <class T, int NR>
__device__ __forceinline__ void g16_mac(T (&acc)[NR], const T (&m)[16], const T (&xr)[NR]) {
    g16_mac_terms<T, NR>(acc, m, xr, std::make_integer_sequence<int, 16 * NR>{});
}
// sum over the four column groups: (p0 + p2) + (p1 + p3), the same bits in all four
template <class T, int NR>
__device__ __forceinline__ void g16_reduce(T (&acc)[NR]) {
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] += __shfl_xor(acc[r], 32);
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] += __shfl_xor(acc[r], 16);
}
// m[c] = B[i + ld (16cg + c)] of a rows x cols block, indices clamped into it
// (the vector block is zero beyond cols, so clamped columns add nothing)
template <class T>
__device__ __forceinline__ void g16_load(T (&m)[16], const T *B, int64_t ld, int rows, int cols, int i, int cg) {
    const int64_t ic = min(i, rows - 1);
#pragma unroll
    for (int c = 0; c < 16; ++c) m[c] = B[ic + ld * min(16 * cg + c, cols - 1)];
}
// transposed: m[k] = B[(16cg + k) + ld i], i.e. row i of B^T
template <class T>
__device__ __forceinline__ void g16_load_t(T (&m)[16], const T *B, int64_t ld, int rows, int cols, int i, int cg) {
    const int64_t ic = min(i, cols - 1);
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = B[min(16 * cg + k, rows - 1) + ld * ic];
}
// (column, rhs) of the value lane (cg, j) holds in xr[g]
template <int NR>
__device__ __forceinline__ int g16_col(int g, int cg, int j) { return 16 * cg + (16 * g + j) / NR; }
template <int NR>
__device__ __forceinline__ int g16_rhs(int g, int j) { return (16 * g + j) % NR; }
// xr from a vector block in LDS (V[c][r]), zero beyond cnt columns
template <class T, int NR>
__device__ __forceinline__ void g16_from_lds(T (&xr)[NR], const T (*V)[NR], int cnt, int cg, int j) {
#pragma unroll
    for (int g = 0; g < NR; ++g) {
        const int c = g16_col<NR>(g, cg, j), r = g16_rhs<NR>(g, j);
        xr[g] = c < cnt ? V[c][r] : T(0);
    }
}
// xr from device memory once published by another task: column c at src_c(c)
// (stride-1 right-hand sides); zero beyond cnt columns or nrhs; false on
// abort / time-out
template <class T, int NR, class Addr>
__device__ __forceinline__ bool g16_load_wait(T (&xr)[NR], Addr src_c, int cnt, int nrhs, int cg, int j,
                                              int32_t *abort_flag) {
    const T *p[NR];
    bool live[NR];
#pragma unroll
    for (int g = 0; g < NR; ++g) {
        const int c = g16_col<NR>(g, cg, j), r = g16_rhs<NR>(g, j);
        p[g] = src_c(min(c, cnt - 1)) + min(r, nrhs - 1);
        live[g] = c < cnt && r < nrhs;
        xr[g] = ld_dev(p[g]);
    }
    const bool ok = await_values<T, NR>(xr, p, live, abort_flag);
#pragma unroll
    for (int g = 0; g < NR; ++g) xr[g] = live[g] ? xr[g] : T(0);
    return ok;
}
// the front row indices behind this lane's xr entries for a tile of cnt update
// rows starting at rows[0] (indices clamped into the tile); loaded ahead so that
// the awaited X loads do not wait for them
template <int NR>
__device__ __forceinline__ void g16_rows(int32_t (&ridx)[NR], const int32_t *rows, int cnt, int cg, int j) {
#pragma unroll
    for (int g = 0; g < NR; ++g) ridx[g] = rows[min(g16_col<NR>(g, cg, j), cnt - 1)];
}
// g16_load_wait for values X[ridx[g]] (stride-1 right-hand sides)
template <class T, int NR>
__device__ __forceinline__ bool g16_load_wait_rows(T (&xr)[NR], const T *X, const int32_t (&ridx)[NR], int cnt, int nrhs,
                                                   int cg, int j, int32_t *abort_flag) {
    const T *p[NR];
    bool live[NR];
#pragma unroll
    for (int g = 0; g < NR; ++g) {
        const int c = g16_col<NR>(g, cg, j), r = g16_rhs<NR>(g, j);
        p[g] = X + (int64_t)ridx[g] * nrhs + min(r, nrhs - 1);
        live[g] = c < cnt && r < nrhs;
        xr[g] = ld_dev(p[g]);
    }
    const bool ok = await_values<T, NR>(xr, p, live, abort_flag);
#pragma unroll
    for (int g = 0; g < NR; ++g) xr[g] = live[g] ? xr[g] : T(0);
    return ok;
}

// this lane's share of a row's NR sums when storing: right-hand sides
// [cg NR/4, (cg+1) NR/4)
template <int NR>
__device__ __forceinline__ int g16_r0(int cg) { return cg * (NR / 4); }

// acc[r] += P_u[i][r] for the npart partial sums u = 0, 1, ... (kTile rows x nrhs
// each, consecutive from P) of output row i, in that order; called after the
// reduction, so every lane of a row adds the same values.  The loads of several
// partials go out together (one round trip per batch).  False on abort / time-out.
template <class T, int NR>
__device__ __forceinline__ bool add_partials(T (&acc)[NR], const T *P, int npart, int i, int nr, int nrhs,
                                             int32_t *abort_flag) {
    constexpr int PB = NR >= 16 ? 1 : 16 / NR;
    bool ok = true;
    for (int u0 = 0; u0 < npart; u0 += PB) {
        T v[PB * NR];
        const T *pp[PB * NR];
        bool live[PB * NR];
#pragma unroll
        for (int u = 0; u < PB; ++u)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const int q = u * NR + r;
                pp[q] = P + ((int64_t)min(u0 + u, npart - 1) * kTile + min(i, nr - 1)) * nrhs + min(r, nrhs - 1);
                live[q] = u0 + u < npart && i < nr && r < nrhs;
                v[q] = ld_dev(pp[q]);
            }
        ok = await_values<T, PB * NR>(v, pp, live, abort_flag) && ok;
#pragma unroll
        for (int u = 0; u < PB; ++u)
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (u0 + u < npart) acc[r] += live[u * NR + r] ? v[u * NR + r] : T(0);
    }
    return ok;
}
// a partial sum task's result: rows i < nr of this lane's share of right-hand
// sides into slot P (kTile rows x nrhs)
template <class T, int NR>
__device__ __forceinline__ void store_partial(T *P, const T (&acc)[NR], int i, int nr, int nrhs, int rq) {
    if (i < nr) {
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) {
            const int r = rq + rr;
            if (r < nrhs) st_dev(P + (int64_t)i * nrhs + r, acc[r]);
        }
    }
}

// Gather of the children's update-vector entries landing in rows [r0, r0+nrows)
// of a front into vs (LDS), nt threads, NR slots per row (right-hand sides
// beyond nrhs are skipped): the (index, value) loads of 8 entries per thread go
// out together (a loop of single dependent loads would wait a memory latency per
// entry); children in child order.
// The update vectors hold the pending mark until their task publishes them, so
// the loads poll (all pending values of a thread per round trip) instead of a
// counter wait followed by the loads.  False on abort / time-out.
// first: the task's first child record, loaded by the caller at the start of the
// task; each next record is loaded while the current child is gathered.
template <class T, int NR>
__device__ __forceinline__ bool gather_children(T (*vs)[NR], const FlowChild *__restrict__ gch, int64_t gb, int nch,
                                                FlowChild first, const int32_t *__restrict__ relmap, const T *W,
                                                int r0, int nrhs, int tid, int nt, int32_t *abort_flag) {
    bool ok = true;
    FlowChild ch = first;
    for (int j = 0; j < nch; ++j) {
        const FlowChild nxt = gch[gb + min(j + 1, nch - 1)];
        const int nq = (ch.i1 - ch.i0) * NR;
        const int32_t *rel = relmap + ch.relofs + ch.i0;
        const T *wc = W + (ch.relofs + ch.i0) * nrhs;
        for (int q0 = 0; q0 < nq; q0 += 8 * nt) {
            int32_t row[8];
            T v[8];
            const T *pp[8];
            bool live[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = min(q0 + nt * u + tid, nq - 1), i = q / NR, r = min(q % NR, nrhs - 1);
                row[u] = rel[i];
                pp[u] = wc + i * nrhs + r;
                live[u] = q0 + nt * u + tid < nq && (q0 + nt * u + tid) % NR < nrhs;
                v[u] = ld_dev(pp[u]);
            }
            ok = await_values<T, 8>(v, pp, live, abort_flag) && ok;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = q0 + nt * u + tid, r = q % NR;
                if (q < nq && r < nrhs) vs[row[u] - r0][r] += v[u];
            }
        }
        __syncthreads();
        ch = nxt;
    }
    return ok;
}

// forward: pivot task (s, p) of a supernode with one panel: z = L_11^{-1}
// (y + children's contributions); of a supernode with np >= 2 panels: v_p = y_p +
// children's contributions to the rows of panel p is published to V, then
// z_p = sum_{k<p} X[p,k] v_k + X_pp v_p with X = L_11^{-1} and the v_k of the
// panels before it polled from V -- the panels of s run side by side, with one
// exchange of v and no chain along the front; update task: w = children's
// contributions - sum_k L[rows, k] z_k, written to s's update vector in W;
// whole-front task: both for a supernode with one panel (rows of the front
// <= kWholeRows), the update rows in blocks of 64.
// The k loops run without LDS or barriers: every lane polls the values it needs
// straight into its registers.  The k loop covers [k0, k1); the partial sums of
// the blocks before k0 (kind 2 tasks) are added after the reduction.
template <class T, int NR>
__device__ __forceinline__ void fwd_main_task(const FlowTask &t, T (*vs)[NR], T (*vx)[NR], int &failed,
                                              const FlowChild *__restrict__ gch, const int32_t *__restrict__ relmap,
                                              const T *__restrict__ L, const T *__restrict__ Linv,
                                              const T *__restrict__ Y, T *W, T *Z, T *V, T *P, int nrhs,
                                              int32_t *cnt, int64_t nsn) {
    int32_t *abort_flag = cnt + nsn;
    const FlowChild first = t.nch > 0 ? gch[t.gb] : FlowChild{};   // (in flight with the task's first loads)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, j = lane & 15, cg = lane >> 4;
    const int i = 16 * w + j;                          // row of the 64-row block
    const int rq = g16_r0<NR>(cg);
    const bool pivot = t.p >= 0, whole = t.kind == 1, part = t.kind == 2;
    const bool multi = pivot && t.xofs >= 0;           // np >= 2: through X and the v of the earlier panels
    const int k0 = t.k0, k1 = t.k1;
    // operand of the k loop: update task L[rows, k] (ld ms), multi-panel pivot X[p rows, k] (ld ns)
    const T *Kb = multi ? Linv + t.xofs + t.r0 : L + t.lofs + t.r0;
    const int64_t kld = multi ? t.ns : t.ms;
    const T *Kv = multi ? V : Z;                       // the vector the k loop polls
    T lm[16], im[16];
    if (k1 > k0) {
        const int64_t col = (int64_t)kPanel * k0;
        g16_load(lm, Kb + kld * col, kld, t.nr, (int)min<int64_t>(kPanel, t.ns - col), i, cg);
    } else if (whole && t.ms > t.ns) {
        g16_load(lm, L + t.lofs + t.ns, t.ms, (int)(t.ms - t.ns), t.nr, i, cg);
    }
    T acc[NR], xr[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = T(0);
    auto kloop = [&]() -> bool {
        bool ok = true;
        for (int k = k0; k < k1; ++k) {
            const int64_t col = (int64_t)kPanel * k;
            const T *zk = Kv + (t.c0 + col) * nrhs;
            ok = g16_load_wait<T, NR>(xr, [&](int c) { return zk + (int64_t)c * nrhs; },
                                      (int)min<int64_t>(kPanel, t.ns - col), nrhs, cg, j, abort_flag) && ok;
            T cur[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) cur[c] = lm[c];
            if (k + 1 < k1)
                g16_load(lm, Kb + kld * (col + kPanel), kld, t.nr, (int)min<int64_t>(kPanel, t.ns - col - kPanel), i, cg);
            g16_mac<T, NR>(acc, cur, xr);
        }
        return ok;
    };
    if (part) {                                        // partial sum over [k0, k1) into slot pofs
        const bool ok = kloop();
        g16_reduce<T, NR>(acc);
        if (ok) store_partial<T, NR>(P + (int64_t)t.pofs * kTile * nrhs, acc, i, t.nr, nrhs, rq);
        return;
    }
    const T *Pc = P + (int64_t)t.pofs * kTile * nrhs;  // the partial sums of [0, k0)
    if (pivot) g16_load(im, Linv + t.iofs, t.nr, t.nr, t.nr, i, cg);
#pragma unroll
    for (int q2 = 0; q2 < kWholeRows * NR / kBS; ++q2) {
        const int q = tid + kBS * q2, row = q / NR, r = q % NR;
        const int64_t yi = pivot ? t.c0 + t.r0 + min(row, t.nr - 1) : 0;
        const T y = Y[yi * nrhs + min(r, nrhs - 1)];
        vs[row][r] = (pivot && row < t.nr && r < nrhs) ? y : T(0);
    }
    if (tid == 0) failed = 0;
    __syncthreads();
    if (t.need > 0 && !gather_children<T, NR>(vs, gch, t.gb, t.nch, first, relmap, W, t.r0, nrhs, tid, kBS, abort_flag))
        failed = 1;
    __syncthreads();
    if (failed) return;
    if (multi && i < t.nr) {                           // v_p for the later panels of s
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) {
            const int r = rq + rr;
            if (r < nrhs) st_dev(V + (t.c0 + t.r0 + i) * nrhs + r, vs[i][r]);
        }
    }
    if (whole) {
        // z = L_11^{-1} v[0:ns], then w = v - L[update rows, :] z in blocks of 64 rows
        g16_from_lds<T, NR>(xr, vs, t.nr, cg, j);      // update rows of the front are not part of v
        g16_mac<T, NR>(acc, im, xr);
        g16_reduce<T, NR>(acc);
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) {
            const int r = rq + rr;
            vx[i][r] = i < t.nr ? acc[r] : T(0);
            if (i < t.nr && r < nrhs) st_dev(Z + (t.c0 + i) * nrhs + r, acc[r]);
        }
        __syncthreads();
        g16_from_lds<T, NR>(xr, vx, t.nr, cg, j);
        const int64_t nb = t.ms - t.ns;                 // block 0 was loaded at the start
        for (int64_t b0 = 0; b0 < nb; b0 += kTile) {
            T cur[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) cur[c] = lm[c];
            if (b0 + kTile < nb) g16_load(lm, L + t.lofs + t.ns + b0 + kTile, t.ms, (int)(nb - b0 - kTile), t.nr, i, cg);
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = T(0);
            g16_mac<T, NR>(acc, cur, xr);
            g16_reduce<T, NR>(acc);
            const int64_t row = t.ns + b0 + i;         // front row
            if (row < t.ms) {
#pragma unroll
                for (int rr = 0; rr < NR / 4; ++rr) {
                    const int r = rq + rr;
                    if (r < nrhs) st_dev(W + (t.relofs + row - t.ns) * nrhs + r, vs[row][r] - acc[r]);
                }
            }
        }
        return;
    }
    bool ok = kloop();
    if (!ok) failed = 1;
    if (multi) {
        // + X_pp v_p (v_p from LDS), one reduction for the whole sum, then the partials
        g16_from_lds<T, NR>(xr, vs, t.nr, cg, j);
        g16_mac<T, NR>(acc, im, xr);
        g16_reduce<T, NR>(acc);
        if (t.npart > 0) ok = add_partials<T, NR>(acc, Pc, t.npart, i, t.nr, nrhs, abort_flag) && ok;
        if (i < t.nr && ok) {
#pragma unroll
            for (int rr = 0; rr < NR / 4; ++rr) {
                const int r = rq + rr;
                if (r < nrhs) st_dev(Z + (t.c0 + t.r0 + i) * nrhs + r, acc[r]);
            }
        }
        return;
    }
    g16_reduce<T, NR>(acc);
    if (!pivot) {
        if (t.npart > 0 && !add_partials<T, NR>(acc, Pc, t.npart, i, t.nr, nrhs, abort_flag)) failed = 1;
        if (i < t.nr) {
#pragma unroll
            for (int rr = 0; rr < NR / 4; ++rr) {
                const int r = rq + rr;
                if (r < nrhs && !failed) st_dev(W + (t.relofs + t.r0 - t.ns + i) * nrhs + r, vs[i][r] - acc[r]);
            }
        }
        return;
    }
    // one-panel supernode (kend = 0): z = L_11^{-1} v
    g16_from_lds<T, NR>(xr, vs, t.nr, cg, j);
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = T(0);
    g16_mac<T, NR>(acc, im, xr);
    g16_reduce<T, NR>(acc);
    if (i < t.nr) {
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) {
            const int r = rq + rr;
            if (r < nrhs) st_dev(Z + (t.c0 + t.r0 + i) * nrhs + r, acc[r]);
        }
    }
}

// backward: task (s, p): t_p = z_p - sum over the update rows i of the front of
// L[i, p]' x_i (ancestors' values), tile by tile from the last -- the order in
// which the ancestors' panels publish them; then x_p = L_pp^{-T} t_p for a
// supernode with one panel, and for np >= 2 panels t_p is published to Tb and
// x_p = sum_{q>p} X[q,p]' t_q + X_pp' t_p (from the last panel, X = L_11^{-1}),
// the t_q of the later panels polled from Tb: one exchange, no chain.
// The products are transposed (lane row = panel column): each tile is loaded
// coalesced, staged through LDS (double-buffered in FP32, so one barrier per
// tile), and read back transposed before its values are awaited.
// Long sums are split over several tasks: kind 2 sums L[rows, p]' x_rows over
// update rows [u0, u1), or X[q,p]' t_q over panels q in [k0, k1), into a partial
// slot; a panel task adds the partials of its far update rows before forming
// t_p; a panel with many later panels has kind 3 form and publish t_p and kind 4
// form x_p from X_pp' t_p (t_p polled from Tb), its own q in [k0, k1) and the
// partials of the panels above k1.
template <class T>
constexpr int kBwdStage = sizeof(T) == 4 ? 2 : 1;      // staging buffers of the four-wave task (LDS budget)
template <class T, int NR>
__device__ __forceinline__ void bwd_main_task(const FlowTask &t, T (*Lt)[kTile][kTile + 1], T (*vx)[NR], int &failed,
                                              const int32_t *__restrict__ rows, const T *__restrict__ L,
                                              const T *__restrict__ Linv, const T *__restrict__ Z, T *X, T *Tb,
                                              T *P, int nrhs, int32_t *cnt, int64_t nsn) {
    constexpr int NB = kBwdStage<T>;
    int32_t *abort_flag = cnt + nsn;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, j = lane & 15, cg = lane >> 4;
    const int c = 16 * w + j;                          // panel column of this lane's output
    const int rq = g16_r0<NR>(cg);
    const int nc = t.nr;
    const bool multi = t.xofs >= 0;
    const bool tphase = t.kind == 0 || t.kind == 3;    // forms t_p from the update rows
    const bool xout = t.kind == 0 || t.kind == 4;      // forms x_p
    const int nchunk = (t.u1 - t.u0 + kTile - 1) / kTile;   // (none for the X tasks)
    const int ntile = nchunk + (multi ? t.k1 - t.k0 : 0);
    const T *Lc = L + t.lofs + t.ms * t.r0;            // front row 0, column r0
    const T *Xc = multi ? Linv + t.xofs + t.ns * t.r0 : Lc;   // X row 0, column r0
    const int32_t *rw = rows + t.rofs;
    // tile e: update rows [u0 + 64(nchunk-1-e), ...) of L for e < nchunk, else the
    // rows of panel q = k1-1-(e-nchunk) of X
    auto tile_row0 = [&](int e) -> int64_t {
        return e < nchunk ? t.u0 + (int64_t)kTile * (nchunk - 1 - e) : (int64_t)kPanel * (t.k1 - 1 - (e - nchunk));
    };
    auto tile_rows = [&](int e) -> int {
        const int64_t r0 = tile_row0(e);
        return (int)min<int64_t>(kTile, (e < nchunk ? (int64_t)t.u1 : t.ns) - r0);
    };
    auto load_tile = [&](T (&v)[16], int e) {
        if (e < nchunk) load_cols16(v, Lc + tile_row0(e), t.ms, tile_rows(e), nc, lane, w);
        else load_cols16(v, Xc + tile_row0(e), t.ns, tile_rows(e), nc, lane, w);
    };
    T lv[16], im[16], zv[NR / 4], acc[NR], xr[NR];
    if (xout) g16_load_t(im, Linv + t.iofs, nc, nc, nc, c, cg); // row c of L_pp^{-T}
    if (tphase) {
        const T *zp = Z + (t.c0 + t.r0 + min(c, nc - 1)) * nrhs;
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) zv[rr] = zp[min(rq + rr, nrhs - 1)];
    }
    if (tid == 0) failed = 0;
    if (ntile > 0) load_tile(lv, 0);
    int32_t ridx[NR];                                  // row indices of the next update-row tile
    if (nchunk > 0) g16_rows<NR>(ridx, rw + tile_row0(0), tile_rows(0), cg, j);
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = T(0);
    // t_p = z_p - (update-row sum, then the partial sums of the far update rows)
    // into vx, published for the earlier panels of s; the accumulator restarts
    // for the later panels' sum
    bool formed = false;
    bool ok = true;
    auto form_t = [&]() {
        g16_reduce<T, NR>(acc);
        if (t.npart > 0)
            ok = add_partials<T, NR>(acc, P + (int64_t)t.pofs * kTile * nrhs, t.npart, c, nc, nrhs, abort_flag) && ok;
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) {
            const int r = rq + rr;
            const T tv = (c < nc && r < nrhs) ? zv[rr] - acc[r] : T(0);
            vx[c][r] = tv;
            if (multi && c < nc && r < nrhs) st_dev(Tb + (t.c0 + t.r0 + c) * nrhs + r, tv);
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] = T(0);
        formed = true;
    };
    for (int e = 0; e < ntile; ++e) {
        if (tphase && e == nchunk) form_t();
        const int nre = tile_rows(e);
        T (*S)[kTile + 1] = Lt[e % NB];
#pragma unroll
        for (int k = 0; k < 16; ++k) S[lane][16 * w + k] = (lane < nre && 16 * w + k < nc) ? lv[k] : T(0);
        __syncthreads();
        if (e + 1 < ntile) load_tile(lv, e + 1);
        T m[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = S[16 * cg + k][c];
        if (NB == 1) __syncthreads();                  // the next tile overwrites the only buffer
        if (e < nchunk) {
            ok = g16_load_wait_rows<T, NR>(xr, X, ridx, nre, nrhs, cg, j, abort_flag) && ok;
            const int en = min(e + 1, nchunk - 1);
            g16_rows<NR>(ridx, rw + tile_row0(en), tile_rows(en), cg, j);
        } else {
            const T *tq = Tb + (t.c0 + tile_row0(e)) * nrhs;
            ok = g16_load_wait<T, NR>(xr, [&](int k) { return tq + (int64_t)k * nrhs; }, nre, nrhs, cg, j,
                                      abort_flag) && ok;
        }
        g16_mac<T, NR>(acc, m, xr);
    }
    if (!ok) failed = 1;
    if (tphase && !formed) form_t();                   // no later panels
    if (t.kind == 3) return;
    if (t.kind == 2) {
        g16_reduce<T, NR>(acc);
        if (ok) store_partial<T, NR>(P + (int64_t)t.pofs * kTile * nrhs, acc, c, nc, nrhs, rq);
        return;
    }
    if (t.kind == 4) {                                 // t_p, published by the panel's kind 3 task
        const T *tp = Tb + (t.c0 + t.r0) * nrhs;
        if (!g16_load_wait<T, NR>(xr, [&](int k) { return tp + (int64_t)k * nrhs; }, nc, nrhs, cg, j, abort_flag))
            failed = 1;
    }
    __syncthreads();
    if (failed) return;
    if (t.kind == 0) g16_from_lds<T, NR>(xr, vx, nc, cg, j);
    g16_mac<T, NR>(acc, im, xr);
    g16_reduce<T, NR>(acc);
    if (t.kind == 4 && t.npart > 0 &&
        !add_partials<T, NR>(acc, P + (int64_t)t.pofs * kTile * nrhs, t.npart, c, nc, nrhs, abort_flag))
        return;
    if (c < nc) {
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) {
            const int r = rq + rr;
            if (r < nrhs) st_dev(X + (t.c0 + t.r0 + c) * nrhs + r, acc[r]);
        }
    }
}

// ---- bottom subtrees: one wavefront per task
// The supernodes of the bottom subtrees (one panel, at most kWholeRows front rows,
// all children such) are most of the tasks and little work each; the four-wave
// kernels above hold two or three of them per CU (their register files), and the
// bottom levels wait on memory.  Here a task is one wavefront that works through
// its rows in passes of 16 -- lane (cg, j) holding row j of the pass and columns
// [16cg, 16cg+16), the same products and the same summation order as the
// four-wave kernels, so the results are bitwise the same -- with the loads of
// the next kLeafRing passes in flight while a pass is computed.  The forward
// leaf launch runs before the main one, the backward after it.
// Ordering within the wavefront: LDS instructions of one wavefront execute in
// issue order, so a wave fence suffices between writing LDS and reading what
// other lanes wrote (a workgroup barrier would also wait for the loads in flight).
constexpr int kLeafRing = 4;

// loads at 32-bit element offsets from a wave-uniform base (a supernode's block
// or a diagonal inverse, < 2^16 entries): one VGPR per address and the SGPR-base
// form of the global load, where 64-bit per-lane addresses would take two VGPRs
// and a 64-bit add each
// a wave-uniform pointer or index, held in SGPRs
template <class T>
__device__ __forceinline__ const T *uniform_ptr(const T *p) {
    const uint64_t a = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
    return (const T *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T>
__device__ __forceinline__ T ld_off(const T *base, uint32_t idx) {
    typedef const __attribute__((address_space(1))) char *global_bytes;
    typedef const __attribute__((address_space(1))) T *global_ptr;
    return *(global_ptr)((global_bytes)base + idx * (uint32_t)sizeof(T));
}
// g16_load, g16_load_t and load_cols16 in that form (ld, rows, cols < 2^16)
template <class T>
__device__ __forceinline__ void g16_load32(T (&m)[16], const T *B, int ld, int rows, int cols, int i, int cg) {
    const uint32_t ic = (uint32_t)min(i, rows - 1);
#pragma unroll
    for (int c = 0; c < 16; ++c) m[c] = ld_off(B, ic + (uint32_t)ld * (uint32_t)min(16 * cg + c, cols - 1));
}
template <class T>
__device__ __forceinline__ void g16_load_t32(T (&m)[16], const T *B, int ld, int rows, int cols, int i, int cg) {
    const uint32_t ic = (uint32_t)min(i, cols - 1) * (uint32_t)ld;
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = ld_off(B, (uint32_t)min(16 * cg + k, rows - 1) + ic);
}
template <class T>
__device__ __forceinline__ void load_cols16_32(T (&v)[16], const T *base, int ld, int rows, int cols, int lane, int w) {
    const uint32_t i = (uint32_t)min(lane, rows - 1);
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = ld_off(base, i + (uint32_t)ld * (uint32_t)min(16 * w + c, cols - 1));
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Children's update-vector entries landing in a whole front, one wavefront, U
// entries per lane per round with every load of the round issued before the
// first addition; children in child order, so each row receives y and then its
// children's values in child order, as in gather_children.
template <class T, int NR, int U>
__device__ __forceinline__ bool gather_children_wave(T (*vs)[NR], const FlowChild *__restrict__ gch, int64_t gb, int nch,
                                                     FlowChild first, const int32_t *__restrict__ relmap, const T *W,
                                                     int nrhs, int lane, int32_t *abort_flag) {
    bool ok = true;
    FlowChild ch = first;
    for (int jc = 0; jc < nch; ++jc) {
        const FlowChild nxt = gch[gb + min(jc + 1, nch - 1)];
        const int nq = (ch.i1 - ch.i0) * NR;
        const int32_t *rel = relmap + ch.relofs + ch.i0;
        const T *wc = W + (ch.relofs + ch.i0) * nrhs;
        for (int q0 = 0; q0 < nq; q0 += 64 * U) {
            int32_t row[U];
            T v[U];
            const T *pp[U];
            bool live[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = min(q0 + 64 * u + lane, nq - 1), i = q / NR, r = min(q % NR, nrhs - 1);
                row[u] = rel[i];
                pp[u] = wc + i * nrhs + r;
                live[u] = q0 + 64 * u + lane < nq && (q0 + 64 * u + lane) % NR < nrhs;
                v[u] = ld_dev(pp[u]);
            }
            ok = await_values<T, U>(v, pp, live, abort_flag) && ok;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = q0 + 64 * u + lane, r = q % NR;
                if (q < nq && r < nrhs) vs[row[u]][r] += v[u];
            }
            wave_sync();
        }
        ch = nxt;
    }
    return ok;
}

// forward whole-front task (s one panel, front rows <= kWholeRows): pass q < npz
// is rows [16q, 16q+16) of z = L_11^{-1} v, pass npz + u is update rows
// [ns + 16u, ns + 16u + 16) of w = v - L[update rows, :] z
template <class T, int NR>
__device__ __forceinline__ void fwd_leaf_task(const FlowTask &t, T (*vs)[NR], T (*vx)[NR], int lane,
                                              const FlowChild *__restrict__ gch, const int32_t *__restrict__ relmap,
                                              const T *__restrict__ L, const T *__restrict__ Linv,
                                              const T *__restrict__ Y, T *W, T *Z, int nrhs, int32_t *cnt,
                                              int64_t nsn) {
    int32_t *abort_flag = cnt + nsn;
    const FlowChild first = t.nch > 0 ? gch[t.gb] : FlowChild{};   // (in flight with the task's first loads)
    const int j = lane & 15, cg = lane >> 4;
    const int rq = g16_r0<NR>(cg);
    const int nr = t.nr, nu = (int)(t.ms - t.ns);
    const int npz = (nr + 15) >> 4, npass = npz + ((nu + 15) >> 4);
    const T *Li = Linv + t.iofs, *Lu = L + t.lofs + t.ns;
    // one load path for both kinds of pass (uniform parameters selected first),
    // so that every load keeps the SGPR-base form
    auto load_pass = [&](T (&m)[16], int q) {
        const bool zp = q < npz;
        const int r0 = zp ? 16 * q : 16 * (q - npz);
        const T *base = uniform_ptr(zp ? Li + r0 : Lu + r0);
        g16_load32(m, base, uniform_int(zp ? nr : (int)t.ms), uniform_int((zp ? nr : nu) - r0), nr, j, cg);
    };
    constexpr int QY = kTile * NR / 64;
    T yv[QY];
#pragma unroll
    for (int q2 = 0; q2 < QY; ++q2) {
        const int q = lane + 64 * q2, row = q / NR, r = q % NR;
        yv[q2] = Y[(t.c0 + min(row, nr - 1)) * nrhs + min(r, nrhs - 1)];
    }
    // the ring loads are issued unconditionally (clamped to the last pass), so the
    // compiler's wait counts stay exact: no global access sits under a branch in
    // the pass loop, and the results go out after it
    T buf[kLeafRing][16];
#pragma unroll
    for (int u = 0; u < kLeafRing; ++u) load_pass(buf[u], min(u, npass - 1));
#pragma unroll
    for (int q2 = 0; q2 < QY; ++q2) {
        const int q = lane + 64 * q2, row = q / NR, r = q % NR;
        vs[row][r] = (row < nr && r < nrhs) ? yv[q2] : T(0);
    }
    for (int q = kTile * NR + lane; q < t.ms * NR; q += 64) vs[q / NR][q % NR] = T(0);
    wave_sync();
    if (t.need > 0 && !gather_children_wave<T, NR, 16>(vs, gch, t.gb, t.nch, first, relmap, W, nrhs, lane, abort_flag))
        return;
    T acc[NR], xr[NR];
    g16_from_lds<T, NR>(xr, vs, nr, cg, j);            // update rows of the front are not part of v
    for (int q0 = 0; q0 < npass; q0 += kLeafRing) {
#pragma unroll
        for (int u = 0; u < kLeafRing; ++u) {
            const int q = q0 + u;
            if (q == npz) {                            // z complete
                wave_sync();
                g16_from_lds<T, NR>(xr, vx, nr, cg, j);
            }
            if (q < npass) {
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] = T(0);
                g16_mac<T, NR>(acc, buf[u], xr);
                g16_reduce<T, NR>(acc);
                if (q < npz) {                         // z rows into vx
                    const int row = 16 * q + j;
#pragma unroll
                    for (int rr = 0; rr < NR / 4; ++rr) vx[row][rq + rr] = row < nr ? acc[rq + rr] : T(0);
                } else {                               // w rows in place of v's update rows
                    const int row = (int)t.ns + 16 * (q - npz) + j;
                    if (row < t.ms) {
#pragma unroll
                        for (int rr = 0; rr < NR / 4; ++rr) vs[row][rq + rr] -= acc[rq + rr];
                    }
                }
            }
            load_pass(buf[u], min(q + kLeafRing, npass - 1));
        }
    }
    wave_sync();
    for (int q = lane; q < nr * NR; q += 64) {
        const int row = q / NR, r = q % NR;
        if (r < nrhs) st_dev(Z + (t.c0 + row) * nrhs + r, vx[row][r]);
    }
    for (int q = lane; q < nu * NR; q += 64) {
        const int row = q / NR, r = q % NR;
        if (r < nrhs) st_dev(W + (t.relofs + row) * nrhs + r, vs[t.ns + row][r]);
    }
}

// The forward launch: workgroups [0, nq) take the bottom subtrees' tasks four at
// a time (task 4b + w on wave w, each wave on its own, on its own quarter of the
// LDS), the rest take one task each of the remaining list.  The bottom tasks
// come first in the list and wait only on earlier ones (or on a wave of their
// own workgroup, which is resident), so in-order dispatch always progresses, and
// the upper tasks start as soon as their own subtrees are done.
template <class T, int NR>
constexpr bool kLeafQuads = NR <= 8;
template <class T, int NR>
__global__ void __launch_bounds__(kBS) fwd_flow_kernel(const FlowTask *__restrict__ tasks,
                                                       const FlowChild *__restrict__ gch,
                                                       const int32_t *__restrict__ relmap, const T *__restrict__ L,
                                                       const T *__restrict__ Linv, const T *__restrict__ Y, T *W, T *Z,
                                                       T *V, T *P, int nrhs, int32_t *cnt, int64_t nsn, int64_t nleaf) {
    constexpr int NQ = kLeafQuads<T, NR> ? 4 : 1;
    __shared__ __attribute__((aligned(16))) T lds[NQ][kWholeRows + kTile][NR];
    __shared__ int failed_one;                         // (with quads: in the main path's unused quarter)
    __shared__ int32_t ticket_slot;
    const int64_t b = draw_ticket(cnt + nsn + 1, &ticket_slot);
    const int64_t nq = (nleaf + 3) / 4;
    if constexpr (kLeafQuads<T, NR>) {
        if (b < nq) {
            const int w = uniform_int((int)(threadIdx.x >> 6));
            const int64_t k = 4 * b + w;
            if (k < nleaf)
                fwd_leaf_task<T, NR>(tasks[k], lds[w], lds[w] + kWholeRows, (int)(threadIdx.x & 63), gch, relmap, L,
                                     Linv, Y, W, Z, nrhs, cnt, nsn);
            return;
        }
    }
    const FlowTask t = tasks[nleaf + b - nq];
    int *failed;
    if constexpr (NQ > 1) failed = (int *)&lds[1][0][0];
    else failed = &failed_one;
    fwd_main_task<T, NR>(t, lds[0], lds[0] + kWholeRows, *failed, gch, relmap, L, Linv, Y, W, Z, V, P, nrhs, cnt, nsn);
}

// backward task of a one-panel supernode: x = L_11^{-T} (z - L[update rows, :]' x_u),
// two wavefronts: wave w takes the output column groups q = w and w + 2
// (columns [16q, 16q+16), acc[h] for q = w + 2h).  The update rows go by 64-row
// tiles (from the last, as the four-wave kernel does); a tile's 64 x 16 column
// chunk q is read coalesced (lane = row) and transposed through the wave's LDS
// buffer; chunk q of the next tile is in flight while the current one is used,
// and after the last tile the same slot receives row block q of L_11^{-T}.
// pair barrier: both waves of a pair count in on *ctr (zeroed before any wait of
// the launch) and wait for two; the other pair of the workgroup is not involved
__device__ __forceinline__ void pair_barrier(int *ctr, int lane) {
    wave_sync();
    if (lane == 0) __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < 2) __builtin_amdgcn_s_sleep(1);
}
template <class T, int NR>
__device__ __forceinline__ void bwd_leaf_task(const FlowTask &t, T (*Sb)[kTile][17], T (*vx)[NR], int *sync, int w,
                                              int lane, const int32_t *__restrict__ rows, const T *__restrict__ L,
                                              const T *__restrict__ Linv, const T *__restrict__ Z, T *X, int nrhs,
                                              int32_t *cnt, int64_t nsn) {
    int32_t *abort_flag = cnt + nsn;
    const int j = lane & 15, cg = lane >> 4;
    const int rq = g16_r0<NR>(cg);
    const int nc = t.nr, npz = (nc + 15) >> 4;
    const int nchunk = (t.u1 - t.u0 + kTile - 1) / kTile;
    const T *Lc = L + t.lofs;                          // front row 0, column 0
    const int32_t *rw = rows + t.rofs;
    T (*S)[17] = Sb[w];                                // this wave's transposition buffer
    auto tile_row0 = [&](int e) -> int { return t.u0 + kTile * (nchunk - 1 - e); };
    auto tile_rows = [&](int e) -> int { return min(kTile, t.u1 - tile_row0(e)); };
    // slot h holds chunk q = w + 2h of the current tile, after the last tile row
    // block q of L_11^{-T} (g16 lane layout); loaded unconditionally (indices
    // clamped), so the compiler's wait counts stay exact
    // entry c: base[a + b min(c0 + c, nc - 1)] -- chunk: a = row (lane), b = ld,
    // c0 = 16q; inverse: a = nc (column 16q + j), b = 1, c0 = 16cg; one load path
    // for both, so that every load keeps the SGPR-base form
    auto load_next = [&](T (&v)[16], int e, int q) {
        const bool tile = e < nchunk;
        const T *base = uniform_ptr(tile ? Lc + tile_row0(e) : Linv + t.iofs);
        const uint32_t a = tile ? (uint32_t)min(lane, tile_rows(e) - 1) : (uint32_t)(nc * min(16 * q + j, nc - 1));
        const uint32_t b = (uint32_t)uniform_int(tile ? (int)t.ms : 1);
        const int c0 = tile ? 16 * q : 16 * cg;
#pragma unroll
        for (int c = 0; c < 16; ++c) v[c] = ld_off(base, a + b * (uint32_t)min(c0 + c, nc - 1));
    };
    T buf[2][16], acc[2][NR], zv[2][NR / 4], xr[NR];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const T *zp = Z + (t.c0 + min(16 * (w + 2 * h) + j, nc - 1)) * nrhs;
#pragma unroll
        for (int rr = 0; rr < NR / 4; ++rr) zv[h][rr] = zp[min(rq + rr, nrhs - 1)];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) load_next(buf[h], 0, w + 2 * h);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[h][r] = T(0);
    int32_t ridx[NR];                                  // row indices of the next tile
    if (nchunk > 0) g16_rows<NR>(ridx, rw + tile_row0(0), tile_rows(0), cg, j);
    bool ok = true;
    for (int e = 0; e < nchunk; ++e) {
        const int nre = tile_rows(e);
        ok = g16_load_wait_rows<T, NR>(xr, X, ridx, nre, nrhs, cg, j, abort_flag) && ok;
        const int en = min(e + 1, nchunk - 1);
        g16_rows<NR>(ridx, rw + tile_row0(en), tile_rows(en), cg, j);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = w + 2 * h;
            if (q < npz) {
#pragma unroll
                for (int k = 0; k < 16; ++k) S[lane][k] = (lane < nre && 16 * q + k < nc) ? buf[h][k] : T(0);
                wave_sync();
                T m[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) m[k] = S[16 * cg + k][j];
                wave_sync();
                g16_mac<T, NR>(acc[h], m, xr);
            }
            load_next(buf[h], e + 1, q);
        }
    }
    if (!ok) sync[1] = 1;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = w + 2 * h;
        if (q < npz) {
            g16_reduce<T, NR>(acc[h]);
            const int c = 16 * q + j;
#pragma unroll
            for (int rr = 0; rr < NR / 4; ++rr) {
                const int r = rq + rr;
                vx[c][r] = (c < nc && r < nrhs) ? zv[h][rr] - acc[h][r] : T(0);
            }
        }
    }
    pair_barrier(&sync[0], lane);
    if (sync[1]) return;
    g16_from_lds<T, NR>(xr, vx, nc, cg, j);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = w + 2 * h;
        if (q < npz) {
            const int c = 16 * q + j;
            T a[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) a[r] = T(0);
            g16_mac<T, NR>(a, buf[h], xr);
            g16_reduce<T, NR>(a);
            if (c < nc) {
#pragma unroll
                for (int rr = 0; rr < NR / 4; ++rr) {
                    const int r = rq + rr;
                    if (r < nrhs) st_dev(X + (t.c0 + c) * nrhs + r, a[r]);
                }
            }
        }
    }
}

// The backward launch: workgroups [0, nmain) take one task each of the list's
// head, the rest the bottom subtrees' tasks (the tail) two at a time, task
// nmain + 2(b - nmain) + p on waves 2p, 2p+1.  A bottom task waits only on earlier
// tasks (the other pair of its workgroup, if at all, is the earlier one and
// resident), so in-order dispatch always progresses.  The pairs' buffers overlay
// the four-wave task's staging tiles.
template <class T, int NR>
constexpr bool kLeafPairs = sizeof(T) * NR <= 64;     // (FP64 with 16 right-hand sides would spill)
template <class T, int NR>
__global__ void __launch_bounds__(kBS) bwd_flow_kernel(const FlowTask *__restrict__ tasks, const int32_t *__restrict__ rows,
                                                       const T *__restrict__ L, const T *__restrict__ Linv,
                                                       const T *__restrict__ Z, T *X, T *Tb, T *P, int nrhs,
                                                       int32_t *cnt, int64_t nsn, int64_t nmain, int64_t nleaf) {
    constexpr int NB = kBwdStage<T>;
    constexpr int kMain = NB * kTile * (kTile + 1) + kTile * NR;
    constexpr int kPair = 2 * kTile * 17 + kTile * NR;
    __shared__ __attribute__((aligned(16))) T lds[kLeafPairs<T, NR> && 2 * kPair > kMain ? 2 * kPair : kMain];
    __shared__ int flags[5];                           // main: failed; pairs: arrivals and failure, per pair
    __shared__ int32_t ticket_slot;
    const int64_t b = draw_ticket(cnt + nsn + 2, &ticket_slot);
    if (kLeafPairs<T, NR> && b >= nmain) {
        const int tid = threadIdx.x, pw = uniform_int(tid >> 6), p = pw >> 1;
        if (tid < 4) flags[1 + tid] = 0;
        __syncthreads();                               // before any wait
        const int64_t k = nmain + 2 * (b - nmain) + p;
        if (k >= nmain + nleaf) return;
        T *base = lds + p * kPair;
        bwd_leaf_task<T, NR>(tasks[k], (T (*)[kTile][17])base, (T (*)[NR])(base + 2 * kTile * 17), &flags[1 + 2 * p],
                             pw & 1, tid & 63, rows, L, Linv, Z, X, nrhs, cnt, nsn);
        return;
    }
    bwd_main_task<T, NR>(tasks[b], (T (*)[kTile][kTile + 1])lds, (T (*)[NR])(lds + NB * kTile * (kTile + 1)),
                         flags[0], rows, L, Linv, Z, X, Tb, P, nrhs, cnt, nsn);
}

// ----------------------------------------------------------------- refinement helpers
// Bp[i*nrhs + r] = b[perm[i] + ndof*r]
__global__ void gather_rhs_kernel(int64_t n, int nrhs, int64_t ndof, const int32_t *__restrict__ perm,
                                  const double *__restrict__ b, double *__restrict__ Bp) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n * nrhs; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = q / nrhs;
        int r = (int)(q % nrhs);
        Bp[q] = b[perm[i] + ndof * r];
    }
}

template <class T>
__global__ void to_t_kernel(int64_t n, const double *__restrict__ a, T *__restrict__ o) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) o[q] = (T)a[q];
}

// The refined iterate is a double-double (hi + lo, |lo| <= ulp(hi)/2) so that its
// residual can fall below the ~2^-53 |K||x| floor of any vector stored in double.
// X += (double) Y in double-double, and scatter to the global interleaved
// vectors xg_hi / xg_lo (ndof x nrhs, zero on fixed DOFs).
// Error-free transformations need every product and sum rounded on its own:
// contraction into fma is switched off where they are used.
__device__ inline void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

template <class T>
__global__ void accumulate_kernel(int64_t n, int nrhs, const int32_t *__restrict__ perm, const T *__restrict__ Y,
                                  double *__restrict__ Xh, double *__restrict__ Xl, double *__restrict__ xgh,
                                  double *__restrict__ xgl, const int32_t *__restrict__ active) {
#pragma clang fp contract(off)
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n * nrhs; q += (int64_t)gridDim.x * blockDim.x) {
        int r = (int)(q % nrhs);
        if (!active[r]) continue;
        double s, e;
        two_sum(Xh[q], (double)Y[q], s, e);
        e += Xl[q];
        const double h = s + e;
        const double l = e - (h - s);
        Xh[q] = h;
        Xl[q] = l;
        const int64_t g = (int64_t)perm[q / nrhs] * nrhs + r;
        xgh[g] = h;
        xgl[g] = l;
    }
}

// (h, l) += (h2, l2) in double-double
__device__ inline void dd_add(double &h, double &l, double h2, double l2) {
#pragma clang fp contract(off)
    double s, e;
    two_sum(h, h2, s, e);
    e += l + l2;
    h = s + e;
    l = e - (h - s);
}

// R = Bp - K_ff x for x = xgh + xgl (xgl may be null: x = xgh), rows of K_ff in
// elimination order, NR right-hand sides.  Eight lanes share a row (coalesced
// column/value reads); each accumulates its products exactly split with fma into
// a double-double, and the eight partial sums are combined in a fixed butterfly,
// so R is the exact residual to ~2^-104 relative, rounded once to double.  xg is
// zero on fixed DOFs, so the full CSR row can be used.
template <int NR>
__global__ void __launch_bounds__(kBS) residual_kernel(int64_t n, const int32_t *__restrict__ perm,
                                                       const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                       const double *__restrict__ val, const double *__restrict__ xgh,
                                                       const double *__restrict__ xgl, const double *__restrict__ Bp,
                                                       double *__restrict__ R) {
#pragma clang fp contract(off)
    const int sub = threadIdx.x & 7;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const bool valid = i < n;
    const int64_t g = valid ? perm[i] : 0;
    const int64_t p0 = valid ? rowptr[g] : 0, p1 = valid ? rowptr[g + 1] : 0;
    double hi[NR], lo[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) { hi[r] = 0.0; lo[r] = 0.0; }
    for (int64_t p = p0 + sub; p < p1; p += 8) {
        const double kv = val[p];
        const int64_t c = col[p];
        const double *xh = xgh + c * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double xv = xh[r];
            const double pr = kv * xv;
            const double pe = fma(kv, xv, -pr);
            double t, e;
            two_sum(hi[r], pr, t, e);
            hi[r] = t;
            lo[r] += e + pe;
        }
        if (xgl) {
            const double *xl = xgl + c * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) lo[r] = fma(kv, xl[r], lo[r]);
        }
    }
#pragma unroll
    for (int o = 4; o >= 1; o >>= 1) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double h2 = __shfl_xor(hi[r], o, 8), l2 = __shfl_xor(lo[r], o, 8);
            dd_add(hi[r], lo[r], h2, l2);
        }
    }
    if (valid && sub == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double t, te;
            two_sum(Bp[i * NR + r], -hi[r], t, te);
            R[i * NR + r] = t + (te - lo[r]);
        }
    }
}

// The block that finishes last (a device-scope ticket, reset by that block for
// the next launch) adds the kBS blocks' partials of each RHS in a fixed tree
// order: one launch per reduction, the same sums whichever block is last.  The
// partials go out as device-coherent stores, completed before the ticket, and
// are read back with device-coherent loads (the protocol of the solve kernels:
// no fence, which would write back and invalidate the whole L2 per block).
// sh: the block's reduction buffer (free once its partials are stored).
template <int NR>
__device__ __forceinline__ void reduce_last_block(double (*sh)[kBS], double *part, double *out,
                                                  unsigned *ticket, double *host_out = nullptr) {
    __shared__ bool last;
    if (threadIdx.x < NR) {
        st_dev(part + (int64_t)threadIdx.x * gridDim.x + blockIdx.x, sh[threadIdx.x][0]);
        complete_stores();
    }
    __syncthreads();
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
#pragma unroll
    for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] = ld_dev(part + (int64_t)r * kBS + threadIdx.x);
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < NR) out[threadIdx.x] = sh[threadIdx.x][0];
    if (host_out && threadIdx.x < NR) host_out[threadIdx.x] = sh[threadIdx.x][0];
    if (threadIdx.x == 0) st_dev(ticket, 0u);
}

// per-block partial sums of squares, per RHS, in tree order; the last block adds
// the partials (reduce_last_block) into out and into the pinned host array
// host_out, which the host reads after a stream synchronization (no copy
// launch).  Launched with kBS blocks.
template <int NR>
__global__ void __launch_bounds__(kBS) sumsq_partial_kernel(int64_t n, const double *__restrict__ v, double *__restrict__ part,
                                                            double *__restrict__ out, unsigned *ticket, double *host_out) {
    __shared__ double sh[NR][kBS];
    double a[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) a[r] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double x = v[i * NR + r];
            a[r] += x * x;
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] = a[r];
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + w];
        }
        __syncthreads();
    }
    reduce_last_block<NR>(sh, part, out, ticket, host_out);
}

// q = K_ff p (plain FP64), p given in the global interleaved layout (zero on fixed
// DOFs), q in elimination order; eight lanes per row, fixed-order lane reduction
template <int NR>
__global__ void __launch_bounds__(kBS) spmv_kernel(int64_t n, const int32_t *__restrict__ perm,
                                                   const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                   const double *__restrict__ val, const double *__restrict__ pg,
                                                   double *__restrict__ q) {
    const int sub = threadIdx.x & 7;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
    const bool valid = i < n;
    const int64_t g = valid ? perm[i] : 0;
    const int64_t p0 = valid ? rowptr[g] : 0, p1 = valid ? rowptr[g + 1] : 0;
    double a[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) a[r] = 0.0;
    // U entries per lane at a time: their (value, column) loads, then their x
    // loads, all in flight before the first product (a loop of single entries
    // waits two memory latencies per entry); products in entry order
    constexpr int U = NR >= 16 ? 2 : 4;
    for (int64_t p = p0 + sub; p < p1; p += 8 * U) {
        double kv[U];
        int64_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t pu = min(p + 8 * u, p1 - 1);
            kv[u] = val[pu];
            c[u] = col[pu];
        }
        double xv[U][NR];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < NR; ++r) xv[u][r] = pg[c[u] * NR + r];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = p + 8 * u < p1;
#pragma unroll
            for (int r = 0; r < NR; ++r) a[r] = in ? fma(kv[u], xv[u][r], a[r]) : a[r];
        }
    }
#pragma unroll
    for (int o = 4; o >= 1; o >>= 1) {
#pragma unroll
        for (int r = 0; r < NR; ++r) a[r] += __shfl_xor(a[r], o, 8);
    }
    if (valid && sub == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) q[i * NR + r] = a[r];
    }
}

// per-block partial dot products a.b per RHS (tree order); the last block adds
// the partials into out.  Launched with kBS blocks.
template <int NR>
__global__ void __launch_bounds__(kBS) dot_partial_kernel(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                                                          double *__restrict__ part, double *__restrict__ out,
                                                          unsigned *ticket) {
    __shared__ double sh[NR][kBS];
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] = fma(a[i * NR + r], b[i * NR + r], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] = acc[r];
    __syncthreads();
    for (int w = kBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int r = 0; r < NR; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + w];
        }
        __syncthreads();
    }
    reduce_last_block<NR>(sh, part, out, ticket);
}

// conjugate-gradient scalars per RHS (one thread per RHS):
// mode 0: alpha = rho / pq;  mode 1: beta = rho_new / rho, rho = rho_new.
// An inactive RHS keeps its rho (and its search direction, see pcg_dir_kernel),
// so that it can continue later from a replaced residual.
__global__ void pcg_scalars_kernel(int nrhs, int mode, const int32_t *__restrict__ act, const double *__restrict__ num,
                                   double *__restrict__ rho, double *__restrict__ coef) {
    const int r = threadIdx.x;
    if (r >= nrhs) return;
    if (mode == 0) {
        const double pq = num[r];
        coef[r] = (act[r] && pq > 0.0) ? rho[r] / pq : 0.0;
    } else {
        const double rn = num[r];
        coef[r] = (act[r] && rho[r] != 0.0) ? rn / rho[r] : 0.0;
        if (act[r]) rho[r] = rn;
    }
}

// d += alpha p;  s -= alpha q
template <int NR>
__global__ void pcg_step_kernel(int64_t n, const double *__restrict__ alpha, const double *__restrict__ p,
                                const double *__restrict__ q, double *__restrict__ d, double *__restrict__ sres) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * NR; e += (int64_t)gridDim.x * blockDim.x) {
        const double a = alpha[e % NR];
        d[e] = fma(a, p[e], d[e]);
        sres[e] = fma(-a, q[e], sres[e]);
    }
}

// mode 0: p = z + beta p for the active RHS (the others keep p);
// mode 1 (start): p = z, d = 0 for every RHS;
// mode 2 (continue after x += d): p = z + beta p for the active RHS, d = 0
template <int NR>
__global__ void pcg_dir_kernel(int64_t n, int mode, const int32_t *__restrict__ act, const double *__restrict__ beta,
                               const double *__restrict__ z, double *__restrict__ p, double *__restrict__ d) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * NR; e += (int64_t)gridDim.x * blockDim.x) {
        if (mode == 1) {
            p[e] = z[e];
            d[e] = 0.0;
        } else {
            if (act[e % NR]) p[e] = fma(beta[e % NR], p[e], z[e]);
            if (mode == 2) d[e] = 0.0;
        }
    }
}

// global interleaved pg[perm[i]] = p[i]
template <int NR>
__global__ void scatter_global_kernel(int64_t n, const int32_t *__restrict__ perm, const double *__restrict__ p,
                                      double *__restrict__ pg) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * NR; e += (int64_t)gridDim.x * blockDim.x)
        pg[(int64_t)perm[e / NR] * NR + e % NR] = p[e];
}

template <class T>
__global__ void from_t_kernel(int64_t n, const T *__restrict__ a, double *__restrict__ o) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) o[q] = (double)a[q];
}

#define JFEM_NR_SWITCH(nrhs, CALL)                                                                            \
    switch (nrhs) {                                                                                           \
    case 1: CALL(1); break;   case 2: CALL(2); break;   case 3: CALL(3); break;   case 4: CALL(4); break;    \
    case 5: CALL(5); break;   case 6: CALL(6); break;   case 7: CALL(7); break;   case 8: CALL(8); break;    \
    case 9: CALL(9); break;   case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break;  \
    case 13: CALL(13); break; case 14: CALL(14); break; case 15: CALL(15); break; default: CALL(16); break; \
    }

// global column-major x (ndof x nrhs) from the interleaved xg
__global__ void export_kernel(int64_t ndof, int nrhs, const double *__restrict__ xg, double *__restrict__ x) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ndof * nrhs; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t g = q % ndof;
        int r = (int)(q / ndof);
        x[q] = xg[g * nrhs + r];
    }
}

__global__ void diag_ratio_kernel(int64_t n, const int32_t *__restrict__ perm, const int32_t *__restrict__ pos2sn,
                                  const int64_t *__restrict__ col0, const int64_t *__restrict__ mrow,
                                  const int64_t *__restrict__ lofs, const int32_t *__restrict__ diagpos,
                                  const double *__restrict__ val, const void *__restrict__ L, int is64,
                                  double *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int s = pos2sn[i];
        const int64_t j = i - col0[s];
        const int64_t off = lofs[s] + j + mrow[s] * j;
        const double l = is64 ? ((const double *)L)[off] : (double)((const float *)L)[off];
        const int32_t dp = diagpos[perm[i]];
        const double kd = dp >= 0 ? fabs(val[dp]) : 0.0;
        out[i] = kd > 1e-30 ? (l * l) / kd : 1.0;
    }
}

// inertia of a signed factorization: the number of negative signs and the
// smallest and largest pivot magnitude D_kk = L_kk^2, one workgroup, fixed
// reduction tree (out: count, min, max)
__global__ void __launch_bounds__(1024) inertia_kernel(int64_t n, const int32_t *__restrict__ pos2sn,
                                                       const int64_t *__restrict__ col0, const int64_t *__restrict__ mrow,
                                                       const int64_t *__restrict__ lofs, const double *__restrict__ L,
                                                       const double *__restrict__ J, double *__restrict__ out) {
    __shared__ double sc[1024], smin[1024], smax[1024];
    double cnt = 0.0, mn = INFINITY, mx = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) {
        const int s = pos2sn[i];
        const int64_t j = i - col0[s];
        const double l = L[lofs[s] + j + mrow[s] * j];
        const double d = l * l;
        cnt += J[i] < 0.0 ? 1.0 : 0.0;
        mn = fmin(mn, d);
        mx = fmax(mx, d);
    }
    sc[threadIdx.x] = cnt; smin[threadIdx.x] = mn; smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sc[threadIdx.x] += sc[threadIdx.x + w];
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sc[0]; out[1] = smin[0]; out[2] = smax[0]; }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBS - 1) / kBS, 65536)); }

// numeric factorization of V matrices sharing S (K values, factors and update
// matrices of variant v at v times the strides in vs), in one launch sequence
// SG: the signed factorization A = L J L' (J into Jv, one sign per elimination
// position); solve_blocks: also the diagonal-block inverses and the X = L_11^{-1}
// of the multi-panel supernodes the triangular solves need
template <class T, bool SG = false>
hipError_t numeric(jfem_ctx *ctx, const jfem_symbolic *S, const double *Kval, T *L, T *U, T *Linv, int V,
                   VarStride vs, T *Jv = nullptr, bool solve_blocks = true) {
    hipStream_t st = ctx->stream;
    const SymbolicHost &H = S->H;
    const bool detail = ctx->timing_level >= 2;
    auto lap = [&](int fam) {
        if (detail) tl_mark(ctx, fam);
    };
    if (detail) tl_begin(ctx);
    const int64_t nLv = V > 1 ? vs.L * (V - 1) + H.L_entries : H.L_entries;
    const int64_t nUv = V > 1 ? vs.U * (V - 1) + H.U_entries : H.U_entries;
    hipLaunchKernelGGL(zero16_kernel<T>, dim3(grid_for(nLv * (int64_t)sizeof(T) / 16)), dim3(kBS), 0, st, L, nLv);
    if (H.U_entries)
        hipLaunchKernelGGL(zero16_kernel<T>, dim3(grid_for(nUv * (int64_t)sizeof(T) / 16)), dim3(kBS), 0, st, U, nUv);
    int64_t na = (int64_t)H.amap_src.size();
    if (na)
        hipLaunchKernelGGL(scatter_kernel<T>, dim3(grid_for(na), V), dim3(kBS), 0, st, na, S->d_amap_src, S->d_amap_dst,
                           Kval, L, vs);
    lap(-1);
    const int64_t nlev = (int64_t)H.level_ptr.size() - 1;
    for (int64_t l = 0; l < nlev; ++l) {
        if (H.step_ptr[l + 1] > H.step_ptr[l]) {
            // diagonal blocks of the level's first panels
            const int64_t stp = H.step_ptr[l];
            const int64_t nd = H.pstep_mid[stp] - H.pstep_off[stp];
            if (nd) {
                hipLaunchKernelGGL((diag_factor_kernel<T, SG>), dim3((unsigned)nd, V), dim3(kBS), 0, st,
                                   S->d_ptasks + H.pstep_off[stp], S->d_ncol, S->d_m, S->d_lofs, S->d_iofs, S->d_col0, L,
                                   Linv, ctx->d_status, vs, Jv);
                lap(FAM_PANEL);
            }
        }
        for (int64_t stp = H.step_ptr[l]; stp < H.step_ptr[l + 1]; ++stp) {
            const int64_t np = H.pstep_off[stp + 1] - H.pstep_off[stp];
            const int64_t nu = H.ustep_off[stp + 1] - H.ustep_off[stp];
            if (np) {
                hipLaunchKernelGGL(tile_solve_kernel<T>, dim3((unsigned)np, V), dim3(kBS), 0, st, S->d_ptasks + H.pstep_off[stp],
                                   S->d_ncol, S->d_m, S->d_lofs, S->d_iofs, L, (const T *)Linv, vs);
                lap(FAM_PANEL);
            }
            if (nu) {
                hipLaunchKernelGGL((update_kernel<T, SG>), dim3((unsigned)nu, V), dim3(kBS), 0, st,
                                   S->d_utasks + H.ustep_off[stp], S->d_ncol, S->d_m, S->d_lofs, S->d_uofs, S->d_iofs, L,
                                   U, Linv, 1, S->d_col0, ctx->d_status, vs, Jv);
                lap(FAM_UPDATE);
            }
        }
        // the level's update matrices, with the full rank of each front, pushed into
        // the parents' fronts (one launch per sibling rank)
        for (int64_t q = H.slev_ptr[l]; q < H.slev_ptr[l + 1]; ++q) {
            const int64_t nsch = H.spush_off[q + 1] - H.spush_off[q];
            if (!nsch) continue;
            hipLaunchKernelGGL((schur_kernel<T, SG>), dim3((unsigned)nsch, V), dim3(kBS), 0, st,
                               S->d_stasks + H.spush_off[q], S->d_relmap, L, U, vs, S->d_col0, (const T *)Jv);
            lap(FAM_SCHUR);
        }
    }
    // inverse diagonal blocks and the X = L_11^{-1} of the multi-panel supernodes
    // (doubling levels, two phases each) for the triangular solves
    if (solve_blocks && !H.flow_b.empty()) {
        hipLaunchKernelGGL(diag_inverse_kernel<T>, dim3((unsigned)H.flow_b.size(), V), dim3(64), 0, st, S->d_flow_b,
                           (const T *)L, Linv, vs);
        for (size_t q = 0; q + 1 < H.istep_off.size(); ++q) {
            const int64_t ni = H.istep_off[q + 1] - H.istep_off[q];
            if (!ni) continue;
            if (q % 2 == 0)
                hipLaunchKernelGGL((inv_tile_kernel<T, 0>), dim3((unsigned)ni, V), dim3(kBS), 0, st,
                                   S->d_itasks + H.istep_off[q], (const T *)L, Linv, vs);
            else
                hipLaunchKernelGGL((inv_tile_kernel<T, 1>), dim3((unsigned)ni, V), dim3(kBS), 0, st,
                                   S->d_itasks + H.istep_off[q], (const T *)L, Linv, vs);
        }
        lap(FAM_INV);
    }
    if (detail) {
        lap(-1);
        tl_flush(ctx);
    }
    return hipGetLastError();
}

// y = (L L')^{-1} y by the dataflow kernels; cnt: ns + 3 words: ns unused, the
// abort flag (cleared once per refinement call), the two launch tickets (cleared
// before every solve); V and Tb: the v and t values the panels of a multi-panel
// supernode exchange (forward and backward); P: the partial-sum slots of the
// forward list, then those of the backward list (kTile x nrhs entries each)
template <class T, int NR>
hipError_t flow_solve_nr(jfem_ctx *ctx, const jfem_symbolic *S, const jfem_factor *F, const T *Y, T *W, T *Z, T *X,
                         T *V, T *Tb, T *P, int nrhs, int32_t *cnt) {
    const SymbolicHost &H = S->H;
    hipStream_t st = ctx->stream;
    const int64_t nsn = H.ns;
    const T *L = (const T *)F->d_L, *Linv = (const T *)F->d_Linv;
    const int64_t nv = H.n * nrhs;
    const int64_t nwv = H.upd_rows * nrhs;     // update vectors: polled by the parents
    // timing level 2: the kernels on the context's event timeline
    const bool detail = ctx->timing_level >= 2;
    if (detail) tl_begin(ctx);
    const int64_t npf = H.flow_f_parts * kTile * nrhs, npv = npf + H.flow_b_parts * kTile * nrhs;
    hipLaunchKernelGGL(solve_marks_kernel<T>, dim3(grid_for(std::max({nv, nwv, npv}))), dim3(kBS), 0, st, Z, X, V, Tb,
                       nv, W, nwv, P, npv, pending_mark<T>(), cnt + nsn + 1);
    auto lap = [&](int fam) { tl_mark(ctx, fam); };
    if (detail) lap(-1);
    // forward: the bottom subtrees' tasks (the head of the list) one wavefront
    // each, then the rest; backward: the rest, then the bottom subtrees (the tail)
    const int64_t nfl = H.flow_f_leaf;
    const int64_t nbl = H.flow_b_leaf, nbm = (int64_t)H.flow_b.size() - nbl;
    if (!H.flow_f.empty()) {
        const int64_t nleaf = kLeafQuads<T, NR> ? nfl : 0;
        const int64_t grid = (nleaf + 3) / 4 + (int64_t)H.flow_f.size() - nleaf;
        hipLaunchKernelGGL((fwd_flow_kernel<T, NR>), dim3((unsigned)grid), dim3(kBS), 0, st, S->d_flow_f, S->d_flow_gb,
                           S->d_relmap, L, Linv, Y, W, Z, V, P, nrhs, cnt, nsn, nleaf);
        if (detail) lap(FAM_FLOW_FWD);
    }
    if (!H.flow_b.empty()) {
        const int64_t nmain = kLeafPairs<T, NR> ? nbm : (int64_t)H.flow_b.size(), nleaf = (int64_t)H.flow_b.size() - nmain;
        const int64_t grid = nmain + (nleaf + 1) / 2;
        hipLaunchKernelGGL((bwd_flow_kernel<T, NR>), dim3((unsigned)grid), dim3(kBS), 0, st, S->d_flow_b, S->d_rows, L,
                           Linv, (const T *)Z, X, Tb, P + npf, nrhs, cnt, nsn, nmain, nleaf);
        if (detail) lap(FAM_FLOW_BWD);
    }
    if (detail) tl_flush(ctx);
    return hipGetLastError();
}

template <class T>
hipError_t flow_solve(jfem_ctx *ctx, const jfem_symbolic *S, const jfem_factor *F, const T *Y, T *W, T *Z, T *X,
                      T *V, T *Tb, T *P, int nrhs, int32_t *cnt) {
    if (nrhs <= 4) return flow_solve_nr<T, 4>(ctx, S, F, Y, W, Z, X, V, Tb, P, nrhs, cnt);
    if (nrhs <= 8) return flow_solve_nr<T, 8>(ctx, S, F, Y, W, Z, X, V, Tb, P, nrhs, cnt);
    return flow_solve_nr<T, 16>(ctx, S, F, Y, W, Z, X, V, Tb, P, nrhs, cnt);
}

template <class V>
hipError_t up(jfem_symbolic *S, V **dst, const std::vector<V> &src, hipStream_t st) {
    *dst = nullptr;
    if (src.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, src.size() * sizeof(V));
    if (e != hipSuccess) return e;
    S->allocs.push_back(*dst);
    S->bytes += (int64_t)(src.size() * sizeof(V));
    return hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, st);
}

hipError_t ensure_scratch(jfem_ctx *ctx, size_t bytes) {
    if (ctx->scratch_bytes >= bytes) return hipSuccess;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    e = hipMalloc(&ctx->scratch, bytes);
    if (e != hipSuccess) return e;
    ctx->scratch_bytes = bytes;
    return hipSuccess;
}

}  // namespace

extern "C" {

int jfem_symbolic_create(jfem_ctx *ctx, const jfem_pattern *P, const jfem_mesh *mesh, const uint8_t *fixed,
                         jfem_symbolic **out, jfem_symbolic_info *info) {
    *out = nullptr;
    if (!mesh || mesh->n_nodes * 6 != P->ndof) return fail(ctx, JFEM_ERR_BAD_ARG, "mesh/pattern mismatch");
    jfem_symbolic *S = new jfem_symbolic();
    S->device = ctx->device;
    S->P = P;
    std::string err;
    if (build_symbolic(P->ndof, P->rowptr.data(), P->col.data(), mesh->xyz.data(), fixed, S->H, err)) {
        delete S;
        return fail(ctx, JFEM_ERR_BAD_ARG, "symbolic: " + err);
    }
    const SymbolicHost &H = S->H;
    if (H.L_entries >= (int64_t)1 << 40) { delete S; return fail(ctx, JFEM_ERR_OOM, "factor too large"); }
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    std::vector<int32_t> level_sn32(H.level_sn.begin(), H.level_sn.end());
    if (e == hipSuccess) e = up(S, &S->d_col0, H.col0, st);
    if (e == hipSuccess) e = up(S, &S->d_ncol, H.ncol, st);
    if (e == hipSuccess) e = up(S, &S->d_m, H.m, st);
    if (e == hipSuccess) e = up(S, &S->d_lofs, H.lofs, st);
    if (e == hipSuccess) e = up(S, &S->d_uofs, H.uofs, st);
    if (e == hipSuccess) e = up(S, &S->d_iofs, H.iofs, st);
    if (e == hipSuccess) e = up(S, &S->d_rofs, H.rofs, st);
    if (e == hipSuccess) e = up(S, &S->d_relofs, H.relofs, st);
    if (e == hipSuccess) e = up(S, &S->d_rows, H.rows, st);
    if (e == hipSuccess) e = up(S, &S->d_relmap, H.relmap, st);
    if (e == hipSuccess) e = up(S, &S->d_child_ptr, H.child_ptr, st);
    if (e == hipSuccess) e = up(S, &S->d_child, H.child, st);
    if (e == hipSuccess) e = up(S, &S->d_perm, H.perm, st);
    if (e == hipSuccess) e = up(S, &S->d_level_sn, level_sn32, st);
    if (e == hipSuccess) e = up(S, &S->d_amap_src, H.amap_src, st);
    if (e == hipSuccess) e = up(S, &S->d_amap_dst, H.amap_dst, st);
    if (e == hipSuccess) e = up(S, &S->d_ptasks, H.ptasks, st);
    if (e == hipSuccess) e = up(S, &S->d_utasks, H.utasks, st);
    if (e == hipSuccess) e = up(S, &S->d_stasks, H.stasks, st);
    if (e == hipSuccess) e = up(S, &S->d_itasks, H.itasks, st);
    if (e == hipSuccess) e = up(S, &S->d_flow_f, H.flow_f, st);
    if (e == hipSuccess) e = up(S, &S->d_flow_b, H.flow_b, st);
    if (e == hipSuccess) e = up(S, &S->d_flow_gb, H.flow_gb, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        jfem_symbolic_free(S);
        return fail(ctx, e == hipErrorOutOfMemory ? JFEM_ERR_OOM : JFEM_ERR_HIP, hipGetErrorString(e));
    }
    if (info) {
        info->n_free = H.n;
        info->n_supernodes = H.ns;
        info->n_levels = (int64_t)H.level_ptr.size() - 1;
        info->nnz_L = H.L_entries;
        info->flops = H.flops;
        info->max_front = H.max_front;
        info->update_entries = H.U_entries;
        info->update_flops = H.update_flops;
        info->update_launches = (int64_t)H.ustep_off.size() - 1;
        info->schur_flops = H.schur_flops;
        int64_t nsl = 0;
        for (size_t q = 0; q + 1 < H.spush_off.size(); ++q) nsl += H.spush_off[q + 1] > H.spush_off[q];
        info->schur_launches = nsl;
        info->inv_entries = H.inv_entries;
        info->upd_rows = H.upd_rows;
        info->flow_fwd_tasks = (int64_t)H.flow_f.size();
        info->flow_bwd_tasks = (int64_t)H.flow_b.size();
        int64_t se = 0;
        for (int64_t k = 0; k < H.ns; ++k)
            for (int64_t c = 0; c < H.ncol[k]; c += kPanel) se += (H.m[k] - c) * std::min<int64_t>(kPanel, H.ncol[k] - c);
        info->solve_entries = se;
    }
    *out = S;
    return JFEM_OK;
}

int jfem_symbolic_free(jfem_symbolic *S) {
    if (!S) return JFEM_OK;
    (void)hipSetDevice(S->device);
    for (void *p : S->allocs) (void)hipFree(p);
    delete S;
    return JFEM_OK;
}

static int do_numeric(jfem_ctx *ctx, jfem_factor *F, const jfem_matrix *K, int64_t *bad_column) {
    const jfem_symbolic *S = F->S;
    if (K->P != S->P) return fail(ctx, JFEM_ERR_BAD_ARG, "matrix and symbolic use different patterns");
    if (bad_column) *bad_column = -1;
    JFEM_CHECK_HIP(ctx, launch_reset_status(ctx->stream, ctx->d_status));
    timer_start(ctx);
    hipError_t e = F->precision == 64
                       ? numeric<double>(ctx, S, K->d_val, (double *)F->d_L, (double *)F->d_U, (double *)F->d_Linv, 1,
                                         VarStride{})
                       : numeric<float>(ctx, S, K->d_val, (float *)F->d_L, (float *)F->d_U, (float *)F->d_Linv, 1,
                                        VarStride{});
    JFEM_CHECK_HIP(ctx, e);
    timer_stop(ctx, FAM_FACTOR, 1);
    int64_t st[2];
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (st[1] != 0) {
        int64_t g = (st[0] >= 0 && st[0] < S->H.n) ? S->H.perm[st[0]] : -1;
        if (bad_column) *bad_column = g;
        return fail(ctx, JFEM_ERR_NOT_SPD, "non-positive pivot at DOF " + std::to_string(g));
    }
    return JFEM_OK;
}

int jfem_factor_create(jfem_ctx *ctx, const jfem_symbolic *S, const jfem_matrix *K, int precision, jfem_factor **out,
                       int64_t *bad_column) {
    clear_stale_hip_error();
    *out = nullptr;
    if (precision != 64 && precision != 32) return fail(ctx, JFEM_ERR_BAD_ARG, "precision must be 32 or 64");
    jfem_factor *F = new jfem_factor();
    F->device = ctx->device;
    F->S = S;
    F->precision = precision;
    const size_t es = precision == 64 ? 8 : 4;
    hipError_t e = hipMalloc(&F->d_L, std::max<int64_t>(S->H.L_entries, 1) * es);
    if (e == hipSuccess) e = hipMalloc(&F->d_U, std::max<int64_t>(S->H.U_entries, 1) * es);
    if (e == hipSuccess) e = hipMalloc(&F->d_Linv, std::max<int64_t>(S->H.inv_entries, 1) * es);
    if (e != hipSuccess) {
        jfem_factor_free(F);
        return fail(ctx, JFEM_ERR_OOM, "factor storage");
    }
    F->bytes = (S->H.L_entries + S->H.U_entries + S->H.inv_entries) * (int64_t)es;
    int st = do_numeric(ctx, F, K, bad_column);
    if (st != JFEM_OK) {
        jfem_factor_free(F);
        return st;
    }
    *out = F;
    return JFEM_OK;
}

int jfem_refactor(jfem_ctx *ctx, jfem_factor *F, const jfem_matrix *K, int64_t *bad_column) {
    return do_numeric(ctx, F, K, bad_column);
}

int jfem_factor_free(jfem_factor *F) {
    if (!F || F->view) return JFEM_OK;     // variants are freed with their batch
    (void)hipSetDevice(F->device);
    if (F->d_L) (void)hipFree(F->d_L);
    if (F->d_U) (void)hipFree(F->d_U);
    if (F->d_Linv) (void)hipFree(F->d_Linv);
    delete F;
    return JFEM_OK;
}

int64_t jfem_factor_device_bytes(const jfem_factor *F) { return F ? F->bytes : 0; }

// numeric factors of every variant of a batch over one symbolic, in one launch
// sequence (each launch of the factorization covers all variants)
int jfem_batch_factor(jfem_ctx *ctx, jfem_batch *B, const jfem_symbolic *S, int precision, int64_t *bad_variant,
                      int64_t *bad_column) {
    if (bad_variant) *bad_variant = -1;
    if (bad_column) *bad_column = -1;
    if (precision != 64 && precision != 32) return fail(ctx, JFEM_ERR_BAD_ARG, "precision must be 32 or 64");
    if (S->P != B->P) return fail(ctx, JFEM_ERR_BAD_ARG, "batch and symbolic use different patterns");
    const SymbolicHost &H = S->H;
    const size_t es = precision == 64 ? 8 : 4;
    if (B->S != S || B->precision != precision) {
        batch_free_factor(B);
        B->L_stride = jfem_stride(std::max<int64_t>(H.L_entries, 1), es);
        B->U_stride = jfem_stride(std::max<int64_t>(H.U_entries, 1), es);
        B->inv_stride = jfem_stride(std::max<int64_t>(H.inv_entries, 1), es);
        hipError_t e = hipMalloc(&B->d_L, (size_t)(B->L_stride * B->V) * es);
        if (e == hipSuccess) e = hipMalloc(&B->d_U, (size_t)(B->U_stride * B->V) * es);
        if (e == hipSuccess) e = hipMalloc(&B->d_Linv, (size_t)(B->inv_stride * B->V) * es);
        if (e != hipSuccess) {
            batch_free_factor(B);
            return fail(ctx, JFEM_ERR_OOM, "batch factor storage");
        }
        B->S = S;
        B->precision = precision;
        for (int64_t v = 0; v < B->V; ++v) {
            jfem_factor *F = new jfem_factor();
            F->device = B->device;
            F->S = S;
            F->precision = precision;
            F->d_L = (char *)B->d_L + v * B->L_stride * es;
            F->d_U = (char *)B->d_U + v * B->U_stride * es;
            F->d_Linv = (char *)B->d_Linv + v * B->inv_stride * es;
            F->bytes = (B->L_stride + B->U_stride + B->inv_stride) * (int64_t)es;
            F->view = true;
            B->fviews.push_back(F);
        }
    }
    VarStride vs;
    vs.K = B->val_stride;
    vs.L = B->L_stride;
    vs.U = B->U_stride;
    vs.inv = B->inv_stride;
    vs.col = H.n;
    JFEM_CHECK_HIP(ctx, launch_reset_status(ctx->stream, ctx->d_status));
    timer_start(ctx);
    hipError_t e = precision == 64
                       ? numeric<double>(ctx, S, B->d_val, (double *)B->d_L, (double *)B->d_U, (double *)B->d_Linv,
                                         (int)B->V, vs)
                       : numeric<float>(ctx, S, B->d_val, (float *)B->d_L, (float *)B->d_U, (float *)B->d_Linv,
                                        (int)B->V, vs);
    JFEM_CHECK_HIP(ctx, e);
    timer_stop(ctx, FAM_FACTOR, 1);
    int64_t st[2];
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(st, ctx->d_status, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (st[1] != 0) {
        const int64_t n = std::max<int64_t>(H.n, 1), v = st[0] / n, c = st[0] % n;
        const int64_t g = (c >= 0 && c < H.n) ? H.perm[c] : -1;
        if (bad_variant) *bad_variant = v;
        if (bad_column) *bad_column = g;
        return fail(ctx, JFEM_ERR_NOT_SPD, "variant " + std::to_string(v) + ": non-positive pivot at DOF " +
                                               std::to_string(g));
    }
    return JFEM_OK;
}

int jfem_batch_factor_view(const jfem_batch *B, int64_t v, jfem_factor **view) {
    *view = nullptr;
    if (!B || v < 0 || v >= (int64_t)B->fviews.size()) return JFEM_ERR_BAD_ARG;
    *view = B->fviews[(size_t)v];
    return JFEM_OK;
}

int jfem_symbolic_order(const jfem_symbolic *S, int64_t *dofs_out) {
    if (!S) return JFEM_ERR_BAD_ARG;
    for (int64_t k = 0; k < S->H.n; ++k) dofs_out[k] = S->H.perm[k];
    return JFEM_OK;
}

int jfem_factor_pivot_ratios(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, double *ratios_out) {
    const jfem_symbolic *S = F->S;
    const SymbolicHost &H = S->H;
    if (K->P != S->P) return fail(ctx, JFEM_ERR_BAD_ARG, "matrix and factor use different patterns");
    if (H.n == 0) return JFEM_OK;
    std::vector<int32_t> pos2sn(H.n), diagpos(K->P->diagpos);
    for (int64_t s = 0; s < H.ns; ++s)
        for (int64_t j = 0; j < H.ncol[s]; ++j) pos2sn[H.col0[s] + j] = (int32_t)s;
    int32_t *d_p2s = nullptr, *d_dp = nullptr;
    double *d_out = nullptr;
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_p2s, H.n * sizeof(int32_t)));
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_dp, diagpos.size() * sizeof(int32_t)));
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_out, H.n * sizeof(double)));
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(d_p2s, pos2sn.data(), H.n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(d_dp, diagpos.data(), diagpos.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(diag_ratio_kernel, dim3(grid_for(H.n)), dim3(kBS), 0, ctx->stream, H.n, S->d_perm, d_p2s, S->d_col0,
                       S->d_m, S->d_lofs, d_dp, K->d_val, F->d_L, F->precision == 64 ? 1 : 0, d_out);
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(ratios_out, d_out, H.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(d_p2s); (void)hipFree(d_dp); (void)hipFree(d_out);
    return JFEM_OK;
}

int jfem_factor_small_pivots(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, double threshold,
                             int64_t *dofs_out, int64_t cap, int64_t *count) {
    const SymbolicHost &H = F->S->H;
    *count = 0;
    if (H.n == 0) return JFEM_OK;
    std::vector<double> ratio(H.n);
    const int rc = jfem_factor_pivot_ratios(ctx, F, K, ratio.data());
    if (rc != JFEM_OK) return rc;
    int64_t c = 0;
    for (int64_t i = 0; i < H.n; ++i)
        if (ratio[i] < threshold) {
            if (c < cap) dofs_out[c] = H.perm[i];
            ++c;
        }
    *count = c;
    return JFEM_OK;
}

// Inertia of A_ff (A on the pattern of S, typically K + sigma Kg): the unpivoted
// signed factorization A_ff = L J L' in FP64 on the symbolic analysis S (the same
// launch sequence as the Cholesky, with the pivots' signs carried through the
// updates), the number of negative pivots, and the smallest and largest pivot
// magnitude.  The reference's interval count factors its (diagonally scaled)
// shifted matrices with an unpivoted LDL' and counts negative pivots
// (_eigb_interval_count, src/solver/helpers.jl:105-143).
int jfem_factor_inertia(jfem_ctx *ctx, const jfem_symbolic *S, const jfem_matrix *A, int64_t *negative,
                        double *min_pivot, double *max_pivot, int64_t *bad_column) {
    if (bad_column) *bad_column = -1;
    if (!ctx || !S || !A || !negative) return fail(ctx, JFEM_ERR_BAD_ARG, "null argument");
    if (A->P != S->P) return fail(ctx, JFEM_ERR_BAD_ARG, "matrix and symbolic use different patterns");
    const SymbolicHost &H = S->H;
    *negative = 0;
    if (H.n == 0) {
        if (min_pivot) *min_pivot = 0.0;
        if (max_pivot) *max_pivot = 0.0;
        return JFEM_OK;
    }
    std::vector<int32_t> pos2sn(H.n);
    for (int64_t s = 0; s < H.ns; ++s)
        for (int64_t j = 0; j < H.ncol[s]; ++j) pos2sn[H.col0[s] + j] = (int32_t)s;
    double *dL = nullptr, *dU = nullptr, *dI = nullptr, *dJ = nullptr, *dOut = nullptr;
    int32_t *dP = nullptr;
    hipStream_t st = ctx->stream;
    hipError_t e = hipMalloc((void **)&dL, std::max<int64_t>(H.L_entries, 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dU, std::max<int64_t>(H.U_entries, 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dI, std::max<int64_t>(H.inv_entries, 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dJ, H.n * sizeof(double) + 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dP, H.n * sizeof(int32_t));
    auto release = [&]() {
        for (void *p : {(void *)dL, (void *)dU, (void *)dI, (void *)dJ, (void *)dP})
            if (p) (void)hipFree(p);
    };
    if (e != hipSuccess) {
        release();
        return fail(ctx, JFEM_ERR_OOM, "inertia factor storage");
    }
    dOut = dJ + H.n;
    e = hipMemcpyAsync(dP, pos2sn.data(), H.n * sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_reset_status(st, ctx->d_status);
    if (e == hipSuccess) e = numeric<double, true>(ctx, S, A->d_val, dL, dU, dI, 1, VarStride{}, dJ, false);
    int64_t stw[2] = {0, 0};
    double out[3] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(stw, ctx->d_status, sizeof(stw), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && stw[1] == 0) {
        hipLaunchKernelGGL(inertia_kernel, dim3(1), dim3(1024), 0, st, H.n, dP, S->d_col0, S->d_m, S->d_lofs, dL, dJ,
                           dOut);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out, dOut, sizeof(out), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    release();
    if (e != hipSuccess) return fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    if (stw[1] != 0) {
        const int64_t g = (stw[0] >= 0 && stw[0] < H.n) ? H.perm[stw[0]] : -1;
        if (bad_column) *bad_column = g;
        return fail(ctx, JFEM_ERR_NOT_SPD, "zero or non-finite pivot at DOF " + std::to_string(g));
    }
    *negative = (int64_t)out[0];
    if (min_pivot) *min_pivot = out[1];
    if (max_pivot) *max_pivot = out[2];
    return JFEM_OK;
}

}  // extern "C"

namespace {

// Solve scratch of one refinement call (in the context's scratch buffer): the
// factor-precision vectors of the dataflow solves (Y, W, Z, X, V, Tb and the
// partial-sum slots), their wait words,
// the FP64 iterate, residual and CG vectors (interleaved, n x NR), the global
// iterate for the SpMV (ndof x NR) and the reduction partials.
struct SolveLayout {
    size_t oY, oW, oZs, oXs, oVs, oTs, oPs, oCnt, oXh, oXl, oR, oB, oD, oSv, oPv, oQv, oZv, oXgh, oXgl, oPg;
    size_t oPart, oNrm, oRho, oNum, oAl, oBe, oActO, oActI, oTicket, bytes = 0;
    SolveLayout(const SymbolicHost &H, int64_t ndof, int NR, size_t es) {
        const size_t nv = (size_t)H.n * NR, ng = (size_t)ndof * NR;
        auto take = [&](size_t b) { size_t o = bytes; bytes += (b + 255) & ~(size_t)255; return o; };
        oY = take(nv * es); oW = take((size_t)std::max<int64_t>(H.upd_rows, 1) * NR * es);
        oZs = take(nv * es); oXs = take(nv * es); oVs = take(nv * es); oTs = take(nv * es);
        oPs = take((size_t)std::max<int64_t>(H.flow_f_parts + H.flow_b_parts, 1) * kTile * NR * es);
        oCnt = take((H.ns + 3) * sizeof(int32_t));
        oXh = take(nv * 8); oXl = take(nv * 8); oR = take(nv * 8); oB = take(nv * 8); oD = take(nv * 8);
        oSv = take(nv * 8); oPv = take(nv * 8); oQv = take(nv * 8); oZv = take(nv * 8);
        oXgh = take(ng * 8); oXgl = take(ng * 8); oPg = take(ng * 8);
        oPart = take((size_t)kBS * 16 * 8); oNrm = take(16 * 8); oRho = take(16 * 8); oNum = take(16 * 8);
        oAl = take(16 * 8); oBe = take(16 * 8); oActO = take(16 * 4); oActI = take(16 * 4);
        oTicket = take(sizeof(unsigned));
    }
};

// Refinement of K_ff x = b (nrhs = NR) on a double-double iterate x = xh + xl:
//   outer: r = b - K_ff x exactly (residual_kernel); stop a right-hand side at
//          ||r|| <= tol ||b|| or when ||r|| no longer halves;
//   inner: d ~= K_ff^{-1} r by conjugate gradients preconditioned with the
//          factor (triangular solves in its precision, FP64 vectors, plain FP64
//          SpMV), until ||r - K_ff d|| <= eta ||r||; then x += d in double-double.
// With an FP64 factor the inner loop ends after one preconditioner application
// (classical refinement); with an FP32 factor the Krylov acceleration keeps the
// outer loop converging where plain refinement would stall.  max_refine bounds
// the preconditioner applications to max_refine + 1; max_refine = 0 is the plain
// solve x = F^{-1} b.
template <int NR, class T>
int solve_refine(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, const double *b_dev, double *x_dev,
                 double tol, int max_refine, int32_t *iters, double *relres, double *xlo_dev, double *relres_out,
                 char *static_base) {
    const jfem_symbolic *S = F->S;
    const SymbolicHost &H = S->H;
    const jfem_pattern *P = K->P;
    const int64_t n = H.n, ndof = P->ndof;
    const int64_t nv = n * NR, ng = ndof * NR;
    const int nblk = kBS;
    // One preconditioned conjugate-gradient iteration runs across the outer
    // steps.  Its recursively updated residual drifts from the true one by about
    // 2^-53 || |K_ff| |dx| || for the correction dx accumulated since the last
    // exact residual (~5e-10 ||b|| on thin shells for dx = x), so the drift has to
    // stay small against the residual: once every running right-hand side is
    // below eta * its last exact residual (a checkpoint) the correction d is added
    // to the double-double x, the exact residual replaces the recursive one, and
    // the iteration continues from it with its search directions -- a restart
    // would discard the Krylov space and cost solves.  With eta = 1e-4 the drift
    // stays below ~1e-3 of the residual even when some right-hand sides are 100x
    // past the checkpoint, so a right-hand side can stop at 0.9 tol ||b||: if its
    // exact residual then misses tol, it simply runs on.
    const double eta = 1e-4, stop = 0.9;
    const SolveLayout lay(H, ndof, NR, sizeof(T));
    char *base = static_base;
    if (!base) {
        JFEM_CHECK_HIP(ctx, ensure_scratch(ctx, lay.bytes));
        base = (char *)ctx->scratch;
    }
    T *Y = (T *)(base + lay.oY), *W = (T *)(base + lay.oW), *Zs = (T *)(base + lay.oZs), *Xs = (T *)(base + lay.oXs);
    T *Vs = (T *)(base + lay.oVs), *Ts = (T *)(base + lay.oTs), *Ps = (T *)(base + lay.oPs);
    int32_t *cnt = (int32_t *)(base + lay.oCnt);
    double *Xh = (double *)(base + lay.oXh), *Xl = (double *)(base + lay.oXl), *R = (double *)(base + lay.oR);
    double *Bp = (double *)(base + lay.oB), *Dd = (double *)(base + lay.oD), *Sv = (double *)(base + lay.oSv);
    double *Pv = (double *)(base + lay.oPv), *Qv = (double *)(base + lay.oQv), *Zv = (double *)(base + lay.oZv);
    double *xgh = (double *)(base + lay.oXgh), *xgl = (double *)(base + lay.oXgl), *pg = (double *)(base + lay.oPg);
    double *part = (double *)(base + lay.oPart), *nrm = (double *)(base + lay.oNrm), *rho = (double *)(base + lay.oRho);
    double *num = (double *)(base + lay.oNum), *alpha = (double *)(base + lay.oAl), *beta = (double *)(base + lay.oBe);
    int32_t *act_o = (int32_t *)(base + lay.oActO), *act_i = (int32_t *)(base + lay.oActI);
    unsigned *ticket = (unsigned *)(base + lay.oTicket);
    hipStream_t st = ctx->stream;
    const dim3 gv(grid_for(nv)), gg(grid_for(ng)), blk(kBS);
    const dim3 grow((unsigned)std::max<int64_t>(1, (n * 8 + kBS - 1) / kBS));
    hipLaunchKernelGGL(fill_kernel<double>, gg, blk, 0, st, xgh, ng);
    hipLaunchKernelGGL(fill_kernel<double>, gg, blk, 0, st, xgl, ng);
    hipLaunchKernelGGL(fill_kernel<double>, gg, blk, 0, st, pg, ng);
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(1), dim3(64), 0, st, cnt + H.ns, (int64_t)1);
    hipLaunchKernelGGL(fill_kernel<int32_t>, dim3(1), dim3(64), 0, st, (int32_t *)ticket, (int64_t)1);
    for (int r = 0; r < NR; ++r) { iters[r] = 0; relres[r] = 0.0; }
    if (relres_out)
        for (int r = 0; r < NR; ++r) relres_out[r] = 0.0;
    bool all_ok = true;
    if (n > 0) {
        hipLaunchKernelGGL(fill_kernel<double>, gv, blk, 0, st, Xh, nv);
        hipLaunchKernelGGL(fill_kernel<double>, gv, blk, 0, st, Xl, nv);
        hipLaunchKernelGGL(gather_rhs_kernel, gv, blk, 0, st, n, NR, ndof, S->d_perm, b_dev, Bp);
        hipLaunchKernelGGL(to_t_kernel<double>, gv, blk, 0, st, nv, Bp, R);
        // squared norms per right-hand side, written by the reduction's last block
        // into the context's pinned host array
        auto norms = [&](const double *v, double *outn) -> hipError_t {
            hipLaunchKernelGGL(sumsq_partial_kernel<NR>, dim3(nblk), blk, 0, st, n, v, part, nrm, ticket, ctx->h_red);
            const hipError_t e = hipStreamSynchronize(st);
            if (e == hipSuccess) std::memcpy(outn, ctx->h_red, NR * sizeof(double));
            return e;
        };
        auto dot = [&](const double *a, const double *b2, double *out_dev) {
            hipLaunchKernelGGL(dot_partial_kernel<NR>, dim3(nblk), blk, 0, st, n, a, b2, part, out_dev, ticket);
        };
        auto prec = [&](const double *sv, double *z) -> hipError_t {
            timer_start(ctx);
            hipLaunchKernelGGL(to_t_kernel<T>, gv, blk, 0, st, nv, sv, Y);
            hipError_t e = flow_solve<T>(ctx, S, F, Y, W, Zs, Xs, Vs, Ts, Ps, NR, cnt);
            hipLaunchKernelGGL(from_t_kernel<T>, gv, blk, 0, st, nv, (const T *)Xs, z);
            timer_stop(ctx, FAM_SOLVE, 1);
            return e;
        };
        double bn[NR], rn[NR], prev[NR];
        int32_t act[NR], inner[NR];
        JFEM_CHECK_HIP(ctx, norms(Bp, bn));
        int n_active = 0;
        for (int r = 0; r < NR; ++r) {
            bn[r] = std::sqrt(bn[r]);
            act[r] = bn[r] > 0.0 && !(1.0 <= tol) ? 1 : 0;     // relres of x = 0 is 1 (0 when b = 0)
            relres[r] = bn[r] > 0.0 ? 1.0 : 0.0;
            prev[r] = relres[r];
            rn[r] = bn[r];
            n_active += act[r];
        }
        int nsolve = 0;
        bool started = false;
        const int max_solves = max_refine + 1;
        while (n_active > 0 && nsolve < max_solves) {
            JFEM_CHECK_HIP(ctx, hipMemcpyAsync(act_o, act, NR * 4, hipMemcpyHostToDevice, st));
            if (max_refine == 0) {
                JFEM_CHECK_HIP(ctx, prec(R, Dd));
                ++nsolve;
            } else {
                // inner preconditioned conjugate gradients on K_ff d = r
                for (int r = 0; r < NR; ++r) inner[r] = act[r];
                JFEM_CHECK_HIP(ctx, hipMemcpyAsync(act_i, inner, NR * 4, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(to_t_kernel<double>, gv, blk, 0, st, nv, R, Sv);
                JFEM_CHECK_HIP(ctx, prec(Sv, Zv));
                ++nsolve;
                if (!started) {
                    hipLaunchKernelGGL(pcg_dir_kernel<NR>, gv, blk, 0, st, n, 1, act_i, beta, Zv, Pv, Dd);
                    dot(Sv, Zv, rho);
                    started = true;
                } else {
                    dot(Sv, Zv, num);
                    hipLaunchKernelGGL(pcg_scalars_kernel, dim3(1), dim3(64), 0, st, NR, 1, act_i, num, rho, beta);
                    hipLaunchKernelGGL(pcg_dir_kernel<NR>, gv, blk, 0, st, n, 2, act_i, beta, Zv, Pv, Dd);
                }
                while (true) {
                    timer_start(ctx);
                    hipLaunchKernelGGL(scatter_global_kernel<NR>, gv, blk, 0, st, n, S->d_perm, Pv, pg);
                    hipLaunchKernelGGL(spmv_kernel<NR>, grow, blk, 0, st, n, S->d_perm, P->d_rowptr, P->d_col, K->d_val, pg, Qv);
                    timer_stop(ctx, FAM_SPMV, 1);
                    dot(Pv, Qv, num);
                    hipLaunchKernelGGL(pcg_scalars_kernel, dim3(1), dim3(64), 0, st, NR, 0, act_i, num, rho, alpha);
                    hipLaunchKernelGGL(pcg_step_kernel<NR>, gv, blk, 0, st, n, alpha, Pv, Qv, Dd, Sv);
                    double ss[NR];
                    JFEM_CHECK_HIP(ctx, norms(Sv, ss));
                    int n_in = 0;
                    bool checkpoint = true;
                    for (int r = 0; r < NR; ++r) {
                        if (!inner[r]) continue;
                        const double sn = std::sqrt(ss[r]);
                        if (!(sn > stop * tol * bn[r])) { inner[r] = 0; continue; }  // at the target: frozen
                        ++n_in;
                        if (sn > eta * rn[r]) checkpoint = false;
                    }
                    if (n_in == 0 || checkpoint || nsolve >= max_solves) break;
                    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(act_i, inner, NR * 4, hipMemcpyHostToDevice, st));
                    JFEM_CHECK_HIP(ctx, prec(Sv, Zv));
                    ++nsolve;
                    dot(Sv, Zv, num);
                    hipLaunchKernelGGL(pcg_scalars_kernel, dim3(1), dim3(64), 0, st, NR, 1, act_i, num, rho, beta);
                    hipLaunchKernelGGL(pcg_dir_kernel<NR>, gv, blk, 0, st, n, 0, act_i, beta, Zv, Pv, Dd);
                }
            }
            // x += d (double-double) for the active right-hand sides, exact residual
            hipLaunchKernelGGL(accumulate_kernel<double>, gv, blk, 0, st, n, NR, S->d_perm, (const double *)Dd, Xh, Xl, xgh,
                               xgl, act_o);
            timer_start(ctx);
            hipLaunchKernelGGL(residual_kernel<NR>, grow, blk, 0, st, n, S->d_perm, P->d_rowptr, P->d_col, K->d_val, xgh,
                               xgl, Bp, R);
            timer_stop(ctx, FAM_SPMV, 1);
            JFEM_CHECK_HIP(ctx, norms(R, rn));
            n_active = 0;
            for (int r = 0; r < NR; ++r) {
                rn[r] = std::sqrt(rn[r]);
                if (!act[r]) continue;
                const double rr = rn[r] / std::max(bn[r], 1e-30);
                relres[r] = rr;
                iters[r] = nsolve;
                if (rr <= tol || max_refine == 0) { act[r] = 0; continue; }
                if (!(rr < 0.5 * prev[r])) { act[r] = 0; continue; }     // no longer converging
                prev[r] = rr;
                ++n_active;
            }
        }
        for (int r = 0; r < NR; ++r)
            if (!(relres[r] <= tol)) all_ok = false;
        if (relres_out) {
            // residual of the double vector handed back (x_hi alone)
            hipLaunchKernelGGL(residual_kernel<NR>, grow, blk, 0, st, n, S->d_perm, P->d_rowptr, P->d_col, K->d_val, xgh,
                               (const double *)nullptr, Bp, R);
            JFEM_CHECK_HIP(ctx, norms(R, rn));
            for (int r = 0; r < NR; ++r) relres_out[r] = std::sqrt(rn[r]) / std::max(bn[r], 1e-30);
        }
    }
    hipLaunchKernelGGL(export_kernel, gg, blk, 0, st, ndof, NR, xgh, x_dev);
    if (xlo_dev) hipLaunchKernelGGL(export_kernel, gg, blk, 0, st, ndof, NR, xgl, xlo_dev);
    int32_t aborted = 0;
    JFEM_CHECK_HIP(ctx, hipGetLastError());
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(&aborted, cnt + H.ns, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(st));
    if (aborted) return fail(ctx, JFEM_ERR_HIP, "triangular solve: a dependency wait timed out");
    if (!all_ok) return fail(ctx, JFEM_ERR_NOT_CONVERGED, "refinement did not reach the requested residual");
    return JFEM_OK;
}

}  // namespace

// ---------------------------------------------------------------- allocation-free solve
// Solve scratch of jfem_solve_dev_static: one array per GPU in device memory of
// the code object, held until the process ends.  In host code `scratch` names a
// host-side stand-in, never device memory: its device address is looked up once
// per device (hipGetSymbolAddress with that device current) and only that address
// reaches kernels.  Every call on a device uses the same array and leaves its
// contents behind; each solve clears what it reads first (pending marks, tickets,
// the abort flag, the iterate), and calls that overlap on one device race.
constexpr size_t kStaticScratchBytes = size_t(128) << 20;
static __device__ __align__(256) char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char *g_scratch_ptr[kMaxDevices] = {};

int static_scratch_ptr(jfem_ctx *ctx, char **out) {
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices) return fail(ctx, JFEM_ERR_BAD_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        int prev = 0;
        JFEM_CHECK_HIP(ctx, hipGetDevice(&prev));
        JFEM_CHECK_HIP(ctx, hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return fail(ctx, JFEM_ERR_HIP, std::string("static solve scratch: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (char *)p;
    }
    *out = g_scratch_ptr[dev];
    return JFEM_OK;
}
}  // namespace

extern "C" {

static int solve_dev_impl(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs, const double *b_dev,
                          double *x_dev, double tol, int max_refine, int32_t *iters, double *relres, double *xlo_dev,
                          double *relres_out, char *static_base) {
    int rc = JFEM_OK;
#define JFEM_SOLVE_CALL(N)                                                                                      \
    rc = F->precision == 64 ? solve_refine<N, double>(ctx, F, K, b_dev, x_dev, tol, max_refine, iters, relres,   \
                                                      xlo_dev, relres_out, static_base)                         \
                            : solve_refine<N, float>(ctx, F, K, b_dev, x_dev, tol, max_refine, iters, relres,    \
                                                     xlo_dev, relres_out, static_base)
    JFEM_NR_SWITCH(nrhs, JFEM_SOLVE_CALL)
#undef JFEM_SOLVE_CALL
    return rc;
}

static int check_solve_args(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs, int max_refine) {
    if (K->P != F->S->P) return fail(ctx, JFEM_ERR_BAD_ARG, "matrix and factor use different patterns");
    if (nrhs < 1 || nrhs > 16) return fail(ctx, JFEM_ERR_BAD_ARG, "nrhs must be in 1..16");
    if (max_refine < 0) return fail(ctx, JFEM_ERR_BAD_ARG, "max_refine must be >= 0");
    return JFEM_OK;
}

int jfem_solve_dev(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs, const double *b_dev,
                   double *x_dev, double tol, int max_refine, int32_t *iters, double *relres, double *xlo_dev,
                   double *relres_out) {
    clear_stale_hip_error();
    const int rc = check_solve_args(ctx, F, K, nrhs, max_refine);
    if (rc != JFEM_OK) return rc;
    return solve_dev_impl(ctx, F, K, nrhs, b_dev, x_dev, tol, max_refine, iters, relres, xlo_dev, relres_out, nullptr);
}

int jfem_solve_dev_static(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs, const double *b_dev,
                          double *x_dev, double tol, int max_refine, int32_t *iters, double *relres, double *xlo_dev,
                          double *relres_out) {
    clear_stale_hip_error();
    int rc = check_solve_args(ctx, F, K, nrhs, max_refine);
    if (rc != JFEM_OK) return rc;
    const int64_t need = jfem_solve_scratch_bytes(F->S, K->P->ndof, nrhs, F->precision);
    if (need > (int64_t)kStaticScratchBytes)
        return fail(ctx, JFEM_ERR_BAD_ARG,
                    "solve needs " + std::to_string(need) + " bytes of scratch; the static scratch holds " +
                        std::to_string((long long)kStaticScratchBytes));
    char *base = nullptr;
    rc = static_scratch_ptr(ctx, &base);
    if (rc != JFEM_OK) return rc;
    return solve_dev_impl(ctx, F, K, nrhs, b_dev, x_dev, tol, max_refine, iters, relres, xlo_dev, relres_out, base);
}

// Device self-test of the matrix-core trailing update against its scalar form on
// one random 200 x 70 front (three panels, L and U parts): returns the largest
// relative difference.  precision = 64 | 32.
int64_t jfem_symbolic_device_bytes(const jfem_symbolic *S) { return S ? S->bytes : 0; }

int64_t jfem_solve_scratch_bytes(const jfem_symbolic *S, int64_t ndof, int nrhs, int precision) {
    if (!S || nrhs < 1 || nrhs > 16) return 0;
    return (int64_t)SolveLayout(S->H, ndof, nrhs, precision == 64 ? 8 : 4).bytes;
}

int jfem_selftest_update(jfem_ctx *ctx, int precision, double *max_rel_diff) {
    const int64_t m = 200, n = 70;
    std::vector<UpdateTask> tasks;
    for (int64_t c0 = 64; c0 < m; c0 += 64)
        for (int64_t r0 = c0; r0 < m; r0 += 64) tasks.push_back({0, 0, (int32_t)r0, (int32_t)c0});
    // boundaries split at n: use the same rule as the symbolic task builder
    tasks.clear();
    std::vector<int32_t> bnd;
    for (int64_t c = 64; c < n; c += 64) bnd.push_back((int32_t)c);
    for (int64_t c = n; c < m; c += 64) bnd.push_back((int32_t)c);
    for (size_t ci = 0; ci < bnd.size(); ++ci)
        for (size_t ri = ci; ri < bnd.size(); ++ri) tasks.push_back({0, 0, bnd[ri], bnd[ci]});
    const int64_t nL = m * n, nU = (m - n) * (m - n);
    std::vector<double> hL(nL), hU(nU);
    uint64_t st = 0x9e3779b97f4a7c15ull;
    auto rnd = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (double)(st % 2000001) / 1e6 - 1.0; };
    for (auto &v : hL) v = rnd();
    for (auto &v : hU) v = rnd();
    const size_t es = precision == 64 ? 8 : 4;
    int64_t h_ncol = n, h_m = m, h_lofs = 0, h_uofs = 0;
    void *dL[2], *dU[2];
    int64_t *d_meta;
    UpdateTask *d_tasks;
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_meta, 4 * sizeof(int64_t)));
    int64_t meta[4] = {h_ncol, h_m, h_lofs, h_uofs};
    JFEM_CHECK_HIP(ctx, hipMemcpy(d_meta, meta, sizeof(meta), hipMemcpyHostToDevice));
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_tasks, tasks.size() * sizeof(UpdateTask)));
    JFEM_CHECK_HIP(ctx, hipMemcpy(d_tasks, tasks.data(), tasks.size() * sizeof(UpdateTask), hipMemcpyHostToDevice));
    std::vector<float> fL(hL.begin(), hL.end()), fU(hU.begin(), hU.end());
    for (int v = 0; v < 2; ++v) {
        JFEM_CHECK_HIP(ctx, hipMalloc(&dL[v], nL * es));
        JFEM_CHECK_HIP(ctx, hipMalloc(&dU[v], nU * es));
        JFEM_CHECK_HIP(ctx, hipMemcpy(dL[v], precision == 64 ? (void *)hL.data() : (void *)fL.data(), nL * es, hipMemcpyHostToDevice));
        JFEM_CHECK_HIP(ctx, hipMemcpy(dU[v], precision == 64 ? (void *)hU.data() : (void *)fU.data(), nU * es, hipMemcpyHostToDevice));
    }
    dim3 g((unsigned)tasks.size());
    if (precision == 64) {
        hipLaunchKernelGGL(update_kernel<double>, g, dim3(kBS), 0, ctx->stream, d_tasks, d_meta, d_meta + 1, d_meta + 2, d_meta + 3, d_meta + 2, (double *)dL[0], (double *)dU[0], (double *)nullptr, 0, d_meta, ctx->d_status, VarStride{}, (double *)nullptr);
        hipLaunchKernelGGL(update_ref_kernel<double>, g, dim3(kBS), 0, ctx->stream, d_tasks, d_meta, d_meta + 1, d_meta + 2, d_meta + 3, (double *)dL[1], (double *)dU[1]);
    } else {
        hipLaunchKernelGGL(update_kernel<float>, g, dim3(kBS), 0, ctx->stream, d_tasks, d_meta, d_meta + 1, d_meta + 2, d_meta + 3, d_meta + 2, (float *)dL[0], (float *)dU[0], (float *)nullptr, 0, d_meta, ctx->d_status, VarStride{}, (float *)nullptr);
        hipLaunchKernelGGL(update_ref_kernel<float>, g, dim3(kBS), 0, ctx->stream, d_tasks, d_meta, d_meta + 1, d_meta + 2, d_meta + 3, (float *)dL[1], (float *)dU[1]);
    }
    JFEM_CHECK_HIP(ctx, hipGetLastError());
    JFEM_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double worst = 0, scale = 0;
    for (int part = 0; part < 2; ++part) {
        int64_t cnt = part == 0 ? nL : nU;
        std::vector<double> a(cnt), r(cnt);
        if (precision == 64) {
            JFEM_CHECK_HIP(ctx, hipMemcpy(a.data(), part ? dU[0] : dL[0], cnt * 8, hipMemcpyDeviceToHost));
            JFEM_CHECK_HIP(ctx, hipMemcpy(r.data(), part ? dU[1] : dL[1], cnt * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> fa(cnt), fr(cnt);
            JFEM_CHECK_HIP(ctx, hipMemcpy(fa.data(), part ? dU[0] : dL[0], cnt * 4, hipMemcpyDeviceToHost));
            JFEM_CHECK_HIP(ctx, hipMemcpy(fr.data(), part ? dU[1] : dL[1], cnt * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < cnt; ++i) { a[i] = fa[i]; r[i] = fr[i]; }
        }
        for (int64_t i = 0; i < cnt; ++i) {
            worst = std::max(worst, std::fabs(a[i] - r[i]));
            scale = std::max(scale, std::fabs(r[i]));
        }
    }
    for (int v = 0; v < 2; ++v) { (void)hipFree(dL[v]); (void)hipFree(dU[v]); }
    (void)hipFree(d_meta);
    (void)hipFree(d_tasks);
    *max_rel_diff = worst / std::max(scale, 1e-300);
    return JFEM_OK;
}

int jfem_solve(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs, const double *b_host, double *x_host,
               double tol, int max_refine, int32_t *iters, double *relres, double *xlo_host, double *relres_out) {
    clear_stale_hip_error();
    const int64_t ndof = K->P->ndof;
    const int64_t cnt = std::max<int64_t>(ndof * nrhs, 1);
    double *d_b = nullptr, *d_x = nullptr, *d_xl = nullptr;
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_b, cnt * 8));
    JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_x, cnt * 8));
    if (xlo_host) JFEM_CHECK_HIP(ctx, hipMalloc((void **)&d_xl, cnt * 8));
    JFEM_CHECK_HIP(ctx, hipMemcpyAsync(d_b, b_host, ndof * nrhs * 8, hipMemcpyHostToDevice, ctx->stream));
    int st = jfem_solve_dev(ctx, F, K, nrhs, d_b, d_x, tol, max_refine, iters, relres, d_xl, relres_out);
    hipError_t e = hipMemcpyAsync(x_host, d_x, ndof * nrhs * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && xlo_host) e = hipMemcpyAsync(xlo_host, d_xl, ndof * nrhs * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_b);
    (void)hipFree(d_x);
    if (d_xl) (void)hipFree(d_xl);
    if (e != hipSuccess) return fail(ctx, JFEM_ERR_HIP, hipGetErrorString(e));
    return st;
}

}  // extern "C"

namespace jfem {
void batch_free_factor(jfem_batch *B) {
    for (jfem_factor *F : B->fviews) delete F;
    B->fviews.clear();
    for (void *p : {B->d_L, B->d_U, B->d_Linv})
        if (p) (void)hipFree(p);
    B->d_L = B->d_U = B->d_Linv = nullptr;
    B->S = nullptr;
    B->precision = 0;
}
}  // namespace jfem
