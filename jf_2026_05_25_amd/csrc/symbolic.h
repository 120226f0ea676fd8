// Symbolic analysis for the supernodal multifrontal Cholesky of K_ff.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace jfem {

constexpr int kPanel = 64;   // panel width of the blocked partial factorization
constexpr int kTile = 64;    // tile edge of panel/update tasks

struct PanelTask { int32_t s, k, r0, pad; };          // supernode, panel column, first row
struct UpdateTask { int32_t s, k, r0, c0; };          // supernode, panel column, tile origin
// Tile (r0, c0) of the update matrix U_s (front rows / columns >= ns) pushed into
// the parent's front: U_s - L[r0.., 0:ns] L[c0.., 0:ns]' added to the parent's
// entries at the relmap positions (pivot columns of the parent in its L, the rest
// in its U).  The supernode's and its parent's shapes and offsets are inlined.
struct PushTask {
    int32_t s, r0, c0, ns, ms, nsp, msp, pad;
    int64_t lofs, uofs, relofs, lofsp, uofsp;
};
// Dataflow solve task: one workgroup each; a list is in dependency order, so a
// task waits only on tasks listed before it.  The supernode's metadata is
// inlined so that a task starts without dependent loads.
// Forward, kind 0: p >= 0: pivot rows [r0, r0+nr) (panel p) of the front of s;
//   p < 0: update rows [r0, r0+nr).  Kind 1: the whole front of a supernode with
//   one panel and at most kWholeRows rows (pivot and update rows in one task).
// Backward: columns [r0, r0+nr) (panel p) of s; update rows [u0, u1) = [ns, ms).
// A block sum over more than kSolveSplit panels (kSolveSplitRows 64-row tiles of
// update rows) is cut into chunks of at most kSolveChunk blocks: a task streams
// its own chunk and adds the npart partial sums (64 rows x nrhs each, partial
// slots [pofs, pofs + npart)) that the partial tasks publish, so the blocks are
// streamed side by side and the consumer waits one exchange for the rest.
// Forward: kind 0 pivot (p >= 0) sums X[p, k] v_k, update (p < 0) L[rows, k] z_k,
//   over k in [k0, k1) plus the partials of [0, k0); kind 2: the partial sum over
//   [k0, k1) of the task that follows, into slot pofs.
// Backward: kind 0 panel task: t_p over the update rows [u0, u1) plus the
//   partials of [u1, ms), then x_p over the later panels q in [k0, k1) (= [p+1, np));
//   a panel with more than kSolveSplit later panels is split: kind 3 forms and
//   publishes t_p only, kind 4 forms x_p from X_pp t_p, q in [k0, k1) and the
//   partials of [k1, np).  Kind 2: a partial sum, over update rows [u0, u1) (of a
//   kind 0 or 3 task, listed just before it) or over panels q in [k0, k1) (of a
//   kind 4 task).
constexpr int kWholeRows = 256;  // front rows of a whole-front forward task
// (the host-side task-list tests also build with smaller values)
#ifndef JFEM_SOLVE_CHUNK
#define JFEM_SOLVE_CHUNK 4
#endif
#ifndef JFEM_SOLVE_SPLIT
#define JFEM_SOLVE_SPLIT 4
#endif
#ifndef JFEM_SOLVE_SPLIT_ROWS
#define JFEM_SOLVE_SPLIT_ROWS 16
#endif
constexpr int kSolveChunk = JFEM_SOLVE_CHUNK;
constexpr int kSolveSplit = JFEM_SOLVE_SPLIT;
constexpr int kSolveSplitRows = JFEM_SOLVE_SPLIT_ROWS;
// first block of a sum over [0, nb) that its own task streams (the rest in
// partials): nb - kSolveChunk if nb exceeds the split threshold, else 0
inline int32_t solve_own_first(int32_t nb, int32_t split) { return nb > split ? nb - kSolveChunk : 0; }
struct FlowTask {
    int32_t s, p, r0, nr;
    int32_t np, need;               // panels of s; update tasks of s's children (forward)
    int32_t parent, pnp;            // parent supernode (-1 for a root) and its panels
    int32_t nch, kind;              // children of s; task kind
    int32_t u0, u1;                 // backward: update rows [u0, u1) of the front
    int32_t k0, k1;                 // block range of the task's own sum
    int32_t npart, pofs;            // partial sums added (consumer) / its slot (kind 2)
    int64_t ns, ms, lofs, c0, relofs, rofs;
    int64_t iofs;                   // inverse of the diagonal block of panel p
    // supernode with np >= 2 panels (pivot and backward tasks; -1 otherwise): the
    // offset in the inverse array of X = L_11^{-1}, ns x ns column-major, of which
    // the solves read the lower blocks (the upper ones are the inversion's scratch)
    int64_t xofs;
    int64_t gb;                     // forward: first FlowChild entry
};
// per (forward task, child of its supernode), in child order: the child's update
// rows [i0, i1) that land in the task's rows, and the child's update-vector offset
struct FlowChild {
    int64_t relofs;
    int32_t i0, i1;
};
// Inversion of L_11 for a supernode with np >= 2 panels, by doubling: at block
// size B panels, for each pair of consecutive diagonal superblocks A = [a, a+B),
// C = [a+B, min(a+2B, np)) (X_AA and X_CC already inverted),
//   phase 0: T = L_CA X_AA,  T[i,j] = sum_{k=j}^{a+B-1} L[i,k] X[k,j],
//   phase 1: X_CA = -X_CC T, X[i,j] = -sum_{k=a+B}^{i} X[i,k] T[k,j],
// one task per 64 x 64 output tile (i in C, j in A, panel indices), k over
// [k0, k1).  T[i,j] is kept transposed in the upper blocks of X (at rows of
// panel j, columns of panel i), which the solves never read.
struct InvTask {
    int64_t lofs, xofs;
    int32_t ns, ms, i, j, k0, k1;
};

struct SymbolicHost {
    int64_t ndof = 0, n = 0;        // all DOFs, free DOFs
    int64_t ns = 0;                 // supernodes
    std::vector<int32_t> perm;      // elimination position -> global DOF
    std::vector<int32_t> iperm;     // global DOF -> elimination position (-1 fixed)
    std::vector<int64_t> col0, ncol, m, lofs, uofs, rofs, relofs;
    std::vector<int32_t> parent;    // -1 for roots
    std::vector<int32_t> rows;      // concatenated row lists (elimination positions)
    std::vector<int32_t> relmap;    // per child: update row -> parent front row
    std::vector<int32_t> child_ptr, child;   // children lists (ascending)
    std::vector<int32_t> level_ptr, level_sn;
    // A scatter (upper triangle of K in original numbering -> lower front)
    std::vector<int32_t> amap_src;
    std::vector<int64_t> amap_dst;
    // per-level task lists
    std::vector<PanelTask> ptasks;
    std::vector<UpdateTask> utasks;
    // update matrices pushed into the parents after each level's last panel step:
    // the level's supernodes in launches by sibling rank (the r-th child of each
    // parent among the level's supernodes goes to launch r), so no two tasks of a
    // launch add into the same parent entry; every entry of a front receives its
    // children's updates in a fixed order (child level, then rank)
    std::vector<PushTask> stasks;
    std::vector<int64_t> spush_off;         // per push launch: offset into stasks (size launches+1)
    std::vector<int64_t> slev_ptr;          // per level: first push launch (size levels+1)
    // ptask/utask ranges per (level, step): offsets into the arrays
    std::vector<int64_t> step_ptr;          // per level: first step index (size levels+1)
    std::vector<int64_t> pstep_off, ustep_off;   // per step: offset (size steps+1)
    std::vector<int64_t> pstep_mid;              // per step: first non-diagonal panel tile
    std::vector<int32_t> pstep_bw;               // per panel step: widest diagonal block
    // dataflow solves: forward list (per supernode its whole front, or its panels
    // and then its update blocks of kTile rows), backward list (per supernode its
    // panels from the last), both ordered by estimated start; per supernode the
    // offset of its inverse diagonal blocks, written by the factorization (panel
    // p: b x b entries at iofs[s] + kPanel^2 p)
    std::vector<FlowTask> flow_f, flow_b;
    // tasks of the bottom subtrees (supernodes with one panel and at most kWholeRows
    // front rows whose children are all such): the first flow_f_leaf entries of
    // flow_f and the last flow_b_leaf entries of flow_b, run by the one-wavefront
    // leaf kernels (forward before, backward after the main launch)
    int64_t flow_f_leaf = 0, flow_b_leaf = 0;
    int64_t flow_f_parts = 0, flow_b_parts = 0;   // partial-sum slots of the two lists
    std::vector<FlowChild> flow_gb;
    std::vector<int64_t> iofs;
    int64_t inv_entries = 0;        // diagonal-block inverses, then the X of the multi-panel supernodes (xofs)
    // inversion tasks: per doubling level, phase 0 then phase 1 (istep_off: one
    // offset per (level, phase), plus the end)
    std::vector<InvTask> itasks;
    std::vector<int64_t> istep_off;
    double inv_flops = 0;      // 2 x multiply-adds of the inversion tiles (full 64 x 64 x 64 products)
    int64_t L_entries = 0, U_entries = 0, upd_rows = 0;
    double flops = 0;
    double update_flops = 0;   // 2 x multiply-adds of the per-step tiles (lower triangles)
    double schur_flops = 0;    // 2 x multiply-adds of the update-matrix tiles (lower triangles)
    int64_t max_front = 0;
};

// Build the symbolic structure.  rowptr/col: CSR pattern of the full K;
// xyz: grid coordinates (3 per node); fixed: ndof mask.
int build_symbolic(int64_t ndof, const int64_t *rowptr, const int32_t *col, const double *xyz,
                   const uint8_t *fixed, SymbolicHost &S, std::string &err);

}  // namespace jfem
