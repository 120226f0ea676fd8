/*
 * jfem_hip.h -- C ABI of the MI355X (gfx950) implementation of OpenJFEM's
 * linear-static hot path: batched element stiffness (CQUAD4, CTRIA3, CTETRA10),
 * deterministic scatter-assembly into a device-resident CSR matrix, and a
 * supernodal sparse Cholesky factor (FP64, or FP32 with FP64 iterative
 * refinement) with its triangular solves.
 *
 * Conventions
 *   - plain pointers and sizes only; host pointers unless a name says "dev";
 *   - matrices column-major (Julia layout), 0-based indices;
 *   - grid point i owns global DOFs 6*i .. 6*i+5 (u, v, w, rx, ry, rz), i = rank of
 *     the GRID ID (src/solver/assembly.jl:2653-2657);
 *   - connectivity is column-major n_elem x nn: conn[k*n_elem + e];
 *   - shell section records are column-major n_elem x 8 (JFEM_SP_*);
 *   - every entry point returns a jfem_status; jfem_last_error(ctx) has the text.
 *   - the library owns all device memory behind opaque handles; handles may be used
 *     from any context of the same device; calls on one context are ordered on that
 *     context's HIP stream; different contexts may run concurrently.
 */
#ifndef JFEM_HIP_H
#define JFEM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    JFEM_OK = 0,
    JFEM_ERR_HIP = 1,             /* HIP runtime error */
    JFEM_ERR_DEGENERATE = 2,      /* degenerate/folded shell: ArgumentError, src/FEMKernels.jl:2862-2867, 4244-4251 */
    JFEM_ERR_INVERTED = 3,        /* non-positive solid Jacobian: DomainError, src/tetra10.jl:31-35 */
    JFEM_ERR_WARPED = 4,          /* CQUAD4 outside the flat-shell differential-stiffness operator (corner
                                     deviation >= 1e-2 of the longer diagonal, src/solver/assembly.jl:6575-6594);
                                     the stiffness path handles warped cells by the finite-warp congruence */
    JFEM_ERR_BAD_ARG = 5,         /* bad sizes / material: ArgumentError, src/tetra10.jl:55-58 */
    JFEM_ERR_NOT_SPD = 6,         /* factorization hit a non-positive pivot: PosDefException */
    JFEM_ERR_OOM = 7,             /* device allocation failed */
    JFEM_ERR_NOT_CONVERGED = 8    /* refinement did not reach the requested residual */
} jfem_status;

/* shell section record columns (per element) */
enum { JFEM_SP_H = 0, JFEM_SP_E = 1, JFEM_SP_NU = 2, JFEM_SP_G = 3, JFEM_SP_BEND_RATIO = 4,
       JFEM_SP_TS_T = 5, JFEM_SP_RIGID_SHEAR = 6, JFEM_SP_MEMBRANE_MODES = 7, JFEM_SP_NCOL = 8 };

enum { JFEM_KIND_CQUAD4 = 0, JFEM_KIND_CTRIA3 = 1, JFEM_KIND_CTETRA10 = 2 };
enum { JFEM_FP64 = 64, JFEM_FP32 = 32 };

typedef struct jfem_ctx jfem_ctx;
typedef struct jfem_mesh jfem_mesh;
typedef struct jfem_pattern jfem_pattern;
typedef struct jfem_matrix jfem_matrix;
typedef struct jfem_symbolic jfem_symbolic;
typedef struct jfem_factor jfem_factor;
typedef struct jfem_loadset jfem_loadset;
typedef struct jfem_batch jfem_batch;

/* ---- contexts ----------------------------------------------------------- */
/* One per host thread or request slot: a HIP stream on `device` plus scratch.
 * (No reference counterpart: the reference is single-process Julia.) */
int jfem_ctx_create(int device, jfem_ctx **out);
int jfem_ctx_destroy(jfem_ctx *ctx);
int jfem_ctx_synchronize(jfem_ctx *ctx);
const char *jfem_last_error(jfem_ctx *ctx);
/* free / total device memory in bytes (admission control) */
int jfem_device_memory(jfem_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);
/* milliseconds accumulated by the named kernel family since the last reset
 * (timed with HIP events on the context's stream); family: 0 element generation,
 * 1 assembly, 2 numeric factor, 3 triangular solves, 4 residual SpMV; with timing
 * level 2 also per launch: 5 factor update kernel, 6 factor panel kernels,
 * 7 (unused), 8 forward dataflow solve kernel, 9 backward dataflow solve
 * kernel, 10 update-matrix kernel (Schur complement pushed into the parent
 * fronts), 11 diagonal-block
 * inverses of the solves (the 64 x 64 inverses and X = L_11^{-1} of the
 * supernodes with two or more panels, one interval for the whole sequence).
 * jfem_ctx_enable_timing(ctx, 0 | 1 | 2). */
int jfem_ctx_kernel_time(jfem_ctx *ctx, int family, double *ms, int64_t *launches);
int jfem_ctx_reset_kernel_time(jfem_ctx *ctx);
int jfem_ctx_enable_timing(jfem_ctx *ctx, int on);

/* ---- element kernels (batched FEM.stiffness_* over a device mesh) -------- */
/* Replaces FEM.stiffness_quad4_matrices (src/FEMKernels.jl:2839),
 * FEM.stiffness_tria3_matrices (:9811) and FEM.stiffness_tetra10
 * (src/tetra10.jl:68) as called from assemble_stiffness
 * (src/solver/assembly.jl:5220-5244, 5474, 5842) INCLUDING the per-element frame,
 * section and T'KT transform (src/solver/assembly.jl:359-440, 5484-5499, 5858-5870):
 * writes every element's GLOBAL-frame matrix (24x24 / 18x18 / 30x30 column-major,
 * element-major) of the given kind into host buffer Ke_out.  bad_element receives
 * the first failing element index (or -1). */
int jfem_element_matrices(jfem_ctx *ctx, const jfem_mesh *mesh, int kind,
                          double *Ke_out, int64_t *bad_element);

/* Local-frame CQUAD4 kernel over n elements, the direct counterpart of
 * stiffness_quad4_matrices(coords, Cm, Cb, Cs, h, E_ref; k6rot, drill_scale, ...)
 * (src/FEMKernels.jl:2839).  lc: n x 8 (x1,y1,...,x4,y4 per element, element-major),
 * Cm, Cb: n x 9, Cs: n x 4 (each element's matrix column-major, element-major);
 * h, k6rot, drill_scale: n; flags: n (bit0 rigid shear, bits 4-7 membrane mode mask,
 * 0 = all four).  Ke_out: n x 576. */
int jfem_q4_local_batch(jfem_ctx *ctx, int64_t n, const double *lc, const double *Cm,
                        const double *Cb, const double *Cs, const double *h,
                        const double *k6rot, const double *drill_scale, const int32_t *flags,
                        double *Ke_out, int64_t *bad_element);

/* Element-frame CTRIA3 kernel over n elements, the counterpart of
 * stiffness_tria3_matrices(coords, Dm, Db, Ds, h, G_ref; bend_ratio, k6rot)
 * (src/FEMKernels.jl:9811 -> stiffness_tria3_matrices_generic :9375-9809, default
 * branches: CST membrane, constant-curvature bending, MITC3 shear with residual
 * bending flexibility and its guard, lumped drilling).  lc: n x 6 (x1,y1,x2,y2,x3,y3
 * per element); Dm, Db: n x 9, Ds: n x 4 (each element's matrix column-major);
 * h, G_ref, bend_ratio, k6rot: n (h, G_ref and bend_ratio are accepted for the
 * reference's signature; its default branch reads none of them).  Ke_out: n x 324
 * (18 x 18 column-major per element; DOFs u,v,w,rx,ry,rz per node). */
int jfem_t3_local_batch(jfem_ctx *ctx, int64_t n, const double *lc, const double *Dm,
                        const double *Db, const double *Ds, const double *h, const double *G_ref,
                        const double *bend_ratio, const double *k6rot, double *Ke_out,
                        int64_t *bad_element);

/* CTETRA10 kernel over n elements, the counterpart of stiffness_tetra10(coords, E, nu)
 * (src/tetra10.jl:68-79): X: n x 30 (each element's 10 x 3 coordinates column-major,
 * corners then edges (1,2),(2,3),(1,3),(1,4),(2,4),(3,4)), E, nu: n.  Ke_out: n x 900
 * (30 x 30 column-major, Matrix(Symmetric(K))).  A non-positive Jacobian at a
 * sample point returns JFEM_ERR_INVERTED (DomainError, src/tetra10.jl:31-35), bad E
 * or nu JFEM_ERR_BAD_ARG (ArgumentError, :55-58); bad_element is the first such. */
int jfem_tet10_batch(jfem_ctx *ctx, int64_t n, const double *X, const double *E, const double *nu,
                     double *Ke_out, int64_t *bad_element);

/* ---- mesh --------------------------------------------------------------- */
/* Upload a model (coordinates SoA, connectivity, sections) to the device.
 * Replaces the per-call model Dict walk of assemble_stiffness
 * (src/solver/assembly.jl:2653-2674, 3224-3473). */
int jfem_mesh_create(jfem_ctx *ctx, int64_t n_nodes, const double *x, const double *y,
                     const double *z, int64_t n_quad, const int32_t *quad_conn,
                     const double *quad_prop, int64_t n_tria, const int32_t *tria_conn,
                     const double *tria_prop, int64_t n_tet, const int32_t *tet_conn,
                     const double *tet_prop, double k6rot, jfem_mesh **out);
/* Replace the section records (same topology): design variants. */
int jfem_mesh_update_props(jfem_ctx *ctx, jfem_mesh *mesh, const double *quad_prop,
                           const double *tria_prop, const double *tet_prop);
int jfem_mesh_free(jfem_mesh *mesh);

/* ---- pattern + assembly -------------------------------------------------- */
/* Symbolic assembly: the CSR pattern of the global K (all 6*n_nodes DOFs) and the
 * element-DOF -> CSR-slot gather lists; the reference analogue is the slot-map
 * CSC refill (src/solver/assembly.jl:5975-6033).  Depends on topology only. */
int jfem_pattern_create(jfem_ctx *ctx, const jfem_mesh *mesh, jfem_pattern **out, int64_t *nnz);
int jfem_pattern_free(jfem_pattern *pattern);

/* Numeric assembly: element generation + deterministic segmented reduction.
 * Replaces assemble_stiffness -> sparse(I,J,V) (src/solver/assembly.jl:2630, 5942):
 * duplicates are summed in triplet order (CQUAD4s, then CTRIA3s, then CTETRA10s,
 * each in element order), so values are reproducible bit for bit run to run. */
int jfem_assemble(jfem_ctx *ctx, const jfem_mesh *mesh, const jfem_pattern *pattern,
                  jfem_matrix **out, int64_t *bad_element);
/* re-assemble into an existing matrix (same pattern), e.g. after update_props */
int jfem_reassemble(jfem_ctx *ctx, const jfem_mesh *mesh, jfem_matrix *K, int64_t *bad_element);
int jfem_matrix_download(jfem_ctx *ctx, const jfem_matrix *K, int64_t *row_ptr, int32_t *col,
                         double *val);
int jfem_matrix_free(jfem_matrix *K);

/* ---- design variants ------------------------------------------------------ */
/* nvar design variants of one topology: the mesh's grid points and connectivity,
 * each variant with its own section records (quad_props: nvar x (n_quad x 8),
 * variant after variant, each in the layout of jfem_mesh_create; likewise
 * tria_props and tet_props).  The reference runs each variant as its own model
 * through assemble_stiffness and cholesky (src/solver/assembly.jl:2630,
 * src/solver/boundary_conditions.jl:1660); here all variants share the pattern
 * and one symbolic factorization, and their element generation, assembly and
 * numeric factorization each run as one launch sequence over all variants. */
int jfem_batch_create(jfem_ctx *ctx, const jfem_mesh *mesh, const jfem_pattern *pattern, int64_t nvar,
                      const double *quad_props, const double *tria_props, const double *tet_props,
                      jfem_batch **out);
int jfem_batch_update_props(jfem_ctx *ctx, jfem_batch *batch, const double *quad_props,
                            const double *tria_props, const double *tet_props);
/* element generation + assembly of every variant (bitwise the same K values as
 * jfem_assemble of that variant alone); on a rejected element bad_variant and
 * bad_element (index in triplet order) name the first one */
int jfem_batch_assemble(jfem_ctx *ctx, jfem_batch *batch, int64_t *bad_variant, int64_t *bad_element);
/* variant v's matrix (owned by the batch; jfem_matrix_free ignores it), usable
 * wherever a jfem_matrix is: BC partition, symbolic, factor, solves */
int jfem_batch_matrix(const jfem_batch *batch, int64_t v, jfem_matrix **view);
/* numeric factors of all variants over S (they must share its free set), in one
 * launch sequence, each bitwise the factor jfem_factor_create gives for that
 * variant alone; a non-positive pivot names its variant and DOF */
int jfem_batch_factor(jfem_ctx *ctx, jfem_batch *batch, const jfem_symbolic *S, int precision,
                      int64_t *bad_variant, int64_t *bad_column);
/* variant v's factor (owned by the batch; jfem_factor_free ignores it), for the
 * solve entry points */
int jfem_batch_factor_view(const jfem_batch *batch, int64_t v, jfem_factor **view);
int jfem_batch_free(jfem_batch *batch);

/* ---- boundary conditions -------------------------------------------------- */
/* SPC + AUTOSPC (zero/negative diagonal) + grid-point singularity processing:
 * src/solver/boundary_conditions.jl:1480-1607, 1028-1138, 745-845.
 * spc: ndof (1 = SPC1); fixed_out: ndof; n_free: number of free DOFs. */
int jfem_bc_partition(jfem_ctx *ctx, const jfem_matrix *K, const uint8_t *spc,
                      double eps_zero, uint8_t *fixed_out, int64_t *n_free);

/* ---- shell recovery -------------------------------------------------------- */
/* Force / stress / strain recovery of a mesh's isotropic PSHELL CQUAD4 and CTRIA3
 * elements from one displacement vector, the counterpart of recover_shell_stresses!
 * (src/solver/stress_recovery.jl:505; called from src/solver/solve_case.jl:1204)
 * for the default (no offset, no SNORM, no thermal) branches.  u: ndof doubles in
 * global DOF order (6 per grid point, basic frame).  Per CQUAD4, q4_out holds 44
 * doubles (element frame): N[3] M[3] Q[2], stress at z = -h/2 [3] and +h/2 [3],
 * strain at -h/2 [3] and +h/2 [3], corner N[4][3], corner M[4][3] (corners in
 * connectivity order; M the printed moment).  Per CTRIA3, t3_out holds the first 20.
 * mean_shear != 0 selects the branch of solution sequences 101/105/106 (MacNeal
 * area-mean shear, corner moments from the section fields,
 * stress_recovery.jl:873-946); 0 that of other sequences (MITC4 centre shear
 * blended with the equilibrium shear of the corner moments, :948-952).
 * Either output may be NULL when the mesh has no element of its kind; a
 * degenerate element returns JFEM_ERR_DEGENERATE with bad_element its index
 * (CQUAD4s first, then CTRIA3s). */
int jfem_shell_recover(jfem_ctx *ctx, const jfem_mesh *mesh, const double *u, int mean_shear,
                       double *q4_out, double *t3_out, int64_t *bad_element);

/* ---- SOL 105: differential stiffness, shifted matrices, block products ------ */
/* Differential (geometric) stiffness of the shells of a mesh under the membrane
 * state of the preload displacements u (host, ndof doubles, global DOF order).
 * CQUAD4: the default operator of a flat isotropic CQUAD4,
 * geometric_stiffness_quad4_reference_kdjj_iso (src/FEMKernels.jl:8810-8942),
 * with the assembly's preload map, warp congruence and element-frame transform
 * (src/solver/assembly.jl:6985-6998, 8063-8075, 8264-8320).  A cell whose corners
 * leave the plane of its diagonals by 1e-2 of the longer diagonal or more is not
 * on this operator (assembly.jl:6575-6594): JFEM_ERR_WARPED, bad_element its
 * index.  CTRIA3: geometric_stiffness_tria3 in the default principal-transverse
 * mode (src/FEMKernels.jl:11142-11288, 8205-8338) under the membrane resultants
 * of stress_strain_tria3 (src/FEMKernels.jl:9859-9937), transformed to global
 * (src/solver/assembly.jl:8343-8498).  Meshes with CTETRA10 elements return
 * JFEM_ERR_BAD_ARG.  jfem_shell_kg_batch: per element the global matrix
 * (column-major, host): nq x 576 doubles of the CQUAD4s, then nt x 324 of the
 * CTRIA3s; bad_element counts in that order.  jfem_assemble_kg: the assembled Kg
 * on the pattern of K (assemble_geometric_stiffness, src/solver/assembly.jl:6106;
 * duplicates summed in element order), a jfem_matrix like K's. */
int jfem_shell_kg_batch(jfem_ctx *ctx, const jfem_mesh *mesh, const double *u, double *kg_out,
                        int64_t *bad_element);
int jfem_assemble_kg(jfem_ctx *ctx, const jfem_mesh *mesh, const jfem_pattern *pattern,
                     const double *u, jfem_matrix **out, int64_t *bad_element);
/* out = K + sigma * Kg (values of two matrices on one pattern): the shifted
 * matrix of the buckling pencil, factored with jfem_factor_create / jfem_refactor
 * on the symbolic analysis of K (the reference's shifted factorizations of
 * K + sigma Kg, src/solver/helpers.jl:105-143, 280-420).  *out == NULL allocates
 * a new matrix; otherwise *out (same pattern) is overwritten. */
int jfem_matrix_shift(jfem_ctx *ctx, const jfem_matrix *K, double sigma, const jfem_matrix *Kg,
                      jfem_matrix **out);
/* Inertia of A_ff on the symbolic analysis S of K (A on K's pattern, e.g. from
 * jfem_matrix_shift or jfem_matrix_shift_scaled): the unpivoted signed
 * factorization A_ff = L J L' (FP64, same launch sequence as the Cholesky),
 * the number of negative pivots (the eigenvalues of A_ff below zero, by
 * Sylvester's law of inertia) and the smallest and largest pivot magnitude.
 * A zero or non-finite pivot returns JFEM_ERR_NOT_SPD with bad_column.  The
 * reference counts buckling roots in an interval this way, with LDL' of the
 * diagonally scaled shifted matrices (_eigb_interval_count,
 * src/solver/helpers.jl:105-143). */
int jfem_factor_inertia(jfem_ctx *ctx, const jfem_symbolic *S, const jfem_matrix *A, int64_t *negative,
                        double *min_pivot, double *max_pivot, int64_t *bad_column);
/* out = D^{-1/2} (K + sigma Kg) D^{-1/2} with D the diagonal of K (entries with
 * K_ii <= 0 scaled by 1): the scaled shifted matrix of that interval count.
 * *out == NULL allocates; otherwise *out (same pattern) is overwritten. */
int jfem_matrix_shift_scaled(jfem_ctx *ctx, const jfem_matrix *K, double sigma, const jfem_matrix *Kg,
                             jfem_matrix **out);
/* Block products of the eigen-extraction, device blocks in full DOF numbering
 * (ndof x ncol column-major, at most 32 columns), sums in a fixed order:
 *   jfem_matrix_spmv:   y = A_ff x (fixed: host mask of ndof bytes, 1 = fixed, or
 *                       NULL; fixed rows of y are zero, fixed columns skipped);
 *   jfem_block_gram:    c_host (p x q column-major) = x' y over n rows;
 *   jfem_block_combine: x = y Q for a host p x q matrix Q (column-major), or
 *                       x += y Q when accumulate != 0. */
int jfem_matrix_spmv(jfem_ctx *ctx, const jfem_matrix *A, const uint8_t *fixed, int ncol,
                     const double *x_dev, double *y_dev);
int jfem_block_gram(jfem_ctx *ctx, int64_t n, int p, const double *x_dev, int q,
                    const double *y_dev, double *c_host);
int jfem_block_combine(jfem_ctx *ctx, int64_t n, int p, const double *y_dev, int q,
                       const double *Q_host, double *x_dev, int accumulate);

/* ---- factor / solve -------------------------------------------------------- */
typedef struct {
    int64_t n_free;          /* order of K_ff */
    int64_t n_supernodes;
    int64_t n_levels;        /* elimination-tree levels (batched launch groups) */
    int64_t nnz_L;           /* stored entries of L (supernodal, incl. explicit zeros) */
    double flops;            /* sum_j (eta_j + 1)^2 over columns of L (counted per column) */
    int64_t max_front;       /* largest front order */
    int64_t update_entries;  /* stored update-matrix entries */
    double update_flops;     /* multiply-adds x 2 of the per-step trailing-update tiles (pivot
                                columns of the fronts, lower triangles) */
    int64_t update_launches; /* update-kernel launches per factorization */
    int64_t inv_entries;      /* entries of the panels' inverse diagonal blocks (solves) */
    int64_t upd_rows;         /* update-vector rows over all supernodes */
    int64_t flow_fwd_tasks, flow_bwd_tasks;   /* workgroups of the two dataflow solve kernels */
    int64_t solve_entries;    /* factor entries one triangular solve reads: per panel its
                                 inverse diagonal block and the rows below that block */
    double schur_flops;       /* multiply-adds x 2 of the update-matrix tiles, each taken once
                                 with the full rank of its front (lower triangles) */
    int64_t schur_launches;   /* update-matrix launches per factorization (one per level) */
} jfem_symbolic_info;

/* Fill-reducing ordering (geometric nested dissection over grid points),
 * elimination tree, supernodes, front maps for K_ff = K[free, free].  Replaces the
 * symbolic half of cholesky(Symmetric(K_ff)) (src/solver/boundary_conditions.jl:1660).
 * xyz: n_nodes x 3 column-major (may be NULL to use the mesh coordinates). */
int jfem_symbolic_create(jfem_ctx *ctx, const jfem_pattern *pattern, const jfem_mesh *mesh,
                         const uint8_t *fixed, jfem_symbolic **out, jfem_symbolic_info *info);
int jfem_symbolic_free(jfem_symbolic *S);

/* Numeric supernodal Cholesky of K_ff in FP64 or FP32 (precision = 64 | 32).
 * Reads the upper triangle of K (Symmetric(K_ff) semantics).  On a non-positive
 * pivot returns JFEM_ERR_NOT_SPD with bad_column = the free-DOF index (global DOF).
 * Replaces _timed_factorization(cholesky, Symmetric(K_ff))
 * (src/solver/boundary_conditions.jl:1660,1709). */
int jfem_factor_create(jfem_ctx *ctx, const jfem_symbolic *S, const jfem_matrix *K,
                       int precision, jfem_factor **out, int64_t *bad_column);
/* refactor in place with new values (same symbolic, same precision) */
int jfem_refactor(jfem_ctx *ctx, jfem_factor *F, const jfem_matrix *K, int64_t *bad_column);
/* post-factor pivot check L_jj^2 / K_jj < threshold
 * (src/solver/boundary_conditions.jl:1662-1710): writes the global DOFs that fail
 * into dofs_out (capacity cap) and their count. */
int jfem_factor_small_pivots(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K,
                             double threshold, int64_t *dofs_out, int64_t cap, int64_t *count);
/* the ratios behind that check, per elimination position k: L_kk^2 / |K_gg| for
 * the global DOF g eliminated at k (1.0 where |K_gg| <= 1e-30); ratios_out holds
 * n_free doubles.  The pivot_ratios loop of src/solver/boundary_conditions.jl:1664-1675. */
int jfem_factor_pivot_ratios(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K,
                             double *ratios_out);
/* the elimination order of a symbolic factorization: dofs_out[k] (n_free entries)
 * is the global DOF eliminated at position k (the reference's F_chol.p,
 * src/solver/boundary_conditions.jl:1666, with the free-set indexing resolved) */
int jfem_symbolic_order(const jfem_symbolic *S, int64_t *dofs_out);
int jfem_factor_free(jfem_factor *F);

/* Solve K_ff x = b for nrhs <= 16 right-hand sides given in FULL DOF numbering
 * (b: ndof x nrhs column-major, fixed rows ignored; x: ndof x nrhs, fixed rows 0).
 * Replaces F \ F_ff (src/solver/boundary_conditions.jl:1442, 1714) and the
 * residual (:1859-1868).  The iterate is a double-double x = x_hi + x_lo refined
 * in an outer loop on the exact residual
 *   r = b - K_ff x  (accumulated exactly: fma-split products, two-sum),
 * each step advancing one conjugate-gradient iteration on K_ff d = r,
 * preconditioned with the factor (triangular solves in its precision), until its
 * residual is below 1e-4 x r for every right-hand side still running (or
 * 0.9 tol x ||b||); then x += d and the iteration continues from the new exact
 * residual, keeping its search directions; until
 * ||K_ff x - b|| / max(||b||, 1e-30) <= tol or the residual no longer halves.
 * max_refine bounds the triangular solves to max_refine + 1 per call;
 * max_refine = 0 is the plain solve x = F^{-1} b.  relres[k] / iters[k]: residual
 * of the double-double iterate and triangular solves used, per RHS.  x receives
 * x_hi; x_lo (optional, may be NULL) the low parts; relres_out (optional) the
 * residual of x_hi alone, i.e. of the double vector handed back, whose floor is
 * about 2^-53 || |K_ff| |x| || / ||b||.  Returns JFEM_ERR_NOT_CONVERGED (x still
 * written) when a right-hand side misses tol.
 * b / x on the host: *_host; the _dev variant takes device pointers. */
int jfem_solve(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs,
               const double *b_host, double *x_host, double tol, int max_refine,
               int32_t *iters, double *relres, double *x_lo_host, double *relres_out);
int jfem_solve_dev(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs,
                   const double *b_dev, double *x_dev, double tol, int max_refine,
                   int32_t *iters, double *relres, double *x_lo_dev, double *relres_out);

/* jfem_solve_dev that allocates nothing: the same solve (same kernels, same
 * results bit for bit) with its scratch in a static device array of the library
 * (128 MiB per GPU, held until the process ends) instead of the context's
 * scratch, which jfem_solve_dev grows with hipMalloc.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  A solve that needs more scratch than
 * the array holds (jfem_solve_scratch_bytes) returns JFEM_ERR_BAD_ARG.  Every
 * call on one GPU uses the same array: calls that overlap on one device, from any
 * context, race. */
int jfem_solve_dev_static(jfem_ctx *ctx, const jfem_factor *F, const jfem_matrix *K, int nrhs,
                          const double *b_dev, double *x_dev, double tol, int max_refine,
                          int32_t *iters, double *relres, double *x_lo_dev, double *relres_out);

/* ---- load vectors ------------------------------------------------------- */
/* A load set: the FORCE / MOMENT cards on grid points and the uniform PLOAD4 cards
 * on shell elements of one load SID, in card order, basic coordinates (CID 0).
 * Replaces the FORCE, MOMENT and shell-PLOAD4 branches of _resolve_loads!
 * (src/solver/loads.jl:517-587) and the point-force application (:1338-1340).
 *   PLOAD4 card i: shell element pload4_elem[i] (CQUAD4 e for e < n_quad of the
 *     mesh, else CTRIA3 e - n_quad), pressure pload4_p[i], direction
 *     pload4_dir3[3i..3i+2] (NULL, or a vector of norm <= 1e-30: the element
 *     normal); elements of area <= 1e-30 take no load, as in the reference;
 *   FORCE / MOMENT card i: grid point *_node[i], magnitude *_mag[i], direction
 *     *_dir3[3i..3i+2], taken as given (the reference's reader rounds both to
 *     single precision: src/parsing/extract_loads.jl:15-29).
 * GRAV on shells is added by jfem_loadset_set_gravity; ACCEL, PLOAD1/2, GRAV on
 * solids, coordinate systems and LOAD combinations stay with the caller (a
 * combination is a sum of load vectors with their factors). */
int jfem_loadset_create(jfem_ctx *ctx, const jfem_mesh *mesh, int64_t n_pload4,
                        const int32_t *pload4_elem, const double *pload4_p,
                        const double *pload4_dir3, int64_t n_force, const int32_t *force_node,
                        const double *force_mag, const double *force_dir3, int64_t n_moment,
                        const int32_t *moment_node, const double *moment_mag,
                        const double *moment_dir3, jfem_loadset **out);
/* GRAV cards of the set on its shells, replacing any set before: card i has
 * acceleration grav_a[i] along grav_n3[3i..3i+2] (basic frame, CID 0); shell_areal
 * (n_quad + n_tria, CQUAD4s then CTRIA3s) is each shell's mass per area rho T + NSM,
 * 0 for a shell without mass.  Each shell adds (areal * area / nn) .* (A scale) N
 * to its nodes' translations, area half the norm of the diagonal (CQUAD4) or edge
 * (CTRIA3) cross product, cards in order and shells in element order, after the
 * PLOAD4 and before the FORCE / MOMENT contributions: the GRAV shell branch of
 * _resolve_loads! (src/solver/loads.jl:680-684, 730-772).  n_grav = 0 removes them. */
int jfem_loadset_set_gravity(jfem_ctx *ctx, jfem_loadset *ls, int64_t n_grav, const double *grav_a,
                             const double *grav_n3, const double *shell_areal);
/* The set's load vector with scale factor `scale` (applied where the reference
 * applies it: area*P*scale per PLOAD4, (Dir*Mag)*scale per FORCE/MOMENT, A*scale
 * per GRAV), all
 * 6*n_nodes DOFs, to the host (F_host) or to device memory (F_dev, on this
 * context's stream).  Sums run in the reference's order: bitwise reproducible.
 * A load set keeps per-call buffers: calls on one load set must not overlap. */
int jfem_load_vector(jfem_ctx *ctx, const jfem_loadset *ls, double scale, double *F_host);
int jfem_load_vector_dev(jfem_ctx *ctx, const jfem_loadset *ls, double scale, double *F_dev);
int jfem_loadset_free(jfem_loadset *ls);

/* device buffers for the _dev entry points (owned by the caller) */
int jfem_dev_alloc(jfem_ctx *ctx, int64_t bytes, void **dev_ptr);
int jfem_dev_free(jfem_ctx *ctx, void *dev_ptr);
int jfem_memcpy_h2d(jfem_ctx *ctx, void *dev_dst, const void *host_src, int64_t bytes);
int jfem_memcpy_d2h(jfem_ctx *ctx, void *host_dst, const void *dev_src, int64_t bytes);

/* Diagnostics: device self-test of the matrix-core trailing-update tile kernel
 * against its scalar form (precision 64 | 32); max_rel_diff receives the largest
 * difference relative to max |entry|.  (No reference counterpart.) */
int jfem_selftest_update(jfem_ctx *ctx, int precision, double *max_rel_diff);

/* memory held by a handle (device bytes); admission control of the serving loop
 * (no reference counterpart) */
int64_t jfem_factor_device_bytes(const jfem_factor *F);
int64_t jfem_matrix_device_bytes(const jfem_matrix *K);
int64_t jfem_mesh_device_bytes(const jfem_mesh *mesh);
int64_t jfem_pattern_device_bytes(const jfem_pattern *pattern);
int64_t jfem_symbolic_device_bytes(const jfem_symbolic *S);
/* solve scratch and BC buffers a context currently holds */
int64_t jfem_ctx_device_bytes(const jfem_ctx *ctx);
/* solve scratch a jfem_solve call with nrhs right-hand sides on a factor of the
 * given precision over S needs in its context (ndof = all DOFs of the pattern) */
int64_t jfem_solve_scratch_bytes(const jfem_symbolic *S, int64_t ndof, int nrhs, int precision);

#ifdef __cplusplus
}
#endif
#endif
