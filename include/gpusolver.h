/*
 * gpusolver: C ABI of the MI355X (gfx950) AMG-PCG solve phase.
 *
 * Drop-in boundary for the solve path of the reference library
 * (include/GPUSolver in the reference tree). The reference's own binding
 * surface is its C++ API: CGSolver (include/GPUSolver/CGSolver.h:19-86),
 * the Preconditioner / Smoother plugins (include/GPUSolver/Preconditioner.h:
 * 13-20, include/GPUSolver/Smoother.h:12-22), the Backend plugin
 * (include/GPUSolver/Backend.h:15-159) and the setup artifacts it consumes
 * (HostCSRMatrix, HostELLMatrix, Permutation, GraphColoring, AMGHierarchy).
 * Every entry point below replaces one of those, as cited on each
 * declaration. Plain pointers and sizes only; no C++ or torch types.
 *
 * Conventions
 *   - Every function returns GS_OK (0) or GS_ERR (-1). On GS_ERR,
 *     gs_last_error() returns the message of the std::runtime_error the
 *     reference would have thrown (thread-local, valid until the next call).
 *   - Matrices are CSR with int32 row pointers / column indices and FP64
 *     values; FP32 device copies are rounded from them once.
 *   - Permutations are given as oldToNew arrays (src/Permutation.cc:8-35).
 *   - Handles are opaque owning pointers; destroy each exactly once.
 */
#ifndef GPUSOLVER_H
#define GPUSOLVER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_OK 0
#define GS_ERR (-1)

/* Precision modes of a solver: CG vectors / V-cycle. */
#define GS_PREC_DOUBLE 0 /* FP64 / FP64 (plain)                        */
#define GS_PREC_MIXED 1  /* FP64 / FP32 (mixed)                        */
#define GS_PREC_SINGLE 2 /* FP32 / FP32 (the reference's precision)    */

/* Smoother kinds, numbered as MultigridSmoother::SmootherType
 * (include/GPUSolver/Smoothers/MultigridSmoother.h:18-23). */
#define GS_SMOOTHER_DAMPED_JACOBI 0
#define GS_SMOOTHER_SYMMETRIC_GAUSS_SEIDEL 1
#define GS_SMOOTHER_BLOCK_JACOBI 2
#define GS_SMOOTHER_BLOCK_GAUSS_SEIDEL 3

const char *gs_last_error(void);
int gs_version(void);

/* ------------------------------------------------------------------ */
/* Host CSR matrices. Replaces HostCSRMatrix's raw-pointer constructor */
/* (include/GPUSolver/HostSparseMatrix.h:34-35, src/HostSparseMatrix.cc */
/* :9-55): the arrays are copied and validated.                        */
/* ------------------------------------------------------------------ */
typedef struct gs_csr_s *gs_csr_h;
int gs_csr_create(int64_t rows, int64_t cols, int64_t nnz, const int32_t *row_ptr,
                  const int32_t *col_idx, const double *values, gs_csr_h *out);
int gs_csr_info(gs_csr_h A, int64_t *rows, int64_t *cols, int64_t *nnz);
int gs_csr_export(gs_csr_h A, int32_t *row_ptr, int32_t *col_idx, double *values);
int gs_csr_destroy(gs_csr_h A);

/* HostCSRMatrix::applyPermutation, A' = P A P^T keeping each row's entry
 * order (src/HostSparseMatrix.cc:118-249). */
int gs_csr_permute_keep_order(gs_csr_h A, const int32_t *old_to_new, gs_csr_h *out);
/* permuteMatrix: rows / columns relabelled, rows sorted, duplicates summed
 * (src/SparseMatrixUtils.cc:8-31). */
int gs_csr_permute_sorted(gs_csr_h A, const int32_t *row_old_to_new,
                          const int32_t *col_old_to_new, gs_csr_h *out);

/* HostELLMatrix (src/HostELLMatrix.cc:10-206): slot-major ELL with CSR
 * overflow. Call once with ell_col == NULL to get *overflow_nnz. */
int gs_ell_pack(gs_csr_h A, int32_t width, int32_t *ell_col, double *ell_val,
                int32_t *ovf_row_ptr, int32_t *ovf_col, double *ovf_val, int64_t *overflow_nnz);

/* Device layout of a scalar operator (SELL-64, no counterpart in the
 * reference): packs A as the solver uploads it, validates it (every stored
 * column inside [0, cols), no padding slot on the diagonal) and reports the
 * slot count and the smallest / largest stored column. */
int gs_csr_sell_layout(gs_csr_h A, int64_t *slots, int32_t *min_col, int32_t *max_col);

/* GraphColoring::compute (src/GraphColoring.cc:9-72). offsets holds
 * ncolors + 1 entries (size it n + 1). */
int gs_graph_coloring(int64_t n, const int32_t *row_ptr, const int32_t *col_idx,
                      int32_t *colors, int32_t *color_vertices, int32_t *color_offsets,
                      int32_t *ncolors);
/* RCMReordering::compute (src/ReorderingStrategies/RCMReordering.cc:11-110). */
int gs_rcm(int64_t n, const int32_t *row_ptr, const int32_t *col_idx, int32_t *old_to_new,
           int32_t *new_to_old);
/* Permutation's validation (src/Permutation.cc:8-35). */
int gs_permutation_check(int64_t n, const int32_t *old_to_new, const int32_t *new_to_old);
/* buildBlockGraph (src/BlockUtils.cc:126-216). Call with g_ptr == NULL for *g_nnz. */
int gs_block_graph(gs_csr_h A, int32_t block_size, int32_t *g_ptr, int32_t *g_col, int64_t *g_nnz);
/* expandBlockPermutation (src/BlockUtils.cc:7-55). */
int gs_expand_block_permutation(int64_t nodes, int32_t block_size, const int32_t *node_old_to_new,
                                int32_t *old_to_new, int32_t *new_to_old);
/* extractInverseDiagonalBlocks3x3 (src/BlockUtils.cc:218-277), 9 per block. */
int gs_inverse_diag_blocks3(gs_csr_h A, double *inverse_blocks);
/* In-place FP64 inverse of a dense SPD matrix (row-major n x n) through its
 * Cholesky factor: the coarsest-level direct solve of the mixed mode
 * (GS_PREC_MIXED, "coarse_direct"), which replaces the reference's coarsest
 * smoother sweeps (src/Smoothers/MultigridSmoother.cc:348-354). Fails when
 * the matrix is not SPD. */
int gs_dense_spd_inverse(int64_t n, double *a);
/* The same inverse built on the GPU (blocked Gauss-Jordan elimination in
 * FP64, the engine's coarse-level setup); a is read and overwritten on the
 * host. */
int gs_dense_spd_inverse_device(int64_t n, double *a);

/* ------------------------------------------------------------------ */
/* Problem generators (replace the reference's mesh inputs).           */
/* ------------------------------------------------------------------ */
int gs_problem_poisson2d(int32_t nx, int32_t ny, gs_csr_h *out);
int gs_problem_poisson3d(int32_t nx, int32_t ny, int32_t nz, gs_csr_h *out);
int gs_problem_varcoef3d(int32_t n, uint64_t seed, double sigma, double scale, gs_csr_h *out);
/* Kuhn-tet elasticity on nv^3 vertices, x = 0 clamped. rhs (rows) and
 * nullspace (rows x 6, row-major) may be NULL. */
int gs_problem_elasticity(int32_t nv, double young, double poisson, double hetero_sigma,
                          double hetero_scale, uint64_t seed, gs_csr_h *out, double *rhs,
                          double *nullspace);
/* The cantilever driver's system on a caller's tet mesh
 * (examples/elasticity/cantileverBeam.cc:275-348 with
 * examples/elasticity/LinearElasticity.ipp:20-79,115-147,239-348,379-413):
 * V nverts x 3 and T ntets x 4 (0-based), row-major; young_per_tet may be
 * NULL (young everywhere; the driver uses 1e5 with poisson 0.35); vertices
 * with x <= x_min + clamp_rel * (x_max - x_min) are clamped (the driver:
 * 1e-4); rhs = density * lumped mass * (0, 0, gravity) (the driver: 1,
 * -9.81). rhs and nullspace (rigid-body modes, rows x 6) may be NULL, else
 * must hold 3 * nverts and 18 * nverts values; the first rows of the
 * returned matrix are written. Errors: "No vertices found on clamped
 * side." (the driver's), empty mesh, vertex index out of range, degenerate
 * tetrahedron. */
int gs_problem_elasticity_mesh(int64_t nverts, const double *V, int64_t ntets, const int32_t *T,
                               double young, double poisson, const double *young_per_tet,
                               double clamp_rel, double density, double gravity, gs_csr_h *out,
                               double *rhs, double *nullspace);
int gs_uniform_pm1(uint64_t seed, int64_t n, double *out);
int gs_gaussian_field(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double sigma, double *out);

/* ------------------------------------------------------------------ */
/* AMG hierarchy. Replaces AMGHierarchy{A, P, R}                       */
/* (include/GPUSolver/AMGHierarchy.h:13-21) and AMGHierarchyBuilder +  */
/* SmoothedAggregationCoarsener (examples/amgSolver/src/               */
/* AMGHierarchyBuilder.cc:58-148, SmoothedAggregationCoarsener.cc:     */
/* 64-108).                                                            */
/* ------------------------------------------------------------------ */
typedef struct gs_hier_s *gs_hier_h;
int gs_hierarchy_build_sa(gs_csr_h A, int32_t block_size, double eps_strong, int32_t max_levels,
                          int64_t min_dofs, const double *nullspace, int32_t nullspace_cols,
                          gs_hier_h *out);
/* The same hierarchy, bit for bit, with the sparse products of each
 * coarsening step on the GPU: the prolongation smoothing P = (I - w
 * D_f^-1 A_f) P_tent, R = P^T and the Galerkin product A_c = R A P
 * (amgcl's transfer_operators / coarse_operator, examples/amgSolver/src/
 * SmoothedAggregationCoarsener.cc:93-97). Aggregation stays on the host:
 * it is sequential by definition. */
int gs_hierarchy_build_sa_device(gs_csr_h A, int32_t block_size, double eps_strong,
                                 int32_t max_levels, int64_t min_dofs, const double *nullspace,
                                 int32_t nullspace_cols, gs_hier_h *out);
/* C = A B with rows sorted by column, each entry summed in the order
 * A's row entries and then B's visit it (the products of the
 * smoothed-aggregation setup, above); on_device 0: host, 1: GPU, the
 * same result bit for bit. */
int gs_csr_spgemm(gs_csr_h A, gs_csr_h B, int32_t on_device, gs_csr_h *out);
/* A^T with rows sorted by column (R = P^T of the setup); on_device as for
 * gs_csr_spgemm. */
int gs_csr_transpose(gs_csr_h A, int32_t on_device, gs_csr_h *out);
int gs_hierarchy_create(gs_hier_h *out);
int gs_hierarchy_push_level(gs_hier_h h, gs_csr_h A);
int gs_hierarchy_push_transfer(gs_hier_h h, gs_csr_h P, gs_csr_h R);
int gs_hierarchy_levels(gs_hier_h h, int32_t *nlev);
/* which: 0 = A[k], 1 = P[k], 2 = R[k]; returns a new host CSR copy. */
int gs_hierarchy_get(gs_hier_h h, int32_t which, int32_t k, gs_csr_h *out);
int gs_hierarchy_aggregates(gs_hier_h h, int32_t k, int32_t *aggregate_of_dof);
/* Rows of A[k] (hierarchy.A[k].rows(), include/GPUSolver/AMGHierarchy.h:
 * 13-21) without copying the matrix. */
int gs_hierarchy_level_rows(gs_hier_h h, int32_t k, int64_t *rows);
/* HierarchicalReordering::compute (include/GPUSolver/ReorderingStrategies/
 * HierarchicalReordering.h:10-16, src/ReorderingStrategies/
 * HierarchicalReordering.cc:217-278): one permutation per level, the
 * coarsest by RCM, every finer level grouped by the rank of each DOF's
 * strongest coarse parent in R (largest FP32 |value|, first on a tie; DOFs
 * without one last) and by its own level's RCM rank within a group.
 * old_to_new / new_to_old receive the levels back to back, finest first
 * (sum of the level sizes each), ready to pass level by level to
 * gs_solver_setup. orphans (levels - 1 entries, may be NULL) receives per
 * transfer the number of fine DOFs without a parent, which the reference
 * prints to standard output instead. Errors are the reference's messages,
 * plus two checks it lacks (R's rows against the coarse level, a square
 * coarsest matrix). */
int gs_hierarchical_reordering(gs_hier_h h, int32_t *old_to_new, int32_t *new_to_old, int64_t *orphans);
int gs_hierarchy_destroy(gs_hier_h h);
/* Binary files, no counterpart in the reference (whose drivers rebuild the
 * hierarchy with amgcl on every run): one CSR matrix, or everything an
 * AMGHierarchy holds (include/GPUSolver/AMGHierarchy.h:13-21: A, P, R) plus
 * the aggregate maps. Little-endian: an 8-byte tag ("GSCSR001" /
 * "GSHIER01"); for a hierarchy int64 levels L and transfers T, then L A
 * records, T P records, T R records, int64 count of aggregate maps and each
 * as int64 n + int32[n]; a matrix record is int64 rows, cols, nnz, int32
 * ptr[rows + 1], int32 col[nnz], float64 val[nnz]. Loading checks sizes
 * against the file before allocating and validates every CSR record
 * ("truncated file", "corrupt ... header", "is not a ... file", "trailing
 * bytes", or the CSR check's own message). */
int gs_csr_save(gs_csr_h A, const char *path);
int gs_csr_load(const char *path, gs_csr_h *out);
int gs_hierarchy_save(gs_hier_h h, const char *path);
int gs_hierarchy_load(const char *path, gs_hier_h *out);

/* ------------------------------------------------------------------ */
/* Solver: CGSolver (src/CGSolver.cc:18-308) with one of the reference's */
/* preconditioners (src/Preconditioners) on the GPU.                  */
/* ------------------------------------------------------------------ */
typedef struct gs_solver_s *gs_solver_h;
int gs_solver_create(int32_t precision, gs_solver_h *out);
/* Integer options (set before gs_solver_setup):
 *   "preconditioner"  0 identity, 1 multigrid (default)
 *   "smoother"        fine-level GS_SMOOTHER_* (default SGS)
 *   "coarse_smoother" levels > 0 (default: same as "smoother")
 *   "pre", "post"     sweeps (default 1, 1)
 *   "coarse_steps"    coarsest-level sweeps (default 2 = pre + post,
 *                     src/Preconditioners/MultigridGaussSeidelPreconditioner.cc:21)
 *   "vcycles"         V-cycles per application (default 1)
 *   "coarse_direct"   1: FP64 direct solve on the coarsest level (explicit
 *                     inverse built on the device, applied as a product
 *                     over its lower 64 x 64 tiles)
 *   "graph_iters"     PCG iterations per captured hipGraph (0: no graph)
 *   "graph_tail"      1 (default): when the last chunk's convergence rate says
 *                     fewer than graph_iters iterations remain, replay a
 *                     one-iteration graph that many times instead of a
 *                     whole chunk (same results; fewer no-op launches).
 *                     Off for a true-residual period above graph_iters that
 *                     is not a multiple of it.
 *   "cg_block3"       1: store the CG operator as 3x3 blocks
 *   "reference_vcycle" 1: reproduce the reference's coarse-level correction
 *                     aliasing (src/Smoothers/MultigridSmoother.cc:401-404:
 *                     below the finest level the prolongated correction
 *                     overwrites the level's solution and is then added to
 *                     itself); 0 (default): x += P e
 *   "max_direct"      largest level solved by the dense FP64 inverse (16384)
 *   "direct_truncate" with "coarse_direct": solve directly at the first
 *                     level of at most "max_direct" unknowns, leaving the
 *                     levels below unused (0)
 *   "group_max_transfer" largest lane group (16, 32, 64) for the rows of the
 *                     restriction / prolongation operators (32)
 *   "transfer_sigma"  restriction / prolongation rows kept one per thread
 *                     are sorted by length within windows of this many rows
 *                     (SELL-C-sigma, less slice padding; same sums); 0: the
 *                     natural order (1024)
 *   "transfer_level_order" 1: restriction / prolongation rows run in the
 *                     level's own (pre-colouring) row order, neighbouring
 *                     rows together; outputs go through a row map, sums are
 *                     unchanged; 0: the colour order (1)
 *   "block_transfers" 1: restriction / prolongation between two levels
 *                     whose orderings keep node triples (block smoothers)
 *                     are stored as 3x3 blocks, same sums (1)
 *   "cg_level_order"  1: the 3x3-block CG operator's rows run in level-0
 *                     order through a row map (each row's sum unchanged;
 *                     the reductions group rows differently) (1)
 *   "cg_pair_layout"  1: the FP64 3x3-block CG operator stores its values
 *                     as planes of value pairs (16-byte loads) and its SpMV
 *                     runs as a persistent launch over the slices (row sums
 *                     unchanged; p.Ap grouped by workgroup) (1)
 *   "tail_rows"       the scalar Gauss-Seidel levels below the finest with
 *                     at most this many unknowns, down to the coarsest, run
 *                     their part of the V-cycle (sweeps, residual,
 *                     transfers, coarse sweeps) as one single-workgroup
 *                     kernel instead of one launch per colour; same results
 *                     bit for bit; 0: off (4096)
 *   "zero_start_lower" k: Gauss-Seidel levels below index k keep a second
 *                     operator with only the entries of earlier colours; the
 *                     forward colours of a sweep from x = 0 read it (the
 *                     entries it leaves out multiply exact zeros; same
 *                     results); 0: none (1: level 0, where the sweeps are
 *                     bandwidth-bound)
 *   "device_pack"     1: 3x3-block operators (block-smoother levels, a
 *                     3x3-block CG operator) are relabelled and packed on
 *                     the GPU from one upload of the level matrix, into the
 *                     layouts the host packers produce; 0: on the host (1)
 *   "persist_grid"    workgroups of the persistent FP64 3x3-block CG SpMV
 *                     (a positive multiple of 8; 4096): one share of the
 *                     slices per XCD, each wavefront walking slices at a
 *                     stride of the waves in its share
 *   "cg_vcycle_operator" 1 (GS_PREC_MIXED): the CG SpMV reads level 0's
 *                     FP32 V-cycle operator, products and sums in FP64; the
 *                     FP64 operator recomputes r = b - A x every
 *                     "cg_replace_every" iterations (8) and confirms every
 *                     stop on convergence; needs level 0 smoothed and
 *                     stored one row / block row per thread, else ignored (0)
 *   "cg_true_residual_every" (GS_PREC_MIXED, FP64 CG operator): the true
 *                     residual b - A x is recomputed in FP64 every that many
 *                     iterations and before every stop on convergence, and
 *                     replaces the recursive residual; 0: only the final
 *                     true residual after the solve (32)
 * Real options: "omega" (default 1.0); "group_min_scalar" (24) and
 * "group_min_block" (20): operators whose average row length reaches the
 * value (scalar entries per row, 3x3 blocks per block row) are processed
 * by lane groups, one row per group of 16/32/64 lanes, with the same
 * arithmetic in the same order; negative values keep one thread per row.
 * "full_group_min_scalar" (40) and "full_group_min_block" (20): the same
 * thresholds for the operators read only by whole-level kernels (the
 * transfers P and R, and the operators of Jacobi-smoothed levels).
 * Takes effect at the next gs_solver_setup. */
int gs_solver_set_int(gs_solver_h s, const char *key, int64_t value);
int gs_solver_set_real(gs_solver_h s, const char *key, double value);
/* Per-level damping and pre = post sweep counts (learned weights). An omega
 * entry <= 0 keeps the global setting; a sweep entry < 0 keeps it, 0 turns
 * that level's smoothing off (a no-op for block Gauss-Seidel and Jacobi
 * levels, as the reference's 0-sweep smoothers; rejected for symmetric
 * Gauss-Seidel, which requires a sweep). */
int gs_solver_set_level_weights(gs_solver_h s, int32_t nlev, const double *omega,
                                const int32_t *sweeps);
/* Upload the hierarchy. level_old_to_new[k] permutes level k
 * (MultigridSmoother ctor, src/Smoothers/MultigridSmoother.cc:22-335); NULL
 * (or a NULL entry) means identity. The CG operator is A[0] permuted by
 * level 0's permutation (CGSolver ctor, src/CGSolver.cc:44-70). */
int gs_solver_setup(gs_solver_h s, gs_hier_h h, const int32_t *const *level_old_to_new);
/* Learned-weight guard (no counterpart in the reference): when set, a
 * solve of s that breaks down (p.Ap <= 0 or r.z <= 0), produces a
 * non-finite residual, or lags the plain rate is restarted from the same
 * initial guess with `plain`; "fell_back" in gs_solver_result reports it.
 * The plain rate is the mean relres^2 reduction per iteration of plain's
 * last converged solve: at every host check of the stop word (each
 * "graph_iters" iterations) the learned solve must have reduced its
 * residual as much as plain does in lag times fewer iterations, and it may
 * take at most lag times plain's last iteration count. A solve with a
 * fallback fails when plain is not set up, was set up for another size, or
 * has not yet converged once. */
int gs_solver_set_fallback(gs_solver_h s, gs_solver_h plain, double lag);
/* CGSolver::solve (src/CGSolver.cc:87-306): b and x (initial guess in,
 * solution out) in the caller's ordering; b_len and x_len are their
 * lengths, checked against the matrix as src/CGSolver.cc:90-98 checks the
 * Eigen vectors' sizes (same messages). iters follows the reference's
 * convention (getNbOfIterations, src/CGSolver.cc:257-268,308). */
int gs_solver_solve(gs_solver_h s, int64_t b_len, const double *b, int64_t x_len, double *x,
                    double tol, int64_t maxit, int64_t *iters, double *relres2,
                    double *true_relres, double *solve_ms);
/* Rows of the matrix the solver was set up for (CGSolver's matrix_.rows(),
 * include/GPUSolver/CGSolver.h:19-86). */
int gs_solver_size(gs_solver_h s, int64_t *n);
/* Result fields of the last solve: "converged", "loop_iterations",
 * "launched_iterations" (iterations launched, no-op ones past convergence
 * included),
 * "breakdown", "wall_ms", "zero_rhs", "fell_back", "lagging" (ended by the
 * fallback's rate check), "fallback_reason" (0 none, 1 lagging the plain
 * rate, 2 CG breakdown, 3 non-finite residual, 4 iteration budget spent),
 * "first_iterations" (PCG iterations of the guarded attempt), and
 * "log_rate" (this solver's rate as a fallback, 0 before its first
 * converged solve). */
int gs_solver_result(gs_solver_h s, const char *key, double *value);
int gs_solver_history(gs_solver_h s, double *relres2, int64_t cap, int64_t *len);
/* One preconditioner application z = M r (Preconditioner::apply,
 * include/GPUSolver/Preconditioner.h:13-20), vectors of length n (the
 * finest level's size, else an error) in level-0 ordering. */
int gs_solver_vcycle(gs_solver_h s, int64_t n, const double *r, double *z);
/* y = A x with the CG operator (Backend::encodeSpmv, include/GPUSolver/
 * Backend.h:68-69), vectors of length n in level-0 ordering; a length other
 * than the matrix size fails with encodeSpmv's "size mismatch"
 * (src/CPU/CPUBackend.cc:220-222). */
int gs_solver_spmv(gs_solver_h s, int64_t n, const double *x, double *y);
/* Average ms per launch of a hot kernel timed with HIP events on the
 * solver's stream: 0 CG SpMV, 1 V-cycle, 2 PCG iteration, 3 / 4 level-0 /
 * level-1 SGS or block-GS sweep, 5 level-0 restriction, 6 level-0
 * prolongation, 7 level-0 residual (rhs - A x in the V-cycle type), 8 the
 * single-workgroup V-cycle tail (its input as the last V-cycle left it).
 * *bytes: algorithmic bytes per launch. */
int gs_solver_time_kernel(gs_solver_h s, int32_t which, int32_t reps, double *ms, double *bytes);
int gs_solver_describe(gs_solver_h s, char *buf, int64_t cap);
int gs_solver_destroy(gs_solver_h s);

/* ------------------------------------------------------------------ */
/* Backend plugin (include/GPUSolver/Backend.h:15-159) on HIP.          */
/* ------------------------------------------------------------------ */
#define GS_DTYPE_F32 0 /* the reference's value type                 */
#define GS_DTYPE_F64 1 /* FP64 extension of the interface            */

typedef struct gs_backend_s *gs_backend_h;
typedef struct gs_encoder_s *gs_encoder_h;
typedef struct gs_dvec_s *gs_dvec_h;     /* DeviceVector      (DeviceVector.h:6-20)      */
typedef struct gs_divec_s *gs_divec_h;   /* DeviceIndexVector (DeviceIndexVector.h:8-26) */
typedef struct gs_dscalar_s *gs_dscalar_h; /* DeviceScalar    (DeviceScalar.h:18-20)     */
typedef struct gs_dcsr_s *gs_dcsr_h;     /* DeviceCSRMatrix   (DeviceSparseMatrix.h:5-18) */
typedef struct gs_dell_s *gs_dell_h;     /* DeviceELLMatrix   (DeviceELLMatrix.h:8-25)   */

int gs_backend_create(int32_t dtype, gs_backend_h *out);
int gs_backend_destroy(gs_backend_h b);
const char *gs_backend_last_error(void);
/* createEncoder / submit / submitAndWait (Backend.h:19-23). */
int gs_backend_create_encoder(gs_backend_h b, gs_encoder_h *out);
int gs_backend_submit(gs_backend_h b, gs_encoder_h e);
int gs_backend_submit_and_wait(gs_backend_h b, gs_encoder_h e);
int gs_encoder_destroy(gs_encoder_h e);
/* The encoder's HIP stream (a hipStream_t) for work the caller enqueues in
 * order with it (CudaEncoder::stream, include/GPUSolver/cuda/
 * CudaEncoder.h:17). Owned by the encoder. */
int gs_encoder_stream(gs_encoder_h e, void **stream);
/* createVector(size[, values]) (Backend.h:26-28); values in the backend's dtype. */
int gs_backend_create_vector(gs_backend_h b, int64_t n, const void *values, gs_dvec_h *out);
int gs_dvec_size(gs_dvec_h v, int64_t *n);
int gs_dvec_update(gs_dvec_h v, const void *values, int64_t n);
int gs_dvec_download(gs_dvec_h v, void *out);
int gs_dvec_destroy(gs_dvec_h v);
/* Device address of the vector's values (CudaDeviceVector::data,
 * include/GPUSolver/cuda/CudaDeviceVector.h:38-40); valid until destroy. */
int gs_dvec_data(gs_dvec_h v, void **ptr);
/* createIndexVector(size[, values]) (Backend.h:30-33). */
int gs_backend_create_index_vector(gs_backend_h b, int64_t n, const int32_t *values,
                                   gs_divec_h *out);
int gs_divec_download(gs_divec_h v, int32_t *out);
int gs_divec_destroy(gs_divec_h v);
/* size() / updateValues(values, n) (DeviceIndexVector.h:14-20); n must be
 * the vector's size (src/CPU/CPUDeviceIndexVector.cc:24-31). */
int gs_divec_size(gs_divec_h v, int64_t *n);
int gs_divec_update(gs_divec_h v, const int32_t *values, int64_t n);
/* createScalar([value]) (Backend.h:35-37). */
int gs_backend_create_scalar(gs_backend_h b, double value, gs_dscalar_h *out);
int gs_dscalar_download(gs_dscalar_h s, double *value);
int gs_dscalar_destroy(gs_dscalar_h s);
/* Device address of the scalar (CudaDeviceScalar::data,
 * include/GPUSolver/cuda/CudaDeviceScalar.h:15-16). */
int gs_dscalar_data(gs_dscalar_h s, void **ptr);
/* createCSRMatrix(const HostCSRMatrix&) (Backend.h:39); updateValues. */
int gs_backend_create_csr(gs_backend_h b, gs_csr_h host, gs_dcsr_h *out);
int gs_dcsr_update_values(gs_dcsr_h m, const double *values, int64_t nnz);
int gs_dcsr_destroy(gs_dcsr_h m);
/* rows() / cols() / nnz() (DeviceSparseMatrix.h:8-10); null outputs skipped. */
int gs_dcsr_info(gs_dcsr_h m, int64_t *rows, int64_t *cols, int64_t *nnz);
/* createELLMatrix(HostELLMatrix(csr, width)) (Backend.h:41). */
int gs_backend_create_ell(gs_backend_h b, gs_csr_h host, int32_t width, gs_dell_h *out);
int gs_dell_destroy(gs_dell_h m);
/* createELLMatrix(const HostELLMatrix&) (Backend.h:41) from the arrays the
 * host matrix holds (HostELLMatrix.h:19-31, src/HostELLMatrix.cc:10-206):
 * rows x width slot-major ELL (index slot * rows + row; padding column 0,
 * value 0), then the overflow CSR (ovf_ptr: rows + 1 entries; ovf_col,
 * ovf_val: ovf_nnz entries, may be null when ovf_nnz is 0). Values in the
 * backend's dtype. Column indices and the overflow row pointer are
 * checked. */
int gs_backend_create_ell_packed(gs_backend_h b, int64_t rows, int64_t cols, int64_t width,
                                 const int32_t *ell_col, const void *ell_val, int64_t ovf_nnz,
                                 const int32_t *ovf_ptr, const int32_t *ovf_col,
                                 const void *ovf_val, gs_dell_h *out);
/* rows() / cols() / ellWidth() / overflowNnz() (DeviceELLMatrix.h:17-21). */
int gs_dell_info(gs_dell_h m, int64_t *rows, int64_t *cols, int64_t *width, int64_t *ovf_nnz);
/* encode* (Backend.h:47-114). */
int gs_encode_scale(gs_encoder_h e, gs_dvec_h x, double alpha);
int gs_encode_scale_dev(gs_encoder_h e, gs_dvec_h x, gs_dscalar_h alpha);
int gs_encode_copy(gs_encoder_h e, gs_dvec_h x, gs_dvec_h y);
int gs_encode_scale_vector(gs_encoder_h e, gs_dvec_h scale, gs_dvec_h out);
int gs_encode_set_zero(gs_encoder_h e, gs_dvec_h x);
int gs_encode_axpy(gs_encoder_h e, double alpha, gs_dvec_h x, gs_dvec_h y);
int gs_encode_axpy_dev(gs_encoder_h e, gs_dscalar_h alpha, gs_dvec_h x, gs_dvec_h y);
int gs_encode_spmv(gs_encoder_h e, gs_dcsr_h A, gs_dvec_h x, gs_dvec_h y);
/* MetalBackend::encodeSpmvELLMetal (include/GPUSolver/metal/MetalBackend.h:229-236). */
int gs_encode_spmv_ell(gs_encoder_h e, gs_dell_h A, gs_dvec_h x, gs_dvec_h y);
int gs_encode_gauss_seidel_color(gs_encoder_h e, gs_dcsr_h A, gs_divec_h color_vertices,
                                 int64_t color_start, int64_t color_count, gs_dvec_h rhs,
                                 gs_dvec_h x, double omega);
int gs_encode_dot(gs_encoder_h e, gs_dvec_h x, gs_dvec_h y, gs_dscalar_h out);
/* No counterpart in the reference: the variant of gs_encode_dot that
 * allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). out = x . y for caller-owned device arrays of n values of
 * dtype (GS_DTYPE_F32 / F64) and a caller-owned device scalar, enqueued on
 * `stream` (a hipStream_t; NULL: the null stream); same kernels and grid as
 * gs_encode_dot, so the same sum bit for bit. Its partial sums live in a
 * 16 KiB array of the library in each GPU's memory (2048 partials: the
 * grid never exceeds 2048 workgroups, so every n fits; negative n is an
 * error). Calls on one GPU share that array: calls that overlap in time on
 * one GPU (different streams, no ordering between them) race. */
int gs_dot_noalloc(int32_t dtype, const void *x, const void *y, int64_t n, void *out,
                   void *stream);
int gs_encode_scalar_divide(gs_encoder_h e, gs_dscalar_h num, gs_dscalar_h den, gs_dscalar_h out);
int gs_encode_scalar_multiply(gs_encoder_h e, gs_dscalar_h a, gs_dscalar_h b, gs_dscalar_h out);
int gs_encode_scalar_set(gs_encoder_h e, double value, gs_dscalar_h out);
int gs_encode_scalar_copy(gs_encoder_h e, gs_dscalar_h in, gs_dscalar_h out);
int gs_encode_scalar_negate(gs_encoder_h e, gs_dscalar_h in, gs_dscalar_h out);
int gs_encode_scalar_sqrt(gs_encoder_h e, gs_dscalar_h in, gs_dscalar_h out);
int gs_encode_apply_block_inverses3x3(gs_encoder_h e, gs_dvec_h inverse_blocks, gs_dvec_h residual,
                                      gs_dvec_h correction);
/* (begin, end) as declared in Backend.h:110-114 and implemented by
 * src/CPU/CPUBackend.cc:364. */
int gs_encode_block_gauss_seidel_color3x3(gs_encoder_h e, gs_dcsr_h A, gs_divec_h color_blocks,
                                          gs_dvec_h inverse_blocks, int64_t begin, int64_t end,
                                          gs_dvec_h rhs, gs_dvec_h x, double omega);

/* ------------------------------------------------------------------ */
/* Learned per-level smoother weights. No counterpart in the reference: */
/* a small MLP (in -> hidden -> hidden -> out, tanh) whose forward,     */
/* backward (MSE) and Adam steps run as HIP kernels in FP64.            */
/* ------------------------------------------------------------------ */
/* Five features per level (mean row nnz, diagonal dominance, coarsening
 * ratio, coefficient contrast, colour count); features holds 5 * levels. */
int gs_hierarchy_features(gs_hier_h h, int32_t block_size, double *features);
typedef struct gs_mlp_s *gs_mlp_h;
int gs_mlp_create(int32_t inputs, int32_t hidden, int32_t outputs, uint64_t seed, gs_mlp_h *out);
int gs_mlp_num_params(gs_mlp_h h, int64_t *n);
int gs_mlp_get_params(gs_mlp_h h, double *params);
int gs_mlp_set_params(gs_mlp_h h, const double *params);
int gs_mlp_forward(gs_mlp_h h, const double *X, int64_t batch, double *Y);
/* Loss 0.5 / B * sum (y - t)^2 and its gradient with respect to the
 * parameters (flat layout W1, b1, W2, b2, W3, b3, row-major). */
int gs_mlp_loss_grad(gs_mlp_h h, const double *X, const double *T, int64_t batch, double *loss,
                     double *grad);
/* One Adam update with the gradient of the last gs_mlp_loss_grad. */
int gs_mlp_adam_step(gs_mlp_h h, double lr, double beta1, double beta2, double eps);
int gs_mlp_train(gs_mlp_h h, const double *X, const double *T, int64_t batch, int64_t epochs,
                 double lr, double *final_loss);
int gs_mlp_destroy(gs_mlp_h h);

#ifdef __cplusplus
}
#endif

#endif /* GPUSOLVER_H */
