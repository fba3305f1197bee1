/*
 * meshfemsparse_amd.h -- C ABI of the MI355X (gfx950) per-Newton-step linear-solve pipeline.
 *
 * Three stages, each a hand-written HIP kernel family behind this boundary:
 *   1. block-CSC Hessian assembly (3x3-block scatter-add through precomputed slot maps),
 *   2. supernodal Cholesky numeric factorization (ordering + symbolic analysis on the host, once),
 *   3. supernodal forward/back solves (+ symmetric block SpMV and FP64 iterative refinement).
 *
 * Conventions
 *   - Every entry point returns an int status: MFS_OK (0) or a negative MFS_ERR_* code;
 *     mfs_last_error() returns a thread-local message for the most recent failure.
 *   - Host arrays are caller-owned and only read during the call.
 *   - Arguments named *_dev are device pointers (hipMalloc'd / torch CUDA tensors) on the
 *     current HIP device; `stream` is a hipStream_t (NULL = default stream).  Device work is
 *     enqueued on `stream`; calls that must report a result to the host (numeric factorization
 *     pivot status, refinement residuals) synchronize that stream before returning.
 *   - Sparse pattern arrays follow the reference's SuiteSparse_long convention (int64).
 *   - Block-CSC value layout ("ContiguousBlocks"): upper-triangle block pattern Ap/Ai over block
 *     indices; block k occupies Ax[9k .. 9k+8] as a column-major 3x3; the diagonal block is the
 *     last entry of its column and only its upper triangle is meaningful.
 *   - A handle must not be used by two host threads at once (it owns mutable scratch), matching
 *     the reference factorizer's threading contract.
 */
#ifndef MESHFEMSPARSE_AMD_H
#define MESHFEMSPARSE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ */
/* Status codes                                                                                */
/* ------------------------------------------------------------------------------------------ */
#define MFS_OK               0
#define MFS_ERR_ARG         -1  /* invalid argument / size mismatch      (std::invalid_argument) */
#define MFS_ERR_STATE       -2  /* missing factorization / wrong phase / internal failure       */
                                /*                                        (std::runtime_error)    */
#define MFS_ERR_NOT_SPD     -3  /* pivot failure; see mfs_num_successful_pivots                  */
#define MFS_ERR_DEVICE      -4  /* HIP runtime error                                             */
#define MFS_ERR_ALLOC       -5  /* host or device allocation failure                             */
#define MFS_ERR_UNSUPPORTED -6  /* e.g. non-uniform block sizes, partial pins in the learned     */
                                /* preconditioner                                                */
#define MFS_ERR_NOT_CONVERGED -7 /* iterative solve stopped at its iteration limit               */

/* Factor precisions */
#define MFS_PREC_FP64 0         /* plain FP64 path: the drop-in                                   */
#define MFS_PREC_FP32 1         /* FP32 factor after symmetric diagonal scaling, FP64 refinement  */
#define MFS_PREC_BF16 2         /* as FP32, Schur updates on bfloat16 MFMA (FP32 accumulation)    */
#define MFS_PREC_BF16X3 3       /* as BF16 with each operand split hi + lo (three bf16 products)  */

/* Element-Hessian layouts accepted by mfs_assemble_hessian */
#define MFS_HE_FULL12   0       /* 144 doubles per element: 12x12 column-major (eval_He(ei))      */
#define MFS_HE_UPPER78  1       /* 78 doubles per element: local-upper blocks, packed (see below) */

typedef struct mfs_pattern_s*    mfs_pattern;
typedef struct mfs_assembler_s*  mfs_assembler;
typedef struct mfs_factorizer_s* mfs_factorizer;
typedef struct mfs_learned_s*    mfs_learned;
typedef struct mfs_nd_s*         mfs_nd;

const char* mfs_last_error(void);
int         mfs_version(void);

/* ------------------------------------------------------------------------------------------ */
/* Block sparsity pattern                                                                      */
/* Replaces SystemAssembler<3>::blockSparsityPattern (src/lib/MeshFEMSparse/SystemAssembler.hh:251-465): */
/* per-element sorted stencils -> column buckets (rows <= column) -> dedupe + sort -> Ap/Ai.   */
/* ------------------------------------------------------------------------------------------ */
int mfs_pattern_build(const int32_t* elem_nodes, int64_t num_elems, int32_t nodes_per_elem,
                      int64_t num_blocks, mfs_pattern* out);
/* nz = number of stored upper blocks; Ap has num_blocks+1 entries, Ai has nz entries.       */
int mfs_pattern_get(mfs_pattern p, int64_t* num_blocks, int64_t* nz,
                    const int64_t** Ap, const int64_t** Ai);
int mfs_pattern_destroy(mfs_pattern p);
/* The same construction on the device (SystemAssembler.hh:251-465 buckets, built by GPU kernels):
 * elem_nodes_dev is device int32 [num_elems * nodes_per_elem]; the result equals
 * mfs_pattern_build's.  mfs_pattern_get returns host copies; mfs_pattern_get_device the device
 * arrays (owned by the handle).  Synchronizes the stream.                                     */
int mfs_pattern_build_device(const int32_t* elem_nodes_dev, int64_t num_elems, int32_t nodes_per_elem,
                             int64_t num_blocks, void* stream, mfs_pattern* out);
int mfs_pattern_get_device(mfs_pattern p, const int64_t** Ap_dev, const int64_t** Ai_dev);

/* ------------------------------------------------------------------------------------------ */
/* Hessian assembly                                                                            */
/* Replaces SystemAssembler<3>::assembleHessian -> assembleHessianBlockAccelerated ->          */
/* ElementHessianContribAssembler<true>::run (SystemAssembler.hh:60-97, 524-565).             */
/* Accumulates: Ax += sum_e scatter(He_e), reading only the local-upper blocks of He_e        */
/* (off-diagonal (a,b) with a<b, transposed when the global order is reversed; upper triangle */
/* of the 4 diagonal blocks).  Contributions to each 3x3 slot are added in ascending element  */
/* order, so the result equals a serial element-order run of the reference bit for bit.       */
/* MFS_HE_UPPER78 packing per element: the 6 off-diagonal blocks (a,b), a<b in order           */
/* (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) as column-major 3x3 (54 doubles), then the 4 diagonal  */
/* blocks' upper triangles in column-major order (0,0),(0,1),(1,1),(0,2),(1,2),(2,2) (24).    */
/* ------------------------------------------------------------------------------------------ */
int mfs_assembler_create(const int64_t* Ap, const int64_t* Ai, int64_t num_blocks,
                         const int32_t* elem_nodes, int64_t num_elems, int32_t nodes_per_elem,
                         mfs_assembler* out);
int mfs_assemble_hessian(mfs_assembler a, const double* He_dev, int he_layout,
                         double* Ax_dev, void* stream);
/* The same into the non-contiguous block layout, BlockCSCHessian<..., ContiguousBlocks=false> */
/* (BlockCSCHessian.hh:52-97; StoreFullDiagonalBlocks = ContiguousBlocks, :570): Ax is the      */
/* scalar upper-triangle CSC of the 3n x 3n matrix (9*nz - 3*n values; block column j at        */
/* 9*Ap[j] - 3*j, scalar column 3j+c holding 3 rows per off-diagonal block, then rows 3j..3j+c  */
/* of the diagonal block).  Every column's last block must be its diagonal block               */
/* (MFS_ERR_ARG otherwise).  Same element order and sums as mfs_assemble_hessian.               */
int mfs_assemble_hessian_noncontiguous(mfs_assembler a, const double* He_dev, int he_layout,
                                       double* Ax_dev, void* stream);
/* Ax = 0 over the 9*nz values (BlockCSCHessianBase::setZero, BlockCSCHessian.hh:638-649).    */
int mfs_assembler_zero(mfs_assembler a, double* Ax_dev, void* stream);
int mfs_assembler_destroy(mfs_assembler a);

/* Element kernel (the reference's caller-supplied eval_He / eval_ge, SystemAssembler.hh:559-565,
 * 700-735, computed on the device): compressible neo-Hookean
 * psi = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2 per linear tet of the assembler's mesh, at
 * rest positions Xr_dev and current positions X_dev (num_blocks x 3).  E_dev / nu_dev per element
 * may be NULL (then E, nu).  He_dev (layout MFS_HE_FULL12 or MFS_HE_UPPER78) and g12_dev (12 per
 * element: dE/dx of the 4 nodes) may each be NULL.  The math is the host generator's.       */
int mfs_neohookean_elements(mfs_assembler a, const double* Xr_dev, const double* X_dev, const double* E_dev,
                            const double* nu_dev, double E, double nu, double* He_dev, int32_t he_layout,
                            double* g12_dev, void* stream);
/* assembleGradient (SystemAssembler.hh:700-735, the gather variant :785-883): accumulates
 * g[3 v + i] += sum over the elements of node v, in ascending element order, of g12[12 e + 3 a + i]
 * (a = the node's local index), i.e. the result of a serial element-order loop.              */
int mfs_assemble_gradient(mfs_assembler a, const double* g12_dev, double* g_dev, void* stream);

/* ------------------------------------------------------------------------------------------ */
/* Factorizer                                                                                  */
/* Replaces CholeskyFactorizerBase / CatamariFactorizer                                        */
/* (src/lib/MeshFEMSparse/Solvers/CholeskyFactorizerBase.hh:50-443,                            */
/*  src/lib/MeshFEMSparse/Solvers/CatamariFactorizer.cc:136-948).                              */
/* ------------------------------------------------------------------------------------------ */
int mfs_factorizer_create(int precision, mfs_factorizer* out);
int mfs_factorizer_destroy(mfs_factorizer f);

/* factorizeSymbolic(const BlockCSCHessianBase&, pinnedScalarVars) (CatamariFactorizer.cc:136-616):
 * whole-block pins -> block row/col removal (SparseMatrices.hh:2020-2076) -> nested-dissection
 * ordering -> elimination tree, relaxed supernodes, row structures, extend-add maps, injection
 * plan, level schedule.  Host arrays.  block_size must be 3.  A partially pinned block stays in
 * the block pattern with its pinned components replaced by the identity at injection; the
 * reference falls back to a scalar factorization there (CatamariFactorizer.cc:207-216); both
 * factor the same reduced matrix, so solves, sizes (n_reduced = free scalars), the reduced and
 * permuted vectors and the exported factor are those of the free scalars.                    */
int mfs_factorize_symbolic(mfs_factorizer f, const int64_t* Ap, const int64_t* Ai,
                           int64_t num_blocks, int32_t block_size,
                           const int64_t* pinned_scalar_vars, int64_t num_pins);

/* factorizeNumeric / factorizeNumericWithShift(A, sigma[, B]) (CholeskyFactorizerBase.hh:93-110,
 * CatamariFactorizer.cc:618-714): factor P(A + sigma*I)P^T, or P(A + sigma*B)P^T when Bx_dev is
 * non-NULL (B has A's pattern and layout).  Ax_dev holds the unreduced 9*nz block values.
 * On a pivot failure returns MFS_ERR_NOT_SPD and drops back to the symbolic state.           */
int mfs_factorize_numeric(mfs_factorizer f, const double* Ax_dev, double sigma,
                          const double* Bx_dev, void* stream);

/* The SuiteSparseMatrix overloads of CholeskyFactorizerBase (CholeskyFactorizerBase.hh:95-113:
 * factorizeSymbolic(const SuiteSparseMatrix&, pins), factorizeNumeric(const SuiteSparseMatrix&),
 * factorizeNumericWithShift(A, sigma[, B])): a scalar symmetric matrix as its upper triangle in
 * CSC (SuiteSparse UPPER_TRIANGLE; Ap[n+1], Ai[nnz] int64, rows sorted, row <= column).  Scalars
 * are grouped three to a block (the last block padded; padding is masked like a pinned scalar)
 * and factored by the block path.  After mfs_factorize_symbolic_scalar the factorizer takes the
 * matrix's nnz scalar values (device) in mfs_factorize_numeric_scalar, mfs_spmv and
 * mfs_solve_refined; full-length vectors have n entries.  Bx_dev (may be NULL) holds the values
 * of a matrix with A's pattern.                                                               */
int mfs_factorize_symbolic_scalar(mfs_factorizer f, const int64_t* Ap, const int64_t* Ai, int64_t n,
                                  const int64_t* pinned_scalar_vars, int64_t num_pins);
int mfs_factorize_numeric_scalar(mfs_factorizer f, const double* Ax_dev, double sigma,
                                 const double* Bx_dev, void* stream);

/* CholeskyFactorizerBase::solve (CholeskyFactorizerBase.hh:262-287) with the fused pin removal
 * and permutation (removeFixedEntries/extractFullSolution, :298-363): b, x are full-length
 * (num_blocks*3 per column, column-major, nrhs columns); pinned entries of x are set to 0.
 * solveMultiRHS (CatamariFactorizer.cc:862-907) is the nrhs > 1 case.  b and x may alias.   */
int mfs_solve(mfs_factorizer f, const double* b_dev, double* x_dev, int32_t nrhs, void* stream);

/* mfs_solve without any device allocation.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels as mfs_solve, with every scratch vector (factor-order
 * right-hand sides, forward results, update vectors, backward partial sums) carved out of a
 * library-owned static device array of MFS_SOLVE_NOALLOC_MAX_BYTES per GPU (96 MiB: one
 * right-hand side of the 64^3 cube, 811,200 unknowns, needs 68.4 MB in FP64).  A solve whose
 * scratch exceeds it returns MFS_ERR_ARG with the size it needs.  The array keeps the last call's
 * contents and is shared by every call on that GPU: calls that overlap on one GPU race.  Kernel
 * profiling (mfs_kernel_profiling) must be off.                                               */
#define MFS_SOLVE_NOALLOC_MAX_BYTES ((size_t)96 << 20)
int mfs_solve_noalloc(mfs_factorizer f, const double* b_dev, double* x_dev, int32_t nrhs, void* stream);

/* solveRawReduced / solveRawReducedInPlace (CatamariFactorizer.cc:788-860): reduced-length
 * vectors; already_permuted = vectors are in factor order.  b and x may alias (in place).    */
int mfs_solve_raw_reduced(mfs_factorizer f, const double* b_dev, double* x_dev, int32_t nrhs,
                          int32_t already_permuted, void* stream);

/* removeFixedEntries / extractFullSolution (CholeskyFactorizerBase.hh:298-363).              */
int mfs_remove_fixed_entries(mfs_factorizer f, const double* x_full_dev, double* x_reduced_dev,
                             int32_t nrhs, int32_t permute, void* stream);
int mfs_extract_full_solution(mfs_factorizer f, const double* x_reduced_dev, double* x_full_dev,
                              int32_t nrhs, int32_t permute, void* stream);

/* Symmetric block SpMV from the upper blocks: y = A x over the full (unreduced) pattern the
 * factorizer was analysed with (BlockCSCHessian::applyRaw, BlockCSCHessian.hh:1059-1120).
 * Each output row sums its blocks in a fixed order (deterministic).  After a scalar analysis
 * Ax_dev holds the scalar values and x, y have n entries.                                     */
int mfs_spmv(mfs_factorizer f, const double* Ax_dev, const double* x_dev, double* y_dev,
             void* stream);

/* Solve A x = b (full-length, one RHS) to ||b - A x||/||b|| <= tol over the free rows, by the
 * factor plus FP64 iterative refinement with the FP64 SpMV residual.  For an FP32 factor this
 * is the mixed-precision path; for FP64 it verifies (and if needed refines) the direct solve.
 * iters = refinement corrections applied after the first solve; relres = final residual.   */
int mfs_solve_refined(mfs_factorizer f, const double* Ax_dev, const double* b_dev, double* x_dev,
                      double tol, int32_t max_iters, int32_t* iters, double* relres, void* stream);

/* Batched systems (many independent meshes assembled as one block-diagonal matrix; each mesh is
 * a connected component and gets its own nested dissection): refinement until every segment
 * [seg_ptr[s], seg_ptr[s+1]) of the full-length vectors has ||b_s - (A x)_s|| / ||b_s|| <= tol.
 * seg_ptr: host, nseg + 1 entries, 0 .. 3 num_blocks.  relres_seg (host, nseg) may be NULL.
 * No reference counterpart (the reference factors one system at a time).                   */
int mfs_solve_refined_segments(mfs_factorizer f, const double* Ax_dev, const double* b_dev, double* x_dev,
                               int64_t nseg, const int64_t* seg_ptr, double tol, int32_t max_iters,
                               int32_t* iters, double* relres_max, double* relres_seg, void* stream);

/* Capture the numeric factorization (all kernels from the zero fill to the block inverses) in a
 * hipGraph on the library's own stream, joined to `stream` by events, and replay it on later
 * calls with the same Ax_dev, Bx_dev and sigma (re-captured when they or the schedule change).
 * No reference counterpart.                                                                  */
int mfs_factorizer_set_graph(mfs_factorizer f, int32_t on);

/* Autotuner (no reference counterpart): analyse once for the ordering, then choose the schedule
 * by timing `reps` numeric factorizations of Ax_dev (the fastest of them, after one warm-up) per
 * candidate, by coordinate descent: first the fused level kernel's bound fuse[0..nfuse) (supernodes
 * one diagonal tile wide with at most that many structure rows, scalars, are factored by the
 * persistent level kernel; 0: none), then for each supernode size bucket (front W + D < 512,
 * < 2048, larger) the largest GEMM tile among tiles[0..ntile) (64: 64x64 tiles only; 128: 128x128
 * tiles where the region is large enough).  The first candidates are the starting point, and a
 * candidate replaces the incumbent only when it is at least 2 % faster.  ms[nfuse + 3 ntile] =
 * milliseconds of every candidate (fuse bounds, then bucket 0, 1, 2 tiles); best[4] = {fuse bound,
 * tile of bucket 0, 1, 2}.  The chosen schedule stays installed (symbolic state: factorize again
 * before solving).                                                                            */
int mfs_autotune(mfs_factorizer f, const int64_t* Ap, const int64_t* Ai, int64_t num_blocks, int32_t block_size,
                 const int64_t* pinned_scalar_vars, int64_t num_pins, const double* Ax_dev, const int32_t* fuse,
                 int32_t nfuse, const int32_t* tiles, int32_t ntile, int32_t reps, double* ms, int32_t* best,
                 void* stream);

/* Schedule statistics: out9 = {fused tasks, fused-kernel launches (levels), tiled supernodes,
 * separate POTRF launches, GEMM launches, extend-add launches, levels, nodes of the captured
 * graph, fused diagonal-tile + TRSM launches}.                                                */
int mfs_schedule_stats(mfs_factorizer f, int64_t* out9);

/* State / statistics (CholeskyFactorizerBase.hh:55-59,193-260; CatamariFactorizer.cc:716-948). */
int mfs_factorization_type(mfs_factorizer f, int32_t* type);   /* 0 none, 1 symbolic, 2 numeric */
int mfs_num_successful_pivots(mfs_factorizer f, int64_t* n);
int mfs_sizes(mfs_factorizer f, int64_t* n_full, int64_t* n_reduced);
int mfs_factor_nnz(mfs_factorizer f, int64_t* nnz);             /* scalar nnz(L), incl. diagonal */
int mfs_flop_estimate(mfs_factorizer f, double* flops);         /* POTRF+TRSM+SYRK, per supernode */
int mfs_get_inverse_permutation(mfs_factorizer f, int64_t* iperm_scalar); /* new -> old reduced */
/* out3 = {supernodal flops (= mfs_flop_estimate), the diagonal-tile POTRF share of it,
 *         sum over scalar columns of c_j^2 for the exact (unrelaxed) column counts}          */
int mfs_flop_breakdown(mfs_factorizer f, double* out3);
int mfs_clear_factors(mfs_factorizer f);
int mfs_stash_factorization(mfs_factorizer f);
int mfs_has_stashed_factorization(mfs_factorizer f, int32_t* has);
int mfs_swap_stashed_factorization(mfs_factorizer f);
int mfs_clear_stashed_factorization(mfs_factorizer f);

/* Supernode statistics of the symbolic analysis: number of supernodes, elimination-tree
 * levels, max supernode width and max row count below a supernode (scalars).               */
int mfs_supernode_stats(mfs_factorizer f, int64_t* num_supernodes, int64_t* num_levels,
                        int64_t* max_width, int64_t* max_degree);

/* Separator tree of the most recent nested-dissection analysis (CholeskyFactorizerBase.hh:51-59,
 * NDOrdering {CParent, CMember, blockSize}): components are postordered (CParent[c] > c, or -1
 * for a root); CMember[v] is the component of reduced ordering variable v (unpermuted reduced
 * block: pins removed), block_size = 3.  Query the sizes with CParent = CMember = NULL.
 * MFS_ERR_STATE when no symbolic analysis exists (or its ordering was not a dissection).      */
int mfs_nd_ordering(mfs_factorizer f, int64_t* num_components, int64_t* num_vars, int64_t* block_size,
                    int64_t* CParent, int64_t* CMember);

/* Temporal reuse of the separator tree between symbolic analyses (CatamariFactorizer.hh:84-90;
 * ParallelNestedDissection/README.md:86-127): with period P > 0 the next P analyses repair only
 * the subtrees whose separators the new pattern violates, then the tree is rebuilt.  P = 0
 * (default) disables reuse; a new period or different pins restart the history.
 * stats4 = {full_rebuild (0/1), graph_vertices, recomputed_subtrees, repartitioned_vertices}
 * of the last analysis, nd_seconds its ordering time (cholmod_nesdis_parallel.hh:33-39).      */
int mfs_set_temporal_reuse_period(mfs_factorizer f, int64_t period);
int mfs_get_temporal_reuse_period(mfs_factorizer f, int64_t* period);
int mfs_reset_temporal_reuse(mfs_factorizer f);
int mfs_temporal_reuse_statistics(mfs_factorizer f, int64_t* stats4, double* nd_seconds);

/* Host nested dissection of a symmetric graph with its own reuse history, as the factorizer
 * uses it (nested_dissection_from_graph + TemporalReuseState, cholmod_nesdis_parallel.hh:41-75).
 * Graph: CSC of both triangles (p[n+1], i[p[n]]; diagonal entries ignored; a graph that does not
 * list every edge in both directions is rejected with MFS_ERR_ARG); vertex ids must keep
 * their meaning between calls.  The dissection knobs of the environment (MFS_ND_*, MFS_LEAF)
 * apply as in the factorizer.  mfs_nd_order writes perm[new] = old and the number of
 * components; mfs_nd_forest then gives CParent[num_components] and CMember[n] of that call.   */
int mfs_nd_create(mfs_nd* out);
int mfs_nd_destroy(mfs_nd h);
int mfs_nd_set_temporal_reuse_period(mfs_nd h, int64_t period);
int mfs_nd_reset(mfs_nd h);
int mfs_nd_order(mfs_nd h, int64_t n, const int64_t* p, const int64_t* i, int64_t* perm,
                 int64_t* num_components);
int mfs_nd_forest(mfs_nd h, int64_t* CParent, int64_t* CMember);
int mfs_nd_statistics(mfs_nd h, int64_t* stats4, double* nd_seconds);

/* Host-only symbolic analysis (no device work): the ordering and supernode statistics the
 * factorizer would use (the MFS_* tuning variables of the environment apply to both; without
 * MFS_LEAF the dissection runs with leaf sizes 96 and 128 and keeps the ordering with fewer
 * unrelaxed factor flops).  iperm_blocks[n_reduced_blocks] = reduced block at factor position k;
 * stats[8] = {n_reduced_blocks, num_supernodes, num_levels, factor_nnz (scalar), max_width,
 * max_degree, L doubles, U arena doubles}; flops[2] = {supernodal POTRF+TRSM+SYRK flops,
 * sum over scalar columns of c_j^2 for the unrelaxed column counts c_j}.  Any output may be NULL. */
int mfs_analyze_host(const int64_t* Ap, const int64_t* Ai, int64_t num_blocks,
                     const int64_t* pinned_scalar_vars, int64_t num_pins, int64_t* iperm_blocks,
                     int64_t* stats, double* flops);

/* Host replay of the device factorization and solve schedules (same analysis, same task lists,
 * plain CPU loops): factor of P(A + sigma I)P^T as CSC like mfs_factor_export_csc (sizes from
 * mfs_analyze_host stats[3]); failed_col = first non-positive pivot column or -1; if x_perm is
 * non-NULL it holds a factor-order right-hand side (3 n_reduced) and is overwritten by the
 * solution of the replayed forward/backward sweeps.  Testing.                               */
int mfs_debug_emulate_factor(const int64_t* Ap, const int64_t* Ai, int64_t num_blocks,
                             const int64_t* pinned_scalar_vars, int64_t num_pins, const double* Ax,
                             double sigma, int64_t* Lp, int64_t* Li, double* Lx, int64_t* failed_col,
                             double* x_perm);

/* Copy the numeric factor L (lower, factor order, incl. diagonal; explicit zeros from supernode
 * relaxation included) to host CSC arrays: Lp[n_reduced+1], Li[nnz], Lx[nnz].  For testing.  */
int mfs_factor_export_csc(mfs_factorizer f, int64_t* Lp, int64_t* Li, double* Lx);

/* Per-stage device time of the most recent numeric factorization / solve, milliseconds, from
 * HIP events on the caller's stream: [0] injection, [1] extend-add, [2] dense panels, [3] total.*/
int mfs_last_timings(mfs_factorizer f, double* ms4);

/* Per-kernel-class device time, accumulated from HIP event pairs placed around every launch while
 * enabled (numeric factorization and mfs_solve_refined; the events cost a few microseconds each).
 * Classes (21): fill L, inject, (unused), extend-add, POTRF tiles, TRSM GEMM tiles, in-panel
 * update GEMM tiles, Schur-update GEMM tiles, forward gather, forward diagonal, forward update
 * GEMV, backward partial GEMV^T, backward diagonal, SpMV, permutation maps, assembly, solve-block
 * inverses, backward gather, precision conversions, refinement vector kernels, fused level
 * kernel, fused diagonal-tile + TRSM steps.  ms[22] and counts[22].  mfs_kernel_profile waits  *
 * for profiled launches still running and folds them in.                                      */
int mfs_kernel_profiling(mfs_factorizer f, int32_t enable, int32_t reset);
int mfs_kernel_profile(mfs_factorizer f, double* ms, int64_t* counts);
/* Executed tile flops of the profiled launches per class (22; the GEMM classes and the fused   */
/* diagonal-tile + TRSM steps carry them).                                                     */
int mfs_kernel_profile_work(mfs_factorizer f, double* work);
/* Bytes the profiled GEMM launches of each class (22) would move with no cache reuse between their
 * tiles: every tile reads its A and B panels and writes its C (reading it first unless beta = 0). */
int mfs_kernel_profile_bytes(mfs_factorizer f, double* bytes);

/* ------------------------------------------------------------------------------------------ */
/* Learned block incomplete-Cholesky preconditioner (no reference counterpart: the reference     */
/* preconditions only with exact factorizations).  A message-passing network (hidden 64, 4       */
/* rounds, bfloat16-rounded weights, FP32 accumulation) over the free-node block graph emits a   */
/* zero-fill lower block factor L on tril(A) in a greedy multicolor ordering; M = L L^T          */
/* preconditions FP32 CG on the Jacobi-scaled system S A S, inside FP64 refinement.  Training on  */
/* the device: loss mean_p |(L L^T - S A S) z_p|^2 / |S A S z_p|^2 over caller probes z_p, own   */
/* forward/backward kernels, Adam.  Pins must be whole blocks (MFS_ERR_UNSUPPORTED otherwise).  */
/* ------------------------------------------------------------------------------------------ */
int mfs_learned_create(const int64_t* Ap, const int64_t* Ai, int64_t num_blocks, const int64_t* pinned_scalar_vars,
                       int64_t num_pins, uint64_t seed, mfs_learned* out);
int mfs_learned_destroy(mfs_learned l);
/* out8 = {free nodes, directed edges, lower blocks, colors, parameters, Adam steps, hidden, rounds} */
int mfs_learned_info(mfs_learned l, int64_t* out8);
/* offsets (floats) of the 22 parameter tensors W0 b0 {Wm bm Wu bu} x 4 Wd1 bd1 Wd2 bd2 Wdd bdd, and the end */
int mfs_learned_param_offsets(mfs_learned l, int64_t* out23);
/* graph arrays (any may be NULL): edge destination/source (free-node ids, grouped by destination,
 * sources ascending), the edge, row and column of each lower block, node colors              */
int mfs_learned_graph(mfs_learned l, int32_t* edge_dst, int32_t* edge_src, int32_t* lower_edge, int32_t* lower_dst,
                      int32_t* lower_src, int32_t* color);
/* the same graph from the pattern on the host (no device): sizes4 = {nodes, edges, lower blocks,
 * colors}; call with NULL arrays first for the sizes                                          */
int mfs_learned_analyze_host(const int64_t* Ap, const int64_t* Ai, int64_t num_blocks,
                             const int64_t* pinned_scalar_vars, int64_t num_pins, int64_t* sizes4,
                             int32_t* edge_dst, int32_t* edge_src, int32_t* lower_edge, int32_t* lower_dst,
                             int32_t* lower_src, int32_t* color);
/* scaling, scaled blocks and node features from the assembled values (9 nz doubles, device)   */
int mfs_learned_set_matrix(mfs_learned l, const double* Ax_dev, void* stream);
/* host copies: node features (6 per node), scaled edge blocks A^(dst, src) (9 per edge,
 * column-major), scaled diagonal blocks (9 per node), scaling (3 per node); any may be NULL   */
int mfs_learned_features(mfs_learned l, float* node_feat, float* edge_feat, float* diag, double* scale);
/* loss for host probes z (nprobe x 3 n floats, free-node order); backward != 0 also fills the
 * parameter gradient                                                                          */
int mfs_learned_loss(mfs_learned l, const float* z, int32_t nprobe, int32_t backward, double* loss, void* stream);
int mfs_learned_adam_step(mfs_learned l, float lr, void* stream);
/* master parameters (FP32): copy out to `get` and/or in from `set` (host arrays, nparams)     */
int mfs_learned_params(mfs_learned l, float* get, const float* set);
int mfs_learned_grads(mfs_learned l, float* out);
/* network forward for the current matrix; host copies of the lower blocks (9 per block) and the
 * diagonal blocks (9 per node), column-major; either may be NULL                              */
int mfs_learned_factor(mfs_learned l, float* L_lower, float* L_diag, void* stream);
/* A x = b (full-length device vectors, pins -> 0): FP32 CG preconditioned by block Jacobi
 * (precond 0) or the network's factor (precond 1) inside FP64 refinement, until the unscaled
 * relative residual <= tol; after max_pcg CG iterations in total the fallback factorizer (with
 * a symbolic factorization of A) factors and solves exactly.  Without a fallback, stopping at
 * the limit returns MFS_ERR_NOT_CONVERGED with x the last iterate.                           */
int mfs_learned_solve(mfs_learned l, mfs_factorizer fallback, const double* Ax_dev, const double* b_dev,
                      double* x_dev, int32_t precond, double tol, int32_t max_pcg, int32_t* pcg_iters,
                      int32_t* outer_iters, double* relres, int32_t* used_fallback, void* stream);

/* ------------------------------------------------------------------------------------------ */
/* BlockCSCHessian binary files (BlockCSCHessianBase::dumpBinaryToStream / constructFromBinaryStream, */
/* BlockCSCHessian.cc:11-103; MatrixRecorder.hh:36-72 names the sequences): the reference's format   */
/* (size_t block size 3, size_t blocks, int64 m n nz snz, uint32 symmetry 1, Ap, Ai, Ax in the      */
/* non-contiguous layout = scalar CSC of the upper triangle).  In memory Ax is contiguous (9 per    */
/* block, strict lower of diagonal blocks 0 on read).  Ax NULL writes / reads the pattern only;    */
/* read with NULL arrays returns the sizes (snz = 0: sparsity only).                              */
/* ------------------------------------------------------------------------------------------ */
int mfs_blockcsc_write(const char* path, const int64_t* Ap, const int64_t* Ai, int64_t num_blocks, const double* Ax);
int mfs_blockcsc_read(const char* path, int64_t* num_blocks, int64_t* nz, int64_t* snz, int64_t* Ap, int64_t* Ai,
                      double* Ax);

/* ------------------------------------------------------------------------------------------ */
/* Workload generator (the reference has no energy code; the caller supplies He).             */
/* Unit cube, N^3 hexes, each split into 6 Kuhn tets; node id = x + (N+1)(y + (N+1) z).       */
/* ------------------------------------------------------------------------------------------ */
int mfs_cube_mesh(int32_t N, int32_t* tets /* 6N^3 x 4 */, double* X /* (N+1)^3 x 3 */);
/* X += sigma_rel * h * N(0,1) per coordinate, from a splitmix64 stream seeded by `seed`.    */
int mfs_perturb_nodes(int32_t N, uint64_t seed, double sigma_rel, double* X);
/* Compressible neo-Hookean psi = mu/2 (I1 - 3) - mu ln J + lambda/2 (ln J)^2 per linear tet:
 * element Hessians (144 doubles each, col-major 12x12, dof 3a+i) and/or the assembled
 * gradient (num_nodes*3).  E_per_elem/nu_per_elem may be NULL to use E, nu.  Host, threaded.*/
int mfs_neohookean_host(const double* X_rest, const double* X, const int32_t* tets, int64_t ne,
                        const double* E_per_elem, const double* nu_per_elem, double E, double nu,
                        double* He_out, double* grad_out, double* energy_out, int32_t num_threads);

/* ------------------------------------------------------------------------------------------ */
/* Test hooks: the device kernels' building blocks exercised directly.                        */
/* ------------------------------------------------------------------------------------------ */
/* C(MxN) = alpha * A(MxK) * B(NxK)^T + beta * C, column-major, through the batched MFMA tile
 * kernel used by the factorization.                                                         */
int mfs_test_gemm_nt(int32_t M, int32_t N, int32_t K, double alpha, const double* A_dev,
                     int32_t lda, const double* B_dev, int32_t ldb, double beta, double* C_dev,
                     int32_t ldc, void* stream);
/* The same through the FP32 instantiation of the tile kernel (float operands).               */
int mfs_test_gemm_nt_f32(int32_t M, int32_t N, int32_t K, float alpha, const float* A_dev,
                         int32_t lda, const float* B_dev, int32_t ldb, float beta, float* C_dev,
                         int32_t ldc, void* stream);
/* The same through the bfloat16 tile kernel of the BF16 variants (float operands rounded to
 * bfloat16, FP32 accumulation; split != 0: hi + lo operand split, three products).          */
int mfs_test_gemm_nt_bf16(int32_t M, int32_t N, int32_t K, float alpha, const float* A_dev,
                          int32_t lda, const float* B_dev, int32_t ldb, float beta, float* C_dev,
                          int32_t ldc, int32_t split, void* stream);
/* trs_kernels::MultiplyLowerBlock semantics (block-compressed row indices, BS=1..3):
 *   B(row_i, :) -= A(i, :) * B(supernode_start : supernode_start + size, :),
 *   row_i = indices[i / BS] + i % BS  for i < degree,
 * through the row-dot device routine of the forward-solve update kernel.                    */
int mfs_test_multiply_lower_block(int32_t BS, const int64_t* indices_dev, int64_t supernode_start,
                                  int64_t supernode_size, int64_t degree, const double* A_dev,
                                  int64_t lda, int32_t nrhs, double* B_dev, int64_t ldb,
                                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MESHFEMSPARSE_AMD_H */
