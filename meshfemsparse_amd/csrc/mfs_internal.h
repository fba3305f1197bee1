// Internal host-side structures shared by the C ABI, the symbolic analysis and the device code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>

namespace mfs {

// Error raised inside the library; the C ABI turns it into a status code + message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// ------------------------------------------------------------------------------------------
// Block sparsity pattern (upper triangle, CSC over block indices, diagonal last per column).
// ------------------------------------------------------------------------------------------
struct Pattern {
    int64_t n = 0;
    std::vector<int64_t> Ap, Ai;
    int64_t nz() const { return n ? Ap[n] : 0; }
};

Pattern build_block_pattern(const int32_t* elem_nodes, int64_t ne, int npe, int64_t nblocks);

// ------------------------------------------------------------------------------------------
// Assembly slot map: for every destination block slot, the list of (element, local pair,
// transpose) contributions sorted by element.  Entry = (e << 5) | (pair << 1) | transpose.
// Pairs 0..5 are the off-diagonal local-upper blocks (0,1),(0,2),(0,3),(1,2),(1,3),(2,3);
// pairs 6..9 are the diagonal blocks 0..3 (upper triangle only).
// The slots are listed in processing order: groups of 256 consecutive columns, and inside each
// group the off-diagonal slots before the diagonal ones (ascending slot within each part), so
// that the slots a workgroup takes together have similar contribution counts.  seg, contrib and
// the per-position arrays follow that order.
// ------------------------------------------------------------------------------------------
struct SlotMap {
    int64_t nslots = 0;
    std::vector<int32_t> seg;          // nslots + 1 (by position)
    std::vector<uint32_t> contrib;     // seg[nslots]
    std::vector<int32_t> slot_of_pos;  // destination slot of each position
    std::vector<uint8_t> is_diag;      // per position
};
SlotMap build_slot_map(const Pattern& P, const int32_t* elem_nodes, int64_t ne, int npe);

// ------------------------------------------------------------------------------------------
// Nested-dissection ordering of a symmetric graph (CSR, no self loops).
// perm[new] = old.
// ------------------------------------------------------------------------------------------
struct NDOptions {
    int leaf_size = 96;
    int coarse_size = 120;
    double imbalance = 0.06;
    uint64_t seed = 0x5eed;
    // A full dissection orders every connected component `trials` times and keeps the ordering
    // with the fewest predicted factor flops.  Trial t uses leaf_alt for odd t, imbalance_alt for
    // t & 2, and seed + t / 4; trial 0 is the plain options above.
    int trials = 8;
    int nseps = 1;          // bisections tried at every level (the smallest separator is kept)
    int node_passes = 0;    // node-based FM passes on every vertex separator (0: none)
    int leaf_alt = 128;
    double imbalance_alt = 0.15;
};
// Predicted factor flops (3x3 blocks, unamalgamated columns: sum of c^2 + (c-1)^2 + (c-2)^2 over
// block column counts c, scalar) of ordering the vertex set `order` (ids into the CSR graph; edges
// to vertices outside the set are ignored) in that order.  pos: scratch of at least n entries,
// all -1 (restored).
double order_flops(const std::vector<int64_t>& xadj, const std::vector<int32_t>& adj,
                   const std::vector<int32_t>& order, std::vector<int32_t>& pos);
// Separator forest of a dissection, postordered (children before parents, parent[c] > c or -1):
// component c owns verts[c] (a separator or a leaf, in ordering order); member[v] = component
// of vertex v; the ordering is the postorder traversal (children in order, then own vertices).
struct SepForest {
    std::vector<int32_t> parent, roots, member;
    std::vector<std::vector<int32_t>> children, verts;
};
struct NDReuseStats {
    bool full_rebuild = false;
    int64_t graph_vertices = 0, recomputed_subtrees = 0, repartitioned_vertices = 0;
    double nd_seconds = 0;
};
// History of one ordering stream (vertex ids must keep their meaning between calls).
struct NDState {
    int64_t period = 0;            // 0: every call dissects from scratch
    int64_t since_rebuild = 0;     // incremental analyses since the last full dissection
    SepForest forest;              // forest of the last call
    NDReuseStats stats;            // statistics of the last call
    void reset() { since_rebuild = 0; forest = SepForest(); stats = NDReuseStats(); }
};
std::vector<int32_t> nested_dissection(int32_t n, const std::vector<int64_t>& xadj,
                                       const std::vector<int32_t>& adj, const NDOptions& opt,
                                       NDState* state);
std::vector<int32_t> nested_dissection(int32_t n, const std::vector<int64_t>& xadj,
                                       const std::vector<int32_t>& adj, const NDOptions& opt,
                                       SepForest* forest = nullptr);
std::vector<int32_t> forest_order(const SepForest& f);

// ------------------------------------------------------------------------------------------
// Symbolic analysis result (all indices in *block* units unless noted "scalar").
// ------------------------------------------------------------------------------------------
struct GemmTask {           // C(MxN) = alpha * A(MxK) * B(NxK)^T + beta * C ; column-major
    int64_t a_off, b_off, c_off;
    int32_t lda, ldb, ldc;
    int32_t M, N, K;
    int32_t flags;          // bit0: C lives in the update arena; bit1: beta = 0 (else 1);
                            // bit2: alpha = +1 (else -1); bit3: A/B live in the inverse scratch
    int32_t pad;
};
static_assert(sizeof(GemmTask) == 56, "GemmTask layout");
enum : int32_t { GT_C_IN_U = 1, GT_BETA0 = 2, GT_ALPHA_POS = 4, GT_B_IN_INV = 8 };
// A Schur-update tile (C in the update arena) may carry the children's update-matrix entries that
// land in it: `pad` is the first of (flags >> GT_PULL_SHIFT) PullEnt entries, in child order.
constexpr int GT_PULL_SHIFT = 8;

// One child's share of a Schur-update tile of its parent p: the child's update-matrix blocks
// (ib, jb) with ib in [ib0, ib0 + nib), jb in [jb0, jb0 + njb), ib >= jb, whose parent rows and
// columns fall in the tile.  Tile-local scalar row of child block row ib, component r:
// 3 rel[rel_off + ib] + r - base_r (columns likewise with base_c), base = 3 W_p + tile origin.
struct PullEnt {
    int64_t u_child;        // child's U offset in the update arena (ld = dc3)
    int32_t dc3;            // 3 x child degree
    int32_t rel_off;        // offset of the child's relative indices
    int32_t ib0, nib;       // child block rows (nib <= 128)
    int32_t jb0, njb;       // child block columns (njb <= 128)
    int32_t base_r, base_c;
};
static_assert(sizeof(PullEnt) == 40, "PullEnt layout");

// A tile (at most 64 x 64) of a parent's panel, lower part, with the children's update-matrix
// blocks that land in it in the panel's columns: the tile is L[c_off + c ldc + r], r < M, c < N;
// PullEnt entries [pull_begin, pull_begin + npull) in child order, base = the tile's origin in
// panel coordinates (scalar row 3 rel + r of the parent's block rows, own columns first).  The
// tile is rebuilt from its matrix values (TileInj entries [inj_begin, inj_begin + ninj): the
// injection entries whose 3x3 block meets it; every other entry of the tile is 0) plus the
// children's blocks, so its previous contents are never read.
struct PanelPullTask {
    int64_t c_off;
    int32_t ldc, M, N;
    int32_t pull_begin, npull;
    int32_t inj_begin, ninj;
    int32_t pad;
};
static_assert(sizeof(PanelPullTask) == 40, "PanelPullTask layout");
// An injection entry of a panel tile: plan index and the block's first scalar row / column
// relative to the tile's origin (-2 .. 63).
struct TileInj {
    int32_t plan;
    int16_t lr, lc;
};
static_assert(sizeof(TileInj) == 8, "TileInj layout");
// A tile (at most 64 x 64) of a panel that no children list rebuilds: L[off + c ld + r] = 0 for
// r < M, c < N at the start of a numeric factorization.
struct FillTile {
    int64_t off;
    int32_t ld, M, N;
    int32_t pad;
};
static_assert(sizeof(FillTile) == 24, "FillTile layout");

struct PotrfTask {          // factor the kb x kb diagonal tile at L[diag_off], ld; inverse -> slot
    int64_t diag_off;
    int32_t ld, kb;
    int32_t slot;
    int32_t col0;           // scalar column (factor order) of the tile's first column
};

struct PanelItem {          // one workgroup of a fused diagonal-tile + TRSM step: factor potrf
                            // task `potrf` (every item of that task redundantly, in LDS), then
                            // solve the rows of TRSM task `trsm` (-1: none); `count` items share
                            // the potrf task and the last of them to load the tile stores it
    int32_t potrf, trsm, count, pad;
};

struct ExtendTask {         // extend-add of child's update matrix into the parent front
    int64_t u_child;        // offset of child's U (D_c x D_c, ld = D_c) in the update arena
    int64_t p_panel;        // parent panel offset in L
    int64_t u_parent;       // parent U offset in arena (or -1 if parent has no U)
    int32_t dc;             // child degree (blocks)
    int32_t ldp;            // parent panel leading dimension (scalars)
    int32_t wp;             // parent width (blocks)
    int32_t dp;             // parent degree (blocks)
    int64_t rel_off;        // offset of child's relative indices (dc entries) in rel array
    int32_t col_block;      // column block kj of the child's U handled by this task
    int32_t row0;           // first scalar row of the column block handled by this task: the
                            // task covers rows [row0, min(3 dc, row0 + kExtendRows))
};
constexpr int32_t kExtendRows = 768;   // rows per extend-add task (a multiple of 3 and of 64)

// Supernodal solves work on column blocks of at most `solve_block` columns of a supernode.  Each
// block's diagonal part is inverted explicitly after the numeric factorization (X = L_bb^{-1},
// lower triangular, nb x nb column-major), so a forward step is y_b = X x_b followed by the
// update of every row below, and a backward step is x_b = X^T (y_b - L(rows, b)^T x(rows)).
struct SolveBlock {
    int64_t l_off;          // L offset of the block's first diagonal entry (c0, c0)
    int64_t xb_off;         // offset of the block inverse X (ld = nb)
    int64_t x_off;          // factor-order scalar offset of the block's first column
    int32_t ld, nb;         // panel leading dimension, block width (scalars)
    int32_t ts_off;         // index of the block's first 64-column tile in the tile-slot list
    int32_t part_begin, part_end;   // backward: partial-sum slots to subtract
    int32_t pad;
};
static_assert(sizeof(SolveBlock) == 48, "SolveBlock layout");

struct SolveTile {          // (block, 64-row or 64-column tile of the block)
    int32_t blk, t;
};

struct SolveRowTask {       // a row range [r0, r0 + nr) of a supernode panel against nc block columns
    int64_t l_off;          // L offset of (first row, first column)
    int64_t v_off;          // forward: offset of y_b in Y; backward: offset of x(first row) in its source
    int64_t d_off;          // forward: x / u offset of the first row; backward: slot * solve_block + column
    int32_t ld, nr, nc;
    int32_t mode;           // forward: 0 x -= L y, 1 u += L y; backward: 0 rows from x, 1 rows from gathered u
};
static_assert(sizeof(SolveRowTask) == 40, "SolveRowTask layout");

// A supernode factored entirely by one workgroup of the fused level kernel (width <= one
// diagonal tile): pull extend-add of its children's update matrices into the panel, diagonal
// tile POTRF + inverse, TRSM of the rows below, U = -L21 L21^T, then the children's columns that
// land in U.
struct FusedTask {
    int64_t panel_off;      // L offset of the panel ((W + D) x W, ld = W + D)
    int64_t u_off;          // U offset (D x D) or -1 when D == 0
    int32_t W, D;           // scalars
    int32_t child_begin, child_end;   // range in the child list
    int32_t slot;           // inverse slot of the diagonal tile
    int32_t col0;           // scalar factor column of the first column
    int32_t wblk;           // width in blocks
    int32_t pad;
};
static_assert(sizeof(FusedTask) == 48, "FusedTask layout");

struct TaskRange {          // [b, m): 128x128-tile GEMM tasks, [m, e): 64x64-tile tasks
    int64_t b = 0, m = 0, e = 0;
};

struct SolveGatherTask {    // forward: gather children update vectors into x / u of parent;
                            // backward: gather x of the supernode's structure rows into u
    int32_t sn;             // supernode
    int32_t child_begin, child_end;   // range in child list
    int32_t w, d;           // scalars
    int32_t pad;
    int64_t x_off;          // x offset of the supernode's first column
    int64_t u_off;          // u offset of the supernode (d scalars)
    int64_t rows_off;       // offset of the structure rows (blocks) in the row list
};

struct LevelPlan {
    // factorization
    std::pair<int64_t,int64_t> fused{0, 0};                  // range in fused tasks (one launch)
    std::vector<std::pair<int64_t,int64_t>> extend_passes;   // ranges into extend tasks (into panels)
    std::vector<std::pair<int64_t,int64_t>> extend_u_passes; // after the Schur update (into U)
    std::pair<int64_t,int64_t> panel_pulls{0, 0};            // the panel part as tile lists (one
                                                             // launch, in place of extend_passes)
    std::vector<std::pair<int64_t,int64_t>> potrf_steps;     // per step: range in potrf tasks
    std::vector<TaskRange> trsm_steps;                       // per step: gemm tasks
    std::vector<std::pair<int64_t,int64_t>> panel_steps;     // per step: range in panel items (empty:
                                                             // separate POTRF and TRSM launches)
    std::vector<TaskRange> update_steps;                     // per step: gemm tasks
    std::vector<TaskRange> rest_steps;                       // per step: trailing-update tasks that
                                                             // may overlap the next outer block
    std::vector<char> trailing_step;                         // per step: 1 if it closes an outer block
    TaskRange syrk;                                          // gemm tasks
    // solves: gather tasks (forward: children -> x/u; backward: x -> u), and per column-block
    // step the diagonal tiles and the row tasks (backward steps in execution order)
    std::pair<int64_t,int64_t> gather{0, 0};
    std::vector<std::pair<int64_t,int64_t>> fdiag_steps, fupd_steps;
    std::vector<std::pair<int64_t,int64_t>> bpart_steps, bdiag_steps;
};

struct Symbolic {
    // sizes
    int64_t n_full_blocks = 0;     // unreduced block count
    int64_t n_red = 0;             // reduced blocks
    int64_t nz_full = 0;           // unreduced stored upper blocks
    // pins
    std::vector<int64_t> fixed_scalar;          // sorted unique scalar vars
    std::vector<int64_t> red_of_full;           // block: full -> reduced or -1
    // ordering: pos[reduced block] = factor position; iperm[pos] = reduced block
    std::vector<int32_t> pos, iperm;
    std::vector<int32_t> nd_perm;              // the fill-reducing ordering before the postorder
    SepForest nd_forest;                       // its separator forest (over reduced blocks)
    bool has_nd_forest = false;
    // supernodes (factor order)
    int32_t ns = 0;
    std::vector<int32_t> sfirst, swidth, sdeg, sparent, slevel;    // blocks
    std::vector<int64_t> srows_off;            // offset of struct rows in `rows`
    std::vector<int32_t> rows;                 // struct rows (block, factor order)
    std::vector<int64_t> panel_off;            // scalar offset of panel in L
    std::vector<int64_t> u_off;                // scalar offset of U in arena (-1 if d == 0)
    std::vector<int64_t> rel_off;              // offset of relative indices in `rel`
    std::vector<int32_t> rel;                  // relative positions in parent's row list
    std::vector<int32_t> children;             // child lists (CSR by parent)
    std::vector<int32_t> child_ptr;
    std::vector<int32_t> super_of;             // block column -> supernode
    int64_t L_size = 0;                        // doubles
    int64_t U_size = 0;                        // doubles (arena)
    int64_t inv_slots = 0;                     // 64x64 inverse scratch slots
    int64_t u_solve_size = 0;                  // doubles per RHS for solve update vectors
    std::vector<int64_t> u_solve_off;          // per supernode
    int64_t factor_nnz = 0;                    // scalar lower nnz incl. diag of the panels
    int32_t schur_tile = 128;                  // large tile of the Schur-update GEMM tasks
    double flops = 0.0;                        // POTRF+TRSM+SYRK over supernodes (scalar)
    double flops_unrelaxed = 0.0;              // sum_j c_j^2 of the unrelaxed column counts
    double flops_potrf = 0.0;                  // diagonal-tile POTRF share of `flops` (sum kb^3/3)
    int32_t num_levels = 0;
    int64_t max_width = 0, max_degree = 0;
    // injection plan: per reduced stored block (in reduced-CSC order)
    std::vector<int32_t> plan_src;             // original block entry index (Ax block)
    std::vector<int64_t> plan_dst;             // (scalar L offset << 2) | mode
    std::vector<int32_t> plan_row, plan_col;   // factor-order block row / column of the lower block
    std::vector<int32_t> diag_src;             // per factor position: Ax block index of its diagonal
    // schedule
    std::vector<LevelPlan> levels;
    std::vector<ExtendTask> extend_tasks;
    std::vector<PotrfTask> potrf_tasks;
    std::vector<PanelItem> panel_items;
    std::vector<GemmTask> gemm_tasks;
    std::vector<PullEnt> pull_ents;            // children's update-matrix shares of the Schur
                                               // tiles and of the panel tiles
    bool pulls_complete = true;                // every Schur tile lists its shares
    std::vector<PanelPullTask> panel_pull_tasks;
    std::vector<TileInj> tile_inj;             // injection entries of the panel tiles
    std::vector<FillTile> fill_tiles;          // with panel lists: the lower-part panel tiles
                                               // they do not rebuild (cleared up front)
    bool panel_pulls_complete = true;          // the panel tiles list every panel share
    std::vector<FusedTask> fused_tasks;
    // solves
    int32_t solve_block = 1024;                // column-block width (scalars, multiple of 64)
    std::vector<int32_t> tile_slot;            // per (supernode, 64-column tile): inverse slot
    std::vector<SolveBlock> sblocks;           // all column blocks (per supernode, in order)
    std::vector<SolveTile> inv_tiles;          // block-inverse work: (block, column tile)
    std::vector<SolveGatherTask> gather_tasks;
    std::vector<SolveTile> fdiag_tiles, bdiag_tiles;
    std::vector<SolveRowTask> fupd_tasks, bpart_tasks;
    int64_t xb_size = 0;                       // block-inverse storage (scalars)
    int64_t bpart_size = 0;                    // partial-sum slots (x solve_block scalars)
    // full-pattern SpMV (block-CSR over both triangles): for block row i, entries
    // (col block, Ax block index, transpose flag); diagonal flagged with 2.
    std::vector<int64_t> spmv_ptr;
    std::vector<int64_t> spmv_ent;             // (blockidx << 2) | mode, mode 0 as-is, 1 transpose, 2 diag
    std::vector<int32_t> spmv_col;
    // Scalar variables of a kept block that are pinned (partially pinned blocks) or padding (a
    // scalar matrix grouped into 3x3 blocks) are "masked": their rows and columns are replaced by
    // the identity at injection, so the factor restricted to the free scalars is the Cholesky
    // factor of the reduced matrix.  Internal vectors have 3 n_red entries in factor order; the
    // caller's vectors have n_full (full) or n_free (reduced, pins removed) entries.
    int64_t n_full = 0;                        // scalar length of full vectors (3 nb, or n of a scalar matrix)
    int64_t n_free = 0;                        // free scalars = reduced length
    std::vector<uint8_t> comp_mask;            // per factor position: bit c = component c masked (empty: none)
    int64_t factor_nnz_free = 0;               // nnz of the factor restricted to the free scalars
    // maps over the caller's index spaces (scalar)
    std::vector<int64_t> full_to_perm;         // full -> internal factor-order index, -1 if pinned/padding
    std::vector<int64_t> full_to_red;          // full -> reduced index (free scalars in order), -1 if pinned
    std::vector<int64_t> full_to_uperm;        // full -> factor-order reduced index (masked positions skipped)
    std::vector<int64_t> red_to_int;           // reduced -> internal factor-order index
    std::vector<int64_t> uperm_to_int;         // factor-order reduced -> internal index
    bool masked() const { return !comp_mask.empty(); }
};

struct SymbolicOptions {
    NDOptions nd;
    // relaxed amalgamation thresholds (scalar columns)
    int nrelax0 = 12, nrelax1 = 48, nrelax2 = 192;
    double zrelax0 = 0.8, zrelax1 = 0.2, zrelax2 = 0.05;
    int tile = 64;          // diagonal tile / inner step width (scalars)
    int outer = 512;        // outer panel block: trailing update with K = outer
    int big_tile = 128;     // GEMM regions at least this large in both dims use 128x128 tiles
    int split_big = 256;    // in-panel and next-block trailing updates with fewer 128x128 tiles use 64x64
    // Supernode size buckets by front size W + D (scalars): [0, 512), [512, 2048), [2048, inf).
    // bucket_tile[b] is the largest GEMM tile of the supernodes of bucket b: 128 (128x128 tiles
    // where the region is large enough; 256 for the bfloat16 Schur kernel) or 64 (64x64 tiles
    // only).  The autotuner picks them per bucket.
    int bucket_tile[3] = {128, 128, 128};
    int schur_tile = 128;   // large tile of the Schur update U = -L21 L21^T: 128, or 256 for the
                            // bfloat16 Schur kernel (regions at least 256 in both dims)
    int tile_group = 8;     // GEMM tiles are emitted in groups of this many tile columns
    int solve_block = 1024; // solve column-block width (multiple of 64, at most 1024)
    int solve_chunk = 1024; // backward partial-sum rows per task (multiple of 64, at most 1024)
    int fuse_degree = 0;    // supernodes of width <= one tile and at most this many structure rows
                            // (scalars) go to the fused level kernel; 0 disables it
    bool lookahead = true;  // split outer trailing updates: the next outer block's columns
                            // first, the rest beside the next outer block's steps
    int panel_max = 384;    // a step whose TRSM needs at most this many row tiles (and POTRF
                            // tasks) runs as one fused diagonal-tile + TRSM launch; 0 disables it
    bool pull_lists = true; // children lists in place of the extend-add passes: Schur-update
                            // tiles list the children's update-matrix blocks that land in them
                            // (added in the tile's epilogue), and
                            // 64 x 64 panel tiles those that land in the parent's panel
    int64_t n_scalar = -1;  // scalar length of the caller's matrix when it is not 3 nb (a scalar
                            // matrix grouped into blocks: scalars n_scalar .. 3 nb - 1 are padding)
    const std::vector<int32_t>* given_perm = nullptr;   // reuse this ordering (perm[new] = old)
    const SepForest* given_forest = nullptr;            // ... and its separator forest, if any
    NDState* nd_state = nullptr;                        // ordering history (temporal reuse)
};

// Options with the tuning knobs of the environment applied: MFS_OUTER (outer panel block),
// MFS_SOLVE_BLOCK (solve column block), MFS_SOLVE_CHUNK (backward rows per task), MFS_LEAF
// (nested-dissection leaf size; setting it makes one trial per component unless MFS_ND_TRIALS
// is also set), MFS_ND_TRIALS / MFS_ND_IMBALANCE / MFS_ND_COARSE / MFS_ND_SEED (dissection
// trials per component, bisection imbalance, coarsest graph size, random seed), MFS_ND_NSEPS
// (bisections tried per level), MFS_ND_NODEFM (node-based FM passes per separator), MFS_RELAX ("n0,n1,n2,z0,z1,z2": relaxed amalgamation
// thresholds), MFS_TILE_GROUP (GEMM tile-column group), MFS_FUSE_DEGREE
// (fused level kernel bound), MFS_BIG_TILE (regions at least this large in both dimensions use
// 128x128 tiles), MFS_PANEL_MAX (fused diagonal-tile + TRSM bound), MFS_LOOKAHEAD (0: no split
// of the outer trailing updates), MFS_PULL (0: extend-add passes instead of the Schur and panel
// tiles' children lists), MFS_BUCKET_TILES ("t0,t1,t2": largest GEMM tile, 64 or 128, per
// supernode size bucket); values are rounded down to multiples of 64 where required.
SymbolicOptions options_from_env();

// A scalar symmetric matrix stored as its upper triangle in CSC (the reference's SuiteSparseMatrix
// with UPPER_TRIANGLE symmetry, CholeskyFactorizerBase.hh:95-113) grouped into 3x3 blocks: scalar
// v belongs to block v / 3 (the last block is padded when n is not a multiple of 3); block (I, J)
// is stored when any scalar entry falls in it, and every diagonal block is stored.
// value_pos[k] = position of scalar entry k in the contiguous block values (9 per block).
struct ScalarBlocking {
    Pattern P;
    std::vector<int64_t> value_pos;
};
ScalarBlocking block_scalar_pattern(const int64_t* Ap, const int64_t* Ai, int64_t n);

// size bucket of a supernode front of W + D scalars (SymbolicOptions::bucket_tile)
inline int front_bucket(int64_t front) { return front < 512 ? 0 : (front < 2048 ? 1 : 2); }

Symbolic analyze(const int64_t* Ap, const int64_t* Ai, int64_t nblocks,
                 const int64_t* pins, int64_t npins, const SymbolicOptions& opt);
// the nested-dissection ordering analyze() computes (perm[new] = old over the reduced blocks)
std::vector<int32_t> nd_order(const Symbolic& S);

// Host replay of the device schedule (returns the first failing pivot column or -1) and the
// supernodal-panel -> CSC conversion.
int64_t emulate_numeric(const Symbolic& S, const double* Ax, double sigma, std::vector<double>& L,
                        std::vector<double>& inv);
void emulate_block_inverses(const Symbolic& S, const std::vector<double>& L, const std::vector<double>& inv,
                            std::vector<double>& Xb);
void emulate_solve(const Symbolic& S, const std::vector<double>& L, const std::vector<double>& Xb,
                   std::vector<double>& x);
void panels_to_csc(const Symbolic& S, const std::vector<double>& h, int64_t* Lp, int64_t* Li, double* Lx);

// BlockCSCHessian binary files (BlockCSCHessian.cc:11-103): contiguous 3x3 values in memory,
// the reference's non-contiguous layout on disk
void write_blockcsc(const char* path, const int64_t* Ap, const int64_t* Ai, int64_t nb, const double* Ax);
void read_blockcsc(const char* path, int64_t* nb, int64_t* nz, int64_t* snz, int64_t* Ap, int64_t* Ai, double* Ax);

// Workload (host)
void cube_mesh(int N, int32_t* tets, double* X);
void perturb_nodes(int N, uint64_t seed, double sigma_rel, double* X);
void neohookean_host(const double* Xr, const double* X, const int32_t* tets, int64_t ne,
                     const double* Ee, const double* nue, double E, double nu,
                     double* He, double* grad, double* energy, int nthreads);

} // namespace mfs
