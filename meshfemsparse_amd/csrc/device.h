// Device-side objects behind the C ABI (implemented in kernels.hip).
#pragma once
#include "mfs_internal.h"

#include <memory>

namespace mfs {

struct DeviceArray {
    void* p = nullptr;
    size_t bytes = 0;
    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray();
    void alloc(size_t b);
    void release();
    template <class T> void upload(const std::vector<T>& v) {
        alloc(std::max<size_t>(v.size() * sizeof(T), 16));
        upload_raw(v.data(), v.size() * sizeof(T));
    }
    void upload_raw(const void* src, size_t b);
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Block sparsity pattern built on the device from device element connectivity (same Ap/Ai as
// build_block_pattern); the arrays are left in Ap_d / Ai_d and returned on the host.
Pattern build_block_pattern_device(const int32_t* elem_nodes_dev, int64_t ne, int npe, int64_t nblocks,
                                   void* stream, DeviceArray& Ap_d, DeviceArray& Ai_d);

class AssemblerDev {
public:
    AssemblerDev(const Pattern& P, const int32_t* elem_nodes, int64_t ne, int npe);
    // contiguous: Ax holds 9 values per block; otherwise the non-contiguous (scalar upper CSC)
    // layout, which needs every diagonal block in the pattern
    void assemble(const double* He, int layout, double* Ax, void* stream, bool contiguous = true);
    void zero(double* Ax, void* stream);
    // neo-Hookean element Hessians (layout 0: 144, 1: 78 per element) and per-element gradients
    // (12 per element); any output may be null
    void neohookean(const double* Xr, const double* X, const double* Ee, const double* nue, double E, double nu,
                    double* He, int layout, double* g12, void* stream);
    // g += sum_e scatter(g12_e), fixed ascending element order per node
    void assemble_gradient(const double* g12, double* g, void* stream);
    int64_t nslots() const { return nslots_; }
private:
    void build_noncontiguous_map();
    int64_t nslots_ = 0, ne_ = 0, nnodes_ = 0;
    int npe_ = 0;
    bool diag_last_ = false;            // every column ends with its diagonal block
    std::vector<int64_t> ap_;
    DeviceArray seg_, contrib_, diag_, pslot_, tets_, nptr_, nent_;
    DeviceArray ncbase_, ncstride_;     // per slot, built on the first non-contiguous assembly
};

// Optional per-kernel-class device timing (HIP event pairs around each launch).
struct KernelProfile {
    enum { FILL_L, INJECT, UNUSED, EXTEND, POTRF, TRSM, UPDATE, SYRK, FGATHER, FDIAG, FGEMV, BGEMV,
           BDIAG, SPMV, MAPS, ASSEMBLE, BLKINV, BGATHER, CONVERT, REFINE, FUSED, PANEL, NCLASS };
    bool on = false;
    double ms[NCLASS] = {};
    int64_t count[NCLASS] = {};
    double work[NCLASS] = {};  // algorithmic flops of the profiled launches (GEMM classes)
    double bytes[NCLASS] = {}; // GEMM classes: operand bytes with no cache reuse across tiles
                               // (every tile reads its A and B panels and reads/writes its C)
    void begin(int cls, void* stream, double work = 0.0, int64_t items = 0, double bytes = 0.0);
    void end(void* stream);
    void collect();            // after the stream has been synchronized
    void reset();
    ~KernelProfile();
private:
    std::vector<void*> pool_;
    std::vector<std::pair<int, size_t>> marks_;
    std::vector<std::pair<double, int64_t>> work_;   // per mark: flops (or bytes), tasks
    std::vector<double> bytes_;                      // per mark: no-reuse operand bytes
};

class FactorizerDev {
public:
    explicit FactorizerDev(int precision);
    ~FactorizerDev();
    void symbolic(const int64_t* Ap, const int64_t* Ai, int64_t nb, int bs, const int64_t* pins,
                  int64_t npins);
    void numeric(const double* Ax, double sigma, const double* Bx, void* stream);
    // a scalar upper-triangular CSC matrix (SuiteSparseMatrix overloads): grouped into 3x3 blocks
    // with padding, then the block path; numeric / spmv / refinement then take the scalar values
    void symbolic_scalar(const int64_t* Ap, const int64_t* Ai, int64_t n, const int64_t* pins, int64_t npins);
    void numeric_scalar(const double* Ax, double sigma, const double* Bx, void* stream);
    bool scalar_mode() const { return scalar_mode_; }
    void solve(const double* b, double* x, int nrhs, void* stream);
    // the same solve with every scratch vector carved out of `base` (solve_scratch_bytes(nrhs)
    // bytes of device memory): allocates nothing
    void solve_in(const double* b, double* x, int nrhs, void* base, size_t bytes, void* stream);
    size_t solve_scratch_bytes(int nrhs) const;
    void solve_raw_reduced(const double* b, double* x, int nrhs, bool already_permuted, void* stream);
    void remove_fixed(const double* xf, double* xr, int nrhs, bool permute, void* stream);
    void extract_full(const double* xr, double* xf, int nrhs, bool permute, void* stream);
    void spmv(const double* Ax, const double* x, double* y, void* stream);
    void solve_refined(const double* Ax, const double* b, double* x, double tol, int max_iters,
                       int* iters, double* relres, void* stream);
    void export_csc(int64_t* Lp, int64_t* Li, double* Lx);
    // refinement until every segment [seg[s], seg[s+1]) of the full-length vectors meets tol
    void solve_refined_segments(const double* Ax, const double* b, double* x, int64_t nseg, const int64_t* seg,
                                double tol, int max_iters, int* iters, double* relres_max, double* relres_seg,
                                void* stream);
    // capture the numeric factorization in a hipGraph (replayed while Ax, Bx, sigma stay the same)
    void set_graph(bool on);
    bool graph_on() const { return graph_on_; }
    int64_t graph_nodes() const { return graph_nodes_; }
    // time the numeric factorization under candidate schedules (fused-kernel bound x big-tile
    // threshold) on the same ordering; keep the fastest.  ms[i * nt + j] for candidate (i, j).
    void autotune(const int64_t* Ap, const int64_t* Ai, int64_t nb, int bs, const int64_t* pins, int64_t npins,
                  const double* Ax, const int* fuse, int nfuse, const int* tiles, int ntile, int reps, double* ms,
                  int* best, void* stream);
    void clear_factors();
    // temporal reuse of the nested-dissection separator forest between symbolic analyses
    void set_reuse_period(int64_t period);
    int64_t reuse_period() const { return nd_state_.period; }
    void reset_reuse() { nd_state_.reset(); }
    const NDReuseStats& reuse_stats() const { return nd_state_.stats; }
    const Symbolic* symbolic_state() const { return ftype >= 1 ? sym.get() : nullptr; }
    void stash();
    void swap_stash();
    void clear_stash();
    bool has_stash() const { return stash_valid_; }

    KernelProfile prof;
    int ftype = 0;                 // 0 none, 1 symbolic, 2 numeric
    int64_t num_successful = 0;
    double timings[4] = {0, 0, 0, 0};
    std::unique_ptr<Symbolic> sym;
    int precision;

    void require(int t) const;
    // the Schur-update tiles add the children's update-matrix blocks (every tile of the analysis
    // carrying its list) instead of the extend-add passes into U
    bool pulls_active() const { return sym && sym->pulls_complete; }
    // the panel part of the extend-add through the panel tiles' children lists (every form)
    bool panel_pulls_active() const { return sym && sym->panel_pulls_complete; }

private:
    void install(std::unique_ptr<Symbolic> S);
    void drop_graph();
    template <typename T> void enqueue_numeric(const double* Ax, double sigma, const double* Bx, void* stream,
                                               void* ev_inject);
    // scratch of one solve (nrhs columns); null members: the factorizer's own buffers
    struct SolveScratch {
        double* xp = nullptr;   // factor-order right-hand sides (3 n_red x nrhs doubles)
        void* Y = nullptr;      // forward results y_b (3 n_red x nrhs, factor type)
        void* u = nullptr;      // update vectors (u_solve_size x nrhs)
        void* part = nullptr;   // backward partial sums (bpart_size x solve_block x nrhs)
        float* xf = nullptr;    // FP32 copy of xp (single-precision factors)
    };
    void solve_permuted_inplace(double* xp, int nrhs, void* stream, const SolveScratch* ws = nullptr);
    void solve_full_with(const double* b, double* x, int nrhs, double* xp, void* stream,
                         const SolveScratch* ws = nullptr);
    void scalar_values(const double* Ax, double* Axb, void* stream);   // scalar -> block values
    template <typename T> void numeric_t(const double* Ax, double sigma, const double* Bx, void* stream);
    template <typename T> void solve_t(T* xp, int nrhs, void* stream, const SolveScratch* ws);
    void ensure_vec(DeviceArray& a, size_t bytes);
    size_t fsize() const;

    DeviceArray L_, U_, inv_, status_;
    DeviceArray planent_, diagsrc_, scale_, xf_, rel_, ext_, potrf_, gemm_, fused_, uoff_, panel_, pcount_;
    DeviceArray pull_;                     // children's update-matrix shares of the Schur and panel tiles
    DeviceArray ppull_;                    // panel tiles with children shares
    DeviceArray tinj_;                     // their injection entries
    DeviceArray fillt_;                    // the panel tiles cleared before the injection
    DeviceArray gath_, sblk_, invt_, fdt_, bdt_, fupd_, bprt_, tslot_, Xb_;
    DeviceArray children_, usoff_, rows_, sdeg_, reloff_;
    DeviceArray spmv_ptr_, spmv_ent_, spmv_col_, f2p_, f2r_, r2p_, f2up_, up2i_, cmask_;
    const uint8_t* cmask() const { return cmask_.as<const uint8_t>(); }
    // scalar matrices: value positions in the block values, block copies of A and B, padded vectors
    bool scalar_mode_ = false;
    int64_t scalar_nnz_ = 0;
    DeviceArray s2b_, Axb_, Bxb_, xpad_, ypad_;
    DeviceArray usolve_, ysolve_, bpart_, scratch_b_, scratch_c_;
    DeviceArray Lstash_, invstash_, scalestash_, xbstash_;
    bool stash_valid_ = false;
    NDState nd_state_;              // ordering history (temporal reuse)
    std::vector<int64_t> nd_pins_;  // sorted pins of the history
    NDState* ordering_state(const int64_t* pins, int64_t npins);
    int schur_tile() const;
    bool graph_on_ = false;
    void* graph_ = nullptr;
    void* graph_exec_ = nullptr;
    void* cstream_ = nullptr;
    void* join_[2] = {nullptr, nullptr};
    void* side_ = nullptr;                 // lookahead stream (default priority)
    std::vector<void*> la_ev_;             // its fork/join events
    const double* gkey_ax_ = nullptr;
    const double* gkey_bx_ = nullptr;
    double gkey_sigma_ = 0.0;
    int64_t graph_nodes_ = 0;
    int fused_grid_ = 512;
    DeviceArray segp_, segsum_;
    int stash_type_ = 0;
    int64_t stash_succ_ = 0;
};

// device address of the allocation-free solve's scratch (MFS_SOLVE_NOALLOC_MAX_BYTES) on the
// current device, looked up on its first use there
void* noalloc_scratch();

// test hooks
template <typename T>
void test_gemm_nt(int M, int N, int K, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* C,
                  int ldc, void* stream);
void test_gemm_nt_bf16(int M, int N, int K, float alpha, const float* A, int lda, const float* B, int ldb,
                       float beta, float* C, int ldc, bool split, void* stream);
void test_multiply_lower_block(int BS, const int64_t* indices, int64_t start, int64_t size,
                               int64_t degree, const double* A, int64_t lda, int nrhs, double* B,
                               int64_t ldb, void* stream);

} // namespace mfs
