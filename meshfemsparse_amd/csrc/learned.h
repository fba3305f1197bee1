// Learned block incomplete-Cholesky preconditioner (no reference counterpart; the reference's
// preconditioning is exact Cholesky only).  A message-passing network over the free-node block
// graph of A emits a zero-fill lower block factor L on the pattern of tril(A) in a multicolor
// ordering; M = L L^T preconditions FP32 CG on the Jacobi-scaled system S A S (S = diag(A)^-1/2),
// with FP64 refinement outside and the supernodal Cholesky as fallback.  The network is trained
// on the device (own forward/backward kernels, Adam) with the loss mean_p |(L L^T - S A S) z_p|^2
// / |S A S z_p|^2 over random probes z_p.
#pragma once
#include "mfs_internal.h"

#include <memory>

namespace mfs {

struct LPGraph {
    int64_t nb_full = 0;
    int32_t n = 0;                          // free nodes (3x3 blocks)
    std::vector<int64_t> full_of;           // free node -> full block
    std::vector<int64_t> free_of_full;      // full block -> free node or -1
    int32_t ncolors = 0;
    std::vector<int32_t> color, pos;        // pos: index in the (color, node) order
    std::vector<int32_t> order;             // nodes in (color, node) order
    std::vector<int32_t> cptr;              // color c = order[cptr[c] .. cptr[c+1])
    // directed edges e = (dst <- src) between adjacent free nodes, CSR by dst, src ascending
    std::vector<int64_t> dptr;              // n + 1
    std::vector<int32_t> edst, esrc;
    std::vector<int64_t> eblk;              // (Ax block << 1) | 1 if A(dst, src) is the transpose
    std::vector<int64_t> diagblk;           // Ax block of each node's diagonal
    std::vector<int64_t> sptr;              // by source: edges sedge[sptr[j] .. sptr[j+1])
    std::vector<int32_t> sedge;
    // lower blocks L(i, j), pos[i] > pos[j]: q = 0..nl-1 in (i, j) order; by row i and column j
    std::vector<int32_t> ledge, ldst, lsrc; // the edge of block q, its row i and column j
    std::vector<int64_t> lrptr;             // row i: blocks q = lrptr[i] .. lrptr[i+1]
    std::vector<int64_t> lcptr;             // column j: blocks lcol[lcptr[j] .. lcptr[j+1]]
    std::vector<int32_t> lcol;
    // node -> full scalar offsets of the free rows (3 n entries)
    std::vector<int64_t> scal;
};

LPGraph build_lp_graph(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins, int64_t npins);

class FactorizerDev;

class LearnedDev {
public:
    static constexpr int H = 64, NL = 4, FN = 6, FE = 9;
    LearnedDev(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins, int64_t npins, uint64_t seed);
    ~LearnedDev();
    const LPGraph& graph() const { return g_; }
    int64_t nparams() const { return np_; }
    const std::vector<int64_t>& param_offsets() const { return poff_; }
    // features, scaling and the scaled matrix from the assembled values (9 nz doubles)
    void set_matrix(const double* Ax, void* stream);
    // forward + loss (+ backward into the gradient buffer) for probes z (host, nprobe x 3 n floats,
    // free-node order); returns the loss
    double loss_grad(const float* z, int nprobe, bool backward, void* stream);
    void adam_step(float lr, void* stream);
    void build_factor(void* stream);        // forward pass only: L for the preconditioner
    // preconditioned CG (FP32, scaled system) inside FP64 refinement on the full-length b, x;
    // precond 0: block Jacobi, 1: the network's factor.  Returns true when the true relative
    // residual reached tol within max_pcg total CG iterations.
    bool solve(const double* Ax, const double* b, double* x, int precond, double tol, int max_pcg, int* pcg_iters,
               int* outer, double* relres, void* stream);
    void get_params(float* h) const;
    void set_params(const float* h);
    void get_grads(float* h) const;
    void export_features(float* node_feat, float* edge_feat, float* diag, double* scale) const;
    void export_factor(float* LE, float* LD) const;
    int64_t steps() const { return t_; }
private:
    void forward(void* stream);
    void backward(void* stream);
    void dense(const float* X, int ldx, int64_t N, int K, const float* W, int ldw, bool wt, const float* b, float* Y,
               int ldy, int O, bool relu, void* stream);
    void wgrad(const float* X, int ldx, const float* Z, int ldz, int64_t N, int K, int O, float* gW, float* gb,
               void* stream);
    LPGraph g_;
    int64_t np_ = 0;
    std::vector<int64_t> poff_;
    int64_t t_ = 0;
    int nprobe_ = 0;
    struct Impl;
    std::unique_ptr<Impl> d_;
};

} // namespace mfs
