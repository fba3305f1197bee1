// extern "C" boundary: status codes + thread-local error message around the C++ objects.
#include "../../include/meshfemsparse_amd.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <memory>

#include "device.h"
#include "learned.h"

namespace mfs {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
} // namespace mfs

struct mfs_pattern_s {
    mfs::Pattern P;
    mfs::DeviceArray Ap_d, Ai_d;     // filled by mfs_pattern_build_device
};
struct mfs_assembler_s { std::unique_ptr<mfs::AssemblerDev> dev; };
struct mfs_factorizer_s { std::unique_ptr<mfs::FactorizerDev> dev; };
struct mfs_learned_s { std::unique_ptr<mfs::LearnedDev> dev; };

#define MFS_TRY try {
#define MFS_CATCH                                                                             \
    }                                                                                         \
    catch (const mfs::Error& e) {                                                             \
        mfs::set_last_error(e.what());                                                        \
        return e.code;                                                                        \
    }                                                                                         \
    catch (const std::bad_alloc&) {                                                           \
        mfs::set_last_error("host allocation failed");                                        \
        return MFS_ERR_ALLOC;                                                                 \
    }                                                                                         \
    catch (const std::invalid_argument& e) {                                                  \
        mfs::set_last_error(e.what());                                                        \
        return MFS_ERR_ARG;                                                                   \
    }                                                                                         \
    catch (const std::exception& e) {                                                         \
        mfs::set_last_error(std::string("internal failure: ") + e.what());                    \
        return MFS_ERR_STATE;                                                                 \
    }                                                                                         \
    return MFS_OK;

#define CHECK_PTR(p)                                                                          \
    if (!(p)) throw mfs::Error(MFS_ERR_ARG, std::string("null argument: ") + #p);

extern "C" {

const char* mfs_last_error(void) { return mfs::g_last_error.c_str(); }
int mfs_version(void) { return 100; }

int mfs_pattern_build(const int32_t* en, int64_t ne, int32_t npe, int64_t nb, mfs_pattern* out) {
    MFS_TRY
    CHECK_PTR(out);
    if (ne > 0) CHECK_PTR(en);
    if (nb <= 0 || ne < 0) throw mfs::Error(MFS_ERR_ARG, "bad pattern sizes");
    auto p = std::make_unique<mfs_pattern_s>();
    p->P = mfs::build_block_pattern(en, ne, npe, nb);
    *out = p.release();
    MFS_CATCH
}

int mfs_pattern_build_device(const int32_t* en_dev, int64_t ne, int32_t npe, int64_t nb, void* stream,
                             mfs_pattern* out) {
    MFS_TRY
    CHECK_PTR(out);
    auto h = std::make_unique<mfs_pattern_s>();
    h->P = mfs::build_block_pattern_device(en_dev, ne, npe, nb, stream, h->Ap_d, h->Ai_d);
    *out = h.release();
    MFS_CATCH
}

int mfs_pattern_get_device(mfs_pattern p, const int64_t** Ap_dev, const int64_t** Ai_dev) {
    MFS_TRY
    CHECK_PTR(p); CHECK_PTR(Ap_dev); CHECK_PTR(Ai_dev);
    if (!p->Ap_d.p) throw mfs::Error(-2, "pattern was built on the host (mfs_pattern_build)");
    *Ap_dev = p->Ap_d.as<const int64_t>();
    *Ai_dev = p->Ai_d.as<const int64_t>();
    MFS_CATCH
}

int mfs_pattern_get(mfs_pattern p, int64_t* nb, int64_t* nz, const int64_t** Ap, const int64_t** Ai) {
    MFS_TRY
    CHECK_PTR(p);
    if (nb) *nb = p->P.n;
    if (nz) *nz = p->P.nz();
    if (Ap) *Ap = p->P.Ap.data();
    if (Ai) *Ai = p->P.Ai.data();
    MFS_CATCH
}

int mfs_pattern_destroy(mfs_pattern p) {
    delete p;
    return MFS_OK;
}

int mfs_assembler_create(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int32_t* en,
                         int64_t ne, int32_t npe, mfs_assembler* out) {
    MFS_TRY
    CHECK_PTR(Ap); CHECK_PTR(Ai); CHECK_PTR(out);
    if (ne > 0) CHECK_PTR(en);
    if (nb <= 0 || ne < 0) throw mfs::Error(MFS_ERR_ARG, "bad assembler sizes");
    mfs::Pattern P;
    P.n = nb;
    P.Ap.assign(Ap, Ap + nb + 1);
    P.Ai.assign(Ai, Ai + Ap[nb]);
    for (int64_t j = 0; j < nb; ++j)
        for (int64_t k = P.Ap[j]; k < P.Ap[j + 1]; ++k)
            if (P.Ai[k] < 0 || P.Ai[k] > j || (k > P.Ap[j] && P.Ai[k] <= P.Ai[k - 1]))
                throw mfs::Error(MFS_ERR_ARG, "pattern must be sorted upper-triangular block CSC");
    auto a = std::make_unique<mfs_assembler_s>();
    a->dev = std::make_unique<mfs::AssemblerDev>(P, en, ne, npe);
    *out = a.release();
    MFS_CATCH
}

int mfs_assemble_hessian(mfs_assembler a, const double* He, int layout, double* Ax, void* stream) {
    MFS_TRY
    CHECK_PTR(a); CHECK_PTR(Ax);
    if (a->dev->nslots() > 0) CHECK_PTR(He);
    a->dev->assemble(He, layout, Ax, stream);
    MFS_CATCH
}

int mfs_assemble_hessian_noncontiguous(mfs_assembler a, const double* He, int layout, double* Ax, void* stream) {
    MFS_TRY
    CHECK_PTR(a); CHECK_PTR(Ax);
    if (a->dev->nslots() > 0) CHECK_PTR(He);
    a->dev->assemble(He, layout, Ax, stream, false);
    MFS_CATCH
}

int mfs_assembler_zero(mfs_assembler a, double* Ax, void* stream) {
    MFS_TRY
    CHECK_PTR(a); CHECK_PTR(Ax);
    a->dev->zero(Ax, stream);
    MFS_CATCH
}

int mfs_neohookean_elements(mfs_assembler a, const double* Xr, const double* X, const double* Ee,
                            const double* nue, double E, double nu, double* He, int32_t layout, double* g12,
                            void* stream) {
    MFS_TRY
    CHECK_PTR(a); CHECK_PTR(Xr); CHECK_PTR(X);
    a->dev->neohookean(Xr, X, Ee, nue, E, nu, He, layout, g12, stream);
    MFS_CATCH
}

int mfs_assemble_gradient(mfs_assembler a, const double* g12, double* g, void* stream) {
    MFS_TRY
    CHECK_PTR(a); CHECK_PTR(g12); CHECK_PTR(g);
    a->dev->assemble_gradient(g12, g, stream);
    MFS_CATCH
}

int mfs_assembler_destroy(mfs_assembler a) {
    delete a;
    return MFS_OK;
}

int mfs_factorizer_create(int precision, mfs_factorizer* out) {
    MFS_TRY
    CHECK_PTR(out);
    if (precision != MFS_PREC_FP64 && precision != MFS_PREC_FP32 && precision != MFS_PREC_BF16 &&
        precision != MFS_PREC_BF16X3)
        throw mfs::Error(MFS_ERR_ARG, "unknown precision");
    auto f = std::make_unique<mfs_factorizer_s>();
    f->dev = std::make_unique<mfs::FactorizerDev>(precision);
    *out = f.release();
    MFS_CATCH
}

int mfs_factorizer_destroy(mfs_factorizer f) {
    delete f;
    return MFS_OK;
}

int mfs_factorize_symbolic(mfs_factorizer f, const int64_t* Ap, const int64_t* Ai, int64_t nb,
                           int32_t bs, const int64_t* pins, int64_t npins) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ap); CHECK_PTR(Ai);
    if (npins > 0) CHECK_PTR(pins);
    f->dev->symbolic(Ap, Ai, nb, bs, pins, npins < 0 ? 0 : npins);
    MFS_CATCH
}

int mfs_factorize_numeric(mfs_factorizer f, const double* Ax, double sigma, const double* Bx, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ax);
    if (f->dev->ftype >= 1 && f->dev->scalar_mode())
        throw mfs::Error(MFS_ERR_ARG, "the symbolic factorization is of a scalar matrix: use mfs_factorize_numeric_scalar");
    f->dev->numeric(Ax, sigma, Bx, stream);
    MFS_CATCH
}

int mfs_factorize_symbolic_scalar(mfs_factorizer f, const int64_t* Ap, const int64_t* Ai, int64_t n,
                                  const int64_t* pins, int64_t npins) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ap); CHECK_PTR(Ai);
    if (npins > 0) CHECK_PTR(pins);
    f->dev->symbolic_scalar(Ap, Ai, n, pins, npins < 0 ? 0 : npins);
    MFS_CATCH
}

int mfs_factorize_numeric_scalar(mfs_factorizer f, const double* Ax, double sigma, const double* Bx, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ax);
    f->dev->numeric_scalar(Ax, sigma, Bx, stream);
    MFS_CATCH
}

int mfs_solve(mfs_factorizer f, const double* b, double* x, int32_t nrhs, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(b); CHECK_PTR(x);
    f->dev->solve(b, x, nrhs, stream);
    MFS_CATCH
}

int mfs_solve_noalloc(mfs_factorizer f, const double* b, double* x, int32_t nrhs, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(b); CHECK_PTR(x);
    if (f->dev->prof.on) throw mfs::Error(MFS_ERR_STATE, "the allocation-free solve needs kernel profiling off");
    f->dev->solve_in(b, x, nrhs, mfs::noalloc_scratch(), MFS_SOLVE_NOALLOC_MAX_BYTES, stream);
    MFS_CATCH
}

int mfs_solve_raw_reduced(mfs_factorizer f, const double* b, double* x, int32_t nrhs,
                          int32_t already_permuted, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(b); CHECK_PTR(x);
    f->dev->solve_raw_reduced(b, x, nrhs, already_permuted != 0, stream);
    MFS_CATCH
}

int mfs_remove_fixed_entries(mfs_factorizer f, const double* xf, double* xr, int32_t nrhs,
                             int32_t permute, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(xf); CHECK_PTR(xr);
    if (nrhs < 1) throw mfs::Error(MFS_ERR_ARG, "nrhs must be >= 1");
    f->dev->remove_fixed(xf, xr, nrhs, permute != 0, stream);
    MFS_CATCH
}

int mfs_extract_full_solution(mfs_factorizer f, const double* xr, double* xf, int32_t nrhs,
                              int32_t permute, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(xf); CHECK_PTR(xr);
    if (nrhs < 1) throw mfs::Error(MFS_ERR_ARG, "nrhs must be >= 1");
    f->dev->extract_full(xr, xf, nrhs, permute != 0, stream);
    MFS_CATCH
}

int mfs_spmv(mfs_factorizer f, const double* Ax, const double* x, double* y, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ax); CHECK_PTR(x); CHECK_PTR(y);
    f->dev->spmv(Ax, x, y, stream);
    MFS_CATCH
}

int mfs_solve_refined(mfs_factorizer f, const double* Ax, const double* b, double* x, double tol,
                      int32_t max_iters, int32_t* iters, double* relres, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ax); CHECK_PTR(b); CHECK_PTR(x);
    int it = 0;
    f->dev->solve_refined(Ax, b, x, tol, max_iters, &it, relres, stream);
    if (iters) *iters = it;
    MFS_CATCH
}

int mfs_solve_refined_segments(mfs_factorizer f, const double* Ax, const double* b, double* x, int64_t nseg,
                               const int64_t* seg_ptr, double tol, int32_t max_iters, int32_t* iters,
                               double* relres_max, double* relres_seg, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ax); CHECK_PTR(b); CHECK_PTR(x); CHECK_PTR(seg_ptr);
    int it = 0;
    f->dev->solve_refined_segments(Ax, b, x, nseg, seg_ptr, tol, max_iters, &it, relres_max, relres_seg, stream);
    if (iters) *iters = it;
    MFS_CATCH
}

int mfs_factorizer_set_graph(mfs_factorizer f, int32_t on) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->set_graph(on != 0);
    MFS_CATCH
}

int mfs_autotune(mfs_factorizer f, const int64_t* Ap, const int64_t* Ai, int64_t nb, int32_t bs, const int64_t* pins,
                 int64_t npins, const double* Ax, const int32_t* fuse, int32_t nfuse, const int32_t* tiles,
                 int32_t ntile, int32_t reps, double* ms, int32_t* best, void* stream) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Ap); CHECK_PTR(Ai); CHECK_PTR(Ax); CHECK_PTR(fuse); CHECK_PTR(tiles);
    if (npins > 0) CHECK_PTR(pins);
    std::vector<int> fz(fuse, fuse + std::max(nfuse, 0)), tz(tiles, tiles + std::max(ntile, 0));
    int b4[4] = {0, 0, 0, 0};
    f->dev->autotune(Ap, Ai, nb, bs, pins, npins, Ax, fz.data(), nfuse, tz.data(), ntile, reps, ms, b4, stream);
    if (best)
        for (int i = 0; i < 4; ++i) best[i] = b4[i];
    MFS_CATCH
}

int mfs_schedule_stats(mfs_factorizer f, int64_t* out) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(out);
    f->dev->require(1);
    const auto& S = *f->dev->sym;
    int64_t fused_launches = 0, potrf_launches = 0, gemm_launches = 0, extend_launches = 0, panel_launches = 0;
    for (const auto& P : S.levels) {
        fused_launches += P.fused.second > P.fused.first;
        for (size_t t = 0; t < P.potrf_steps.size(); ++t) {
            if (P.panel_steps[t].second > P.panel_steps[t].first) {
                ++panel_launches;
                continue;
            }
            potrf_launches += P.potrf_steps[t].second > P.potrf_steps[t].first;
            const auto& r = P.trsm_steps[t];
            gemm_launches += (r.m > r.b) + (r.e > r.m);
        }
        for (const auto& r : P.update_steps) gemm_launches += (r.m > r.b) + (r.e > r.m);
        gemm_launches += (P.syrk.m > P.syrk.b) + (P.syrk.e > P.syrk.m);
        extend_launches += (f->dev->panel_pulls_active() ? (P.panel_pulls.second > P.panel_pulls.first)
                                                          : (int64_t)P.extend_passes.size()) +
                           (f->dev->pulls_active() ? 0 : (int64_t)P.extend_u_passes.size());
    }
    out[0] = (int64_t)S.fused_tasks.size();
    out[1] = fused_launches;
    out[2] = S.ns - (int64_t)S.fused_tasks.size();
    out[3] = potrf_launches;
    out[4] = gemm_launches;
    out[5] = extend_launches;
    out[6] = S.num_levels;
    out[7] = f->dev->graph_nodes();
    out[8] = panel_launches;
    MFS_CATCH
}

int mfs_factorization_type(mfs_factorizer f, int32_t* t) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(t);
    *t = f->dev->ftype;
    MFS_CATCH
}

int mfs_num_successful_pivots(mfs_factorizer f, int64_t* n) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(n);
    *n = f->dev->num_successful;
    MFS_CATCH
}

int mfs_sizes(mfs_factorizer f, int64_t* nf, int64_t* nr) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->require(1);
    if (nf) *nf = f->dev->sym->n_full;
    if (nr) *nr = f->dev->sym->n_free;
    MFS_CATCH
}

int mfs_factor_nnz(mfs_factorizer f, int64_t* nnz) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(nnz);
    f->dev->require(1);
    *nnz = f->dev->sym->factor_nnz_free;
    MFS_CATCH
}

int mfs_flop_estimate(mfs_factorizer f, double* flops) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(flops);
    f->dev->require(1);
    *flops = f->dev->sym->flops;
    MFS_CATCH
}

int mfs_flop_breakdown(mfs_factorizer f, double* out3) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(out3);
    f->dev->require(1);
    const auto& S = *f->dev->sym;
    out3[0] = S.flops;
    out3[1] = S.flops_potrf;
    out3[2] = S.flops_unrelaxed;
    MFS_CATCH
}

int mfs_get_inverse_permutation(mfs_factorizer f, int64_t* ip) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(ip);
    f->dev->require(1);
    const auto& S = *f->dev->sym;
    // factor-order reduced index -> reduced index of the same scalar
    for (int64_t v = 0; v < S.n_full; ++v)
        if (S.full_to_uperm[v] >= 0) ip[S.full_to_uperm[v]] = S.full_to_red[v];
    MFS_CATCH
}

int mfs_clear_factors(mfs_factorizer f) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->clear_factors();
    MFS_CATCH
}

int mfs_stash_factorization(mfs_factorizer f) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->stash();
    MFS_CATCH
}

int mfs_has_stashed_factorization(mfs_factorizer f, int32_t* has) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(has);
    *has = f->dev->has_stash() ? 1 : 0;
    MFS_CATCH
}

int mfs_swap_stashed_factorization(mfs_factorizer f) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->swap_stash();
    MFS_CATCH
}

int mfs_clear_stashed_factorization(mfs_factorizer f) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->clear_stash();
    MFS_CATCH
}

int mfs_supernode_stats(mfs_factorizer f, int64_t* ns, int64_t* nl, int64_t* mw, int64_t* md) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->require(1);
    const auto& S = *f->dev->sym;
    if (ns) *ns = S.ns;
    if (nl) *nl = S.num_levels;
    if (mw) *mw = S.max_width;
    if (md) *md = S.max_degree;
    MFS_CATCH
}

int mfs_nd_ordering(mfs_factorizer f, int64_t* ncomp, int64_t* nvars, int64_t* block_size, int64_t* CParent,
                    int64_t* CMember) {
    MFS_TRY
    CHECK_PTR(f);
    const mfs::Symbolic* S = f->dev->symbolic_state();
    if (!S || !S->has_nd_forest) throw mfs::Error(MFS_ERR_STATE, "no nested-dissection ordering available");
    const mfs::SepForest& F = S->nd_forest;
    if (ncomp) *ncomp = (int64_t)F.parent.size();
    if (nvars) *nvars = (int64_t)F.member.size();
    if (block_size) *block_size = 3;
    if (CParent)
        for (size_t c = 0; c < F.parent.size(); ++c) CParent[c] = F.parent[c];
    if (CMember)
        for (size_t v = 0; v < F.member.size(); ++v) CMember[v] = F.member[v];
    MFS_CATCH
}

static void reuse_stats_out(const mfs::NDReuseStats& st, int64_t* s4, double* secs) {
    if (s4) {
        s4[0] = st.full_rebuild ? 1 : 0;
        s4[1] = st.graph_vertices;
        s4[2] = st.recomputed_subtrees;
        s4[3] = st.repartitioned_vertices;
    }
    if (secs) *secs = st.nd_seconds;
}

int mfs_set_temporal_reuse_period(mfs_factorizer f, int64_t period) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->set_reuse_period(period);
    MFS_CATCH
}

int mfs_get_temporal_reuse_period(mfs_factorizer f, int64_t* period) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(period);
    *period = f->dev->reuse_period();
    MFS_CATCH
}

int mfs_reset_temporal_reuse(mfs_factorizer f) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->reset_reuse();
    MFS_CATCH
}

int mfs_temporal_reuse_statistics(mfs_factorizer f, int64_t* stats4, double* nd_seconds) {
    MFS_TRY
    CHECK_PTR(f);
    reuse_stats_out(f->dev->reuse_stats(), stats4, nd_seconds);
    MFS_CATCH
}

struct mfs_nd_s {
    mfs::NDState state;
};

int mfs_nd_create(mfs_nd* out) {
    MFS_TRY
    CHECK_PTR(out);
    *out = new mfs_nd_s();
    MFS_CATCH
}

int mfs_nd_destroy(mfs_nd h) {
    delete h;
    return MFS_OK;
}

int mfs_nd_set_temporal_reuse_period(mfs_nd h, int64_t period) {
    MFS_TRY
    CHECK_PTR(h);
    if (period < 0) throw mfs::Error(MFS_ERR_ARG, "the temporal reuse period must be >= 0");
    if (period != h->state.period) h->state.reset();
    h->state.period = period;
    MFS_CATCH
}

int mfs_nd_reset(mfs_nd h) {
    MFS_TRY
    CHECK_PTR(h);
    h->state.reset();
    MFS_CATCH
}

int mfs_nd_order(mfs_nd h, int64_t n, const int64_t* p, const int64_t* i, int64_t* perm, int64_t* ncomp) {
    MFS_TRY
    CHECK_PTR(h);
    if (n < 0 || n > INT32_MAX) throw mfs::Error(MFS_ERR_ARG, "bad graph size");
    if (n > 0) { CHECK_PTR(p); CHECK_PTR(i); CHECK_PTR(perm); }
    std::vector<int64_t> xadj(n + 1, 0);
    std::vector<int32_t> adj;
    if (n > 0) {
        if (p[0] != 0) throw mfs::Error(MFS_ERR_ARG, "p[0] must be 0");
        adj.reserve(p[n]);
        for (int64_t v = 0; v < n; ++v) {
            if (p[v + 1] < p[v]) throw mfs::Error(MFS_ERR_ARG, "column pointers must not decrease");
            for (int64_t k = p[v]; k < p[v + 1]; ++k) {
                const int64_t u = i[k];
                if (u < 0 || u >= n) throw mfs::Error(MFS_ERR_ARG, "row index out of range");
                if (u != v) adj.push_back((int32_t)u);
            }
            xadj[v + 1] = (int64_t)adj.size();
        }
    }
    {
        // the graph must list every edge in both directions (a one-triangle pattern would leave
        // vertices without their neighbours): compare the edge set with its transpose
        std::vector<std::pair<int32_t, int32_t>> e, et;
        e.reserve(adj.size());
        et.reserve(adj.size());
        for (int64_t v = 0; v < n; ++v)
            for (int64_t k = xadj[v]; k < xadj[v + 1]; ++k) {
                e.emplace_back((int32_t)v, adj[k]);
                et.emplace_back(adj[k], (int32_t)v);
            }
        std::sort(e.begin(), e.end());
        e.erase(std::unique(e.begin(), e.end()), e.end());
        std::sort(et.begin(), et.end());
        et.erase(std::unique(et.begin(), et.end()), et.end());
        if (e != et) throw mfs::Error(MFS_ERR_ARG, "the graph must be symmetric (both triangles of the pattern)");
    }
    const mfs::NDOptions opt = mfs::options_from_env().nd;     // the factorizer's dissection knobs
    const std::vector<int32_t> pr = mfs::nested_dissection((int32_t)n, xadj, adj, opt, &h->state);
    for (int64_t k = 0; k < n; ++k) perm[k] = pr[k];
    if (ncomp) *ncomp = (int64_t)h->state.forest.parent.size();
    MFS_CATCH
}

int mfs_nd_forest(mfs_nd h, int64_t* CParent, int64_t* CMember) {
    MFS_TRY
    CHECK_PTR(h);
    const mfs::SepForest& F = h->state.forest;
    if (CParent)
        for (size_t c = 0; c < F.parent.size(); ++c) CParent[c] = F.parent[c];
    if (CMember)
        for (size_t v = 0; v < F.member.size(); ++v) CMember[v] = F.member[v];
    MFS_CATCH
}

int mfs_nd_statistics(mfs_nd h, int64_t* stats4, double* nd_seconds) {
    MFS_TRY
    CHECK_PTR(h);
    reuse_stats_out(h->state.stats, stats4, nd_seconds);
    MFS_CATCH
}

int mfs_analyze_host(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins,
                     int64_t npins, int64_t* iperm, int64_t* stats, double* flops) {
    MFS_TRY
    CHECK_PTR(Ap); CHECK_PTR(Ai);
    if (npins > 0) CHECK_PTR(pins);
    mfs::Symbolic S = mfs::analyze(Ap, Ai, nb, pins, npins < 0 ? 0 : npins, mfs::options_from_env());
    if (iperm)
        for (int64_t k = 0; k < S.n_red; ++k) iperm[k] = S.iperm[k];
    if (stats) {
        stats[0] = S.n_red; stats[1] = S.ns; stats[2] = S.num_levels; stats[3] = S.factor_nnz_free;
        stats[4] = S.max_width; stats[5] = S.max_degree; stats[6] = S.L_size; stats[7] = S.U_size;
    }
    if (flops) { flops[0] = S.flops; flops[1] = S.flops_unrelaxed; }
    MFS_CATCH
}

int mfs_debug_emulate_factor(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins,
                             int64_t npins, const double* Ax, double sigma, int64_t* Lp, int64_t* Li,
                             double* Lx, int64_t* failed_col, double* x_perm) {
    MFS_TRY
    CHECK_PTR(Ap); CHECK_PTR(Ai); CHECK_PTR(Ax);
    if (npins > 0) CHECK_PTR(pins);
    mfs::Symbolic S = mfs::analyze(Ap, Ai, nb, pins, npins < 0 ? 0 : npins, mfs::options_from_env());
    std::vector<double> L, inv;
    const int64_t fc = mfs::emulate_numeric(S, Ax, sigma, L, inv);
    if (failed_col) *failed_col = fc;
    if (Lp && Li && Lx) mfs::panels_to_csc(S, L, Lp, Li, Lx);
    if (x_perm && fc < 0) {
        // x_perm holds the free scalars in factor order; masked positions are zero internally
        std::vector<double> x(3 * S.n_red, 0.0), Xb;
        for (int64_t u = 0; u < S.n_free; ++u) x[S.uperm_to_int[u]] = x_perm[u];
        mfs::emulate_block_inverses(S, L, inv, Xb);
        mfs::emulate_solve(S, L, Xb, x);
        for (int64_t u = 0; u < S.n_free; ++u) x_perm[u] = x[S.uperm_to_int[u]];
    }
    if (failed_col && fc >= 0 && S.masked())
        *failed_col = std::lower_bound(S.uperm_to_int.begin(), S.uperm_to_int.end(), fc) - S.uperm_to_int.begin();
    MFS_CATCH
}

int mfs_factor_export_csc(mfs_factorizer f, int64_t* Lp, int64_t* Li, double* Lx) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(Lp); CHECK_PTR(Li); CHECK_PTR(Lx);
    f->dev->export_csc(Lp, Li, Lx);
    MFS_CATCH
}

int mfs_last_timings(mfs_factorizer f, double* ms4) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(ms4);
    for (int i = 0; i < 4; ++i) ms4[i] = f->dev->timings[i];
    MFS_CATCH
}

int mfs_kernel_profiling(mfs_factorizer f, int32_t enable, int32_t reset) {
    MFS_TRY
    CHECK_PTR(f);
    f->dev->prof.on = enable != 0;
    if (reset) f->dev->prof.reset();
    MFS_CATCH
}

int mfs_kernel_profile(mfs_factorizer f, double* ms, int64_t* counts) {
    MFS_TRY
    CHECK_PTR(f);
    auto& P = f->dev->prof;
    P.collect();                       // launches not yet folded in (plain solves), after they end
    for (int c = 0; c < mfs::KernelProfile::NCLASS; ++c) {
        if (ms) ms[c] = P.ms[c];
        if (counts) counts[c] = P.count[c];
    }
    MFS_CATCH
}

int mfs_kernel_profile_work(mfs_factorizer f, double* work) {
    MFS_TRY
    CHECK_PTR(f);
    CHECK_PTR(work);
    for (int c = 0; c < mfs::KernelProfile::NCLASS; ++c) work[c] = f->dev->prof.work[c];
    MFS_CATCH
}

int mfs_kernel_profile_bytes(mfs_factorizer f, double* bytes) {
    MFS_TRY
    CHECK_PTR(f); CHECK_PTR(bytes);
    for (int c = 0; c < mfs::KernelProfile::NCLASS; ++c) bytes[c] = f->dev->prof.bytes[c];
    MFS_CATCH
}

int mfs_cube_mesh(int32_t N, int32_t* tets, double* X) {
    MFS_TRY
    if (N <= 0 || N > 1000) throw mfs::Error(MFS_ERR_ARG, "N out of range");
    mfs::cube_mesh(N, tets, X);
    MFS_CATCH
}

int mfs_perturb_nodes(int32_t N, uint64_t seed, double sigma_rel, double* X) {
    MFS_TRY
    CHECK_PTR(X);
    if (N <= 0) throw mfs::Error(MFS_ERR_ARG, "N out of range");
    mfs::perturb_nodes(N, seed, sigma_rel, X);
    MFS_CATCH
}

int mfs_neohookean_host(const double* Xr, const double* X, const int32_t* tets, int64_t ne,
                        const double* Ee, const double* nue, double E, double nu, double* He,
                        double* grad, double* energy, int32_t nthreads) {
    MFS_TRY
    CHECK_PTR(Xr); CHECK_PTR(X); CHECK_PTR(tets);
    if (grad) {
        int64_t nn = 0;
        for (int64_t q = 0; q < 4 * ne; ++q) nn = std::max<int64_t>(nn, tets[q] + 1);
        std::memset(grad, 0, sizeof(double) * 3 * nn);
    }
    mfs::neohookean_host(Xr, X, tets, ne, Ee, nue, E, nu, He, grad, energy, nthreads);
    MFS_CATCH
}

int mfs_test_gemm_nt(int32_t M, int32_t N, int32_t K, double alpha, const double* A, int32_t lda,
                     const double* B, int32_t ldb, double beta, double* C, int32_t ldc, void* stream) {
    MFS_TRY
    mfs::test_gemm_nt(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream);
    MFS_CATCH
}

int mfs_test_gemm_nt_f32(int32_t M, int32_t N, int32_t K, float alpha, const float* A, int32_t lda,
                         const float* B, int32_t ldb, float beta, float* C, int32_t ldc, void* stream) {
    MFS_TRY
    mfs::test_gemm_nt(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, stream);
    MFS_CATCH
}

int mfs_test_gemm_nt_bf16(int32_t M, int32_t N, int32_t K, float alpha, const float* A, int32_t lda,
                          const float* B, int32_t ldb, float beta, float* C, int32_t ldc, int32_t split,
                          void* stream) {
    MFS_TRY
    mfs::test_gemm_nt_bf16(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, split != 0, stream);
    MFS_CATCH
}

int mfs_test_multiply_lower_block(int32_t BS, const int64_t* indices, int64_t start, int64_t size,
                                  int64_t degree, const double* A, int64_t lda, int32_t nrhs,
                                  double* B, int64_t ldb, void* stream) {
    MFS_TRY
    mfs::test_multiply_lower_block(BS, indices, start, size, degree, A, lda, nrhs, B, ldb, stream);
    MFS_CATCH
}


int mfs_learned_create(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins, int64_t npins,
                       uint64_t seed, mfs_learned* out) {
    MFS_TRY
    CHECK_PTR(Ap); CHECK_PTR(Ai); CHECK_PTR(out);
    if (npins > 0) CHECK_PTR(pins);
    auto l = std::make_unique<mfs_learned_s>();
    l->dev = std::make_unique<mfs::LearnedDev>(Ap, Ai, nb, pins, npins, seed);
    *out = l.release();
    MFS_CATCH
}

int mfs_learned_destroy(mfs_learned l) {
    delete l;
    return MFS_OK;
}

int mfs_learned_info(mfs_learned l, int64_t* out8) {
    MFS_TRY
    CHECK_PTR(l); CHECK_PTR(out8);
    const auto& g = l->dev->graph();
    out8[0] = g.n;
    out8[1] = (int64_t)g.esrc.size();
    out8[2] = (int64_t)g.ledge.size();
    out8[3] = g.ncolors;
    out8[4] = l->dev->nparams();
    out8[5] = l->dev->steps();
    out8[6] = mfs::LearnedDev::H;
    out8[7] = mfs::LearnedDev::NL;
    MFS_CATCH
}

int mfs_learned_param_offsets(mfs_learned l, int64_t* out) {
    MFS_TRY
    CHECK_PTR(l); CHECK_PTR(out);
    const auto& o = l->dev->param_offsets();
    std::copy(o.begin(), o.end(), out);
    MFS_CATCH
}

static void copy_lp_graph(const mfs::LPGraph& g, int32_t* edst, int32_t* esrc, int32_t* ledge, int32_t* ldst,
                          int32_t* lsrc, int32_t* color) {
    if (edst) std::copy(g.edst.begin(), g.edst.end(), edst);
    if (esrc) std::copy(g.esrc.begin(), g.esrc.end(), esrc);
    if (ledge) std::copy(g.ledge.begin(), g.ledge.end(), ledge);
    if (ldst) std::copy(g.ldst.begin(), g.ldst.end(), ldst);
    if (lsrc) std::copy(g.lsrc.begin(), g.lsrc.end(), lsrc);
    if (color) std::copy(g.color.begin(), g.color.end(), color);
}

int mfs_learned_graph(mfs_learned l, int32_t* edst, int32_t* esrc, int32_t* ledge, int32_t* ldst, int32_t* lsrc,
                      int32_t* color) {
    MFS_TRY
    CHECK_PTR(l);
    copy_lp_graph(l->dev->graph(), edst, esrc, ledge, ldst, lsrc, color);
    MFS_CATCH
}

int mfs_learned_analyze_host(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins, int64_t npins,
                             int64_t* sizes4, int32_t* edst, int32_t* esrc, int32_t* ledge, int32_t* ldst,
                             int32_t* lsrc, int32_t* color) {
    MFS_TRY
    CHECK_PTR(Ap); CHECK_PTR(Ai); CHECK_PTR(sizes4);
    if (npins > 0) CHECK_PTR(pins);
    const mfs::LPGraph g = mfs::build_lp_graph(Ap, Ai, nb, pins, npins);
    sizes4[0] = g.n;
    sizes4[1] = (int64_t)g.esrc.size();
    sizes4[2] = (int64_t)g.ledge.size();
    sizes4[3] = g.ncolors;
    copy_lp_graph(g, edst, esrc, ledge, ldst, lsrc, color);
    MFS_CATCH
}

int mfs_learned_set_matrix(mfs_learned l, const double* Ax, void* stream) {
    MFS_TRY
    CHECK_PTR(l); CHECK_PTR(Ax);
    l->dev->set_matrix(Ax, stream);
    MFS_CATCH
}

int mfs_learned_features(mfs_learned l, float* node_feat, float* edge_feat, float* diag, double* scale) {
    MFS_TRY
    CHECK_PTR(l);
    l->dev->export_features(node_feat, edge_feat, diag, scale);
    MFS_CATCH
}

int mfs_learned_loss(mfs_learned l, const float* z, int32_t nprobe, int32_t backward, double* loss, void* stream) {
    MFS_TRY
    CHECK_PTR(l); CHECK_PTR(z);
    const double v = l->dev->loss_grad(z, nprobe, backward != 0, stream);
    if (loss) *loss = v;
    MFS_CATCH
}

int mfs_learned_adam_step(mfs_learned l, float lr, void* stream) {
    MFS_TRY
    CHECK_PTR(l);
    l->dev->adam_step(lr, stream);
    MFS_CATCH
}

int mfs_learned_params(mfs_learned l, float* get, const float* set) {
    MFS_TRY
    CHECK_PTR(l);
    if (get) l->dev->get_params(get);
    if (set) l->dev->set_params(set);
    MFS_CATCH
}

int mfs_learned_grads(mfs_learned l, float* out) {
    MFS_TRY
    CHECK_PTR(l); CHECK_PTR(out);
    l->dev->get_grads(out);
    MFS_CATCH
}

int mfs_learned_factor(mfs_learned l, float* LE, float* LD, void* stream) {
    MFS_TRY
    CHECK_PTR(l);
    l->dev->build_factor(stream);
    l->dev->export_factor(LE, LD);
    MFS_CATCH
}

int mfs_learned_solve(mfs_learned l, mfs_factorizer fallback, const double* Ax, const double* b, double* x,
                      int32_t precond, double tol, int32_t max_pcg, int32_t* pcg_iters, int32_t* outer,
                      double* relres, int32_t* used_fallback, void* stream) {
    MFS_TRY
    CHECK_PTR(l); CHECK_PTR(Ax); CHECK_PTR(b); CHECK_PTR(x);
    int it = 0, ot = 0;
    double rr = 0;
    const bool ok = l->dev->solve(Ax, b, x, precond, tol, max_pcg, &it, &ot, &rr, stream);
    int fb = 0;
    if (!ok && fallback) {   // exact supernodal Cholesky (+ refinement) when CG ran out of iterations
        fallback->dev->numeric(Ax, 0.0, nullptr, stream);
        int fi = 0;
        fallback->dev->solve_refined(Ax, b, x, tol, 10, &fi, &rr, stream);
        fb = 1;
    }
    if (pcg_iters) *pcg_iters = it;
    if (outer) *outer = ot;
    if (relres) *relres = rr;
    if (used_fallback) *used_fallback = fb;
    if (!ok && !fallback) throw mfs::Error(MFS_ERR_NOT_CONVERGED, "preconditioned CG did not reach the tolerance");
    MFS_CATCH
}

int mfs_blockcsc_write(const char* path, const int64_t* Ap, const int64_t* Ai, int64_t nb, const double* Ax) {
    MFS_TRY
    CHECK_PTR(path); CHECK_PTR(Ap); CHECK_PTR(Ai);
    if (nb <= 0) throw mfs::Error(MFS_ERR_ARG, "bad block count");
    mfs::write_blockcsc(path, Ap, Ai, nb, Ax);
    MFS_CATCH
}

int mfs_blockcsc_read(const char* path, int64_t* nb, int64_t* nz, int64_t* snz, int64_t* Ap, int64_t* Ai, double* Ax) {
    MFS_TRY
    CHECK_PTR(path);
    mfs::read_blockcsc(path, nb, nz, snz, Ap, Ai, Ax);
    MFS_CATCH
}
} // extern "C"
