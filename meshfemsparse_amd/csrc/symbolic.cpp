// Host symbolic analysis, computed once per (pattern, pins) and reused across Newton steps.
//
// Mirrors the work CatamariFactorizer::m_factorizeSymbolic does (CatamariFactorizer.cc:177-616):
// whole-block pin removal (:187-246, SparseMatrices.hh:2020-2076), fill-reducing ordering
// (:290-325), symbolic factorization into relaxed supernodes (:516, :102-107 relax_supernodes),
// and the value-injection ("conversion") plan that maps each stored input block to its place in
// the supernodal factor (CatamariConversionPlan.hh:402-541).  Everything is in 3x3 block units:
// all scalar columns of a block share one row structure, so the block analysis is exact for
// uniform block size 3 (the reference's IndexBlockSize compression, CatamariConversionPlan.hh:101-141).
//
// The numeric phase is a multifrontal supernodal LL^T: supernode s owns a column panel
// [L11; L21] of (W+D) x W scalars (W = 3 width, D = 3 degree) and an update matrix U_s (D x D)
// that receives its children's update matrices (extend-add) and then -L21 L21^T.  The level
// schedule (supernode heights in the assembly tree) and every dense tile task are built here,
// so the device side only replays fixed task lists.
#include <cstdio>
#include <cstdlib>
#include "mfs_internal.h"

#include <algorithm>
#include <map>
#include <numeric>

namespace mfs {

namespace {

void etree(int32_t n, const std::vector<int64_t>& Up, const std::vector<int32_t>& Ui,
           std::vector<int32_t>& parent) {
    // Liu's algorithm with path compression; column k of the upper pattern holds rows i < k.
    parent.assign(n, -1);
    std::vector<int32_t> anc(n, -1);
    for (int32_t k = 0; k < n; ++k)
        for (int64_t p = Up[k]; p < Up[k + 1]; ++p) {
            int32_t r = Ui[p];
            while (anc[r] != -1 && anc[r] != k) {
                const int32_t nx = anc[r];
                anc[r] = k;
                r = nx;
            }
            if (anc[r] == -1) { anc[r] = k; parent[r] = k; }
        }
}

std::vector<int32_t> postorder(int32_t n, const std::vector<int32_t>& parent) {
    std::vector<int32_t> head(n, -1), next(n, -1), post;
    post.reserve(n);
    for (int32_t j = n - 1; j >= 0; --j)
        if (parent[j] != -1) { next[j] = head[parent[j]]; head[parent[j]] = j; }
    std::vector<int32_t> stack;
    for (int32_t j = 0; j < n; ++j) {
        if (parent[j] != -1) continue;
        stack.push_back(j);
        while (!stack.empty()) {
            const int32_t p = stack.back();
            const int32_t c = head[p];
            if (c == -1) { stack.pop_back(); post.push_back(p); }
            else { head[p] = next[c]; stack.push_back(c); }
        }
    }
    return post;
}

// Column counts of L (incl. diagonal) for a postordered etree (parent[j] > j), following the
// row-subtree / least-common-ancestor characterisation (Gilbert, Ng, Peyton).
std::vector<int64_t> column_counts(int32_t n, const std::vector<int32_t>& parent,
                                   const std::vector<int64_t>& Lp, const std::vector<int32_t>& Li) {
    std::vector<int64_t> delta(n, 0);
    std::vector<int32_t> first(n, -1), maxfirst(n, -1), prevleaf(n, -1), ancestor(n);
    for (int32_t k = 0; k < n; ++k) {
        int32_t j = k;
        delta[j] = (first[j] == -1) ? 1 : 0;
        for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
    }
    std::iota(ancestor.begin(), ancestor.end(), 0);
    for (int32_t j = 0; j < n; ++j) {
        if (parent[j] != -1) delta[parent[j]]--;
        for (int64_t p = Lp[j]; p < Lp[j + 1]; ++p) {
            const int32_t i = Li[p];            // i > j
            if (first[j] <= maxfirst[i]) continue;
            maxfirst[i] = first[j];
            const int32_t jprev = prevleaf[i];
            prevleaf[i] = j;
            delta[j]++;
            if (jprev == -1) continue;           // first leaf of row subtree i
            int32_t q = jprev;
            while (q != ancestor[q]) q = ancestor[q];
            for (int32_t s = jprev; s != q;) { const int32_t sp = ancestor[s]; ancestor[s] = q; s = sp; }
            delta[q]--;                          // q = lca(jprev, j)
        }
        if (parent[j] != -1) ancestor[j] = parent[j];
    }
    for (int32_t j = 0; j < n; ++j)
        if (parent[j] != -1) delta[parent[j]] += delta[j];
    return delta;
}

// First-fit allocator over a growing arena (U buffers of the multifrontal schedule).
struct Arena {
    std::map<int64_t, int64_t> free_;   // offset -> size
    int64_t top = 0, peak = 0;
    int64_t alloc(int64_t sz) {
        sz = (sz + 31) & ~int64_t(31);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= sz) {
                const int64_t off = it->first, rest = it->second - sz;
                free_.erase(it);
                if (rest > 0) free_[off + sz] = rest;
                return off;
            }
        const int64_t off = top;
        top += sz;
        peak = std::max(peak, top);
        return off;
    }
    void release(int64_t off, int64_t sz) {
        sz = (sz + 31) & ~int64_t(31);
        auto it = free_.emplace(off, sz).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
                it = pv;
            }
        }
        if (it->first + it->second == top) { top = it->first; free_.erase(it); }
    }
};

} // namespace

double order_flops(const std::vector<int64_t>& xadj, const std::vector<int32_t>& adj,
                   const std::vector<int32_t>& order, std::vector<int32_t>& pos) {
    const int32_t m = (int32_t)order.size();
    for (int32_t k = 0; k < m; ++k) pos[order[k]] = k;
    // upper pattern (column k: rows i < k), then the etree postorder
    std::vector<int64_t> Up(m + 1, 0);
    std::vector<int32_t> Ui, parent;
    for (int32_t k = 0; k < m; ++k) {
        const int32_t v = order[k];
        for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
            const int32_t i = pos[adj[e]];
            if (i >= 0 && i < k) Ui.push_back(i);
        }
        Up[k + 1] = (int64_t)Ui.size();
    }
    etree(m, Up, Ui, parent);
    const std::vector<int32_t> post = postorder(m, parent);
    for (int32_t k = 0; k < m; ++k) pos[order[post[k]]] = k;
    std::vector<int64_t> Lp(m + 1, 0);
    std::vector<int32_t> Li;
    Up.assign(m + 1, 0);
    Ui.clear();
    for (int32_t k = 0; k < m; ++k) {             // column k of the postordered matrix
        const int32_t v = order[post[k]];
        for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
            const int32_t i = pos[adj[e]];
            if (i >= 0 && i < k) Ui.push_back(i);
        }
        Up[k + 1] = (int64_t)Ui.size();
    }
    for (int32_t k = 0; k < m; ++k)
        for (int64_t p = Up[k]; p < Up[k + 1]; ++p) Lp[Ui[p] + 1]++;
    for (int32_t k = 0; k < m; ++k) Lp[k + 1] += Lp[k];
    Li.resize(Lp[m]);
    {
        std::vector<int64_t> b(Lp.begin(), Lp.end() - 1);
        for (int32_t k = 0; k < m; ++k)
            for (int64_t p = Up[k]; p < Up[k + 1]; ++p) Li[b[Ui[p]]++] = k;
    }
    etree(m, Up, Ui, parent);
    const std::vector<int64_t> cnt = column_counts(m, parent, Lp, Li);
    for (int32_t v : order) pos[v] = -1;
    double f = 0;
    for (int32_t j = 0; j < m; ++j) {
        const double c = 3.0 * cnt[j];
        f += c * c + (c - 1) * (c - 1) + (c - 2) * (c - 2);
    }
    return f;
}

std::vector<int32_t> nd_order(const Symbolic& S) { return S.nd_perm; }

SymbolicOptions options_from_env() {
    SymbolicOptions opt;
    if (const char* e = getenv("MFS_OUTER")) opt.outer = std::max(64, atoi(e) / 64 * 64);
    if (const char* e = getenv("MFS_SOLVE_BLOCK")) opt.solve_block = std::min(1024, std::max(64, atoi(e) / 64 * 64));
    if (const char* e = getenv("MFS_SOLVE_CHUNK")) opt.solve_chunk = std::min(1024, std::max(64, atoi(e) / 64 * 64));
    if (const char* e = getenv("MFS_LEAF")) { opt.nd.leaf_size = std::max(8, atoi(e)); opt.nd.trials = 1; }
    if (const char* e = getenv("MFS_ND_TRIALS")) opt.nd.trials = std::min(64, std::max(1, atoi(e)));
    if (const char* e = getenv("MFS_ND_NSEPS")) opt.nd.nseps = std::min(16, std::max(1, atoi(e)));
    if (const char* e = getenv("MFS_ND_NODEFM")) opt.nd.node_passes = std::min(16, std::max(0, atoi(e)));
    if (const char* e = getenv("MFS_RELAX")) {
        int n0, n1, n2;
        double z0, z1, z2;
        if (sscanf(e, "%d,%d,%d,%lf,%lf,%lf", &n0, &n1, &n2, &z0, &z1, &z2) == 6) {
            opt.nrelax0 = n0; opt.nrelax1 = n1; opt.nrelax2 = n2;
            opt.zrelax0 = z0; opt.zrelax1 = z1; opt.zrelax2 = z2;
        }
    }
    if (const char* e = getenv("MFS_ND_IMBALANCE")) opt.nd.imbalance = std::min(0.5, std::max(0.0, atof(e)));
    if (const char* e = getenv("MFS_ND_COARSE")) opt.nd.coarse_size = std::max(16, atoi(e));
    if (const char* e = getenv("MFS_ND_SEED")) opt.nd.seed = strtoull(e, nullptr, 0);
    if (const char* e = getenv("MFS_TILE_GROUP")) opt.tile_group = std::max(1, atoi(e));
    if (const char* e = getenv("MFS_FUSE_DEGREE")) opt.fuse_degree = std::max(0, atoi(e));
    if (const char* e = getenv("MFS_BIG_TILE")) opt.big_tile = std::max(128, atoi(e));
    if (const char* e = getenv("MFS_SPLIT_BIG")) opt.split_big = std::max(0, atoi(e));
    if (const char* e = getenv("MFS_PANEL_MAX")) opt.panel_max = std::max(0, atoi(e));
    if (const char* e = getenv("MFS_LOOKAHEAD")) opt.lookahead = atoi(e) != 0;
    if (const char* e = getenv("MFS_PULL")) opt.pull_lists = atoi(e) != 0;
    if (const char* e = getenv("MFS_BUCKET_TILES")) {
        int t0, t1, t2;
        if (sscanf(e, "%d,%d,%d", &t0, &t1, &t2) == 3) {
            opt.bucket_tile[0] = t0 > 64 ? 128 : 64;
            opt.bucket_tile[1] = t1 > 64 ? 128 : 64;
            opt.bucket_tile[2] = t2 > 64 ? 128 : 64;
        }
    }
    return opt;
}

ScalarBlocking block_scalar_pattern(const int64_t* Ap, const int64_t* Ai, int64_t n) {
    if (n <= 0) throw Error(-1, "empty matrix");
    if (n >= (int64_t(1) << 31)) throw Error(-6, "matrix too large");
    if (Ap[0] != 0) throw Error(-1, "column pointers must start at 0");
    for (int64_t j = 0; j < n; ++j) {
        if (Ap[j + 1] < Ap[j]) throw Error(-1, "column pointers must not decrease");
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            if (Ai[k] < 0 || Ai[k] > j)
                throw Error(-1, "the scalar matrix must hold its upper triangle (row <= column)");
            if (k > Ap[j] && Ai[k] <= Ai[k - 1]) throw Error(-1, "row indices must be sorted and unique");
        }
    }
    ScalarBlocking B;
    const int64_t nb = (n + 2) / 3;
    B.P.n = nb;
    B.P.Ap.assign(nb + 1, 0);
    std::vector<int64_t> rows;
    for (int64_t J = 0; J < nb; ++J) {
        rows.clear();
        for (int64_t j = 3 * J; j < std::min(n, 3 * J + 3); ++j)
            for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) rows.push_back(Ai[k] / 3);
        rows.push_back(J);
        std::sort(rows.begin(), rows.end());
        rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
        B.P.Ai.insert(B.P.Ai.end(), rows.begin(), rows.end());
        B.P.Ap[J + 1] = (int64_t)B.P.Ai.size();
    }
    B.value_pos.resize(Ap[n]);
    for (int64_t j = 0; j < n; ++j) {
        const int64_t J = j / 3;
        const int64_t* b0 = B.P.Ai.data() + B.P.Ap[J];
        const int64_t* b1 = B.P.Ai.data() + B.P.Ap[J + 1];
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            const int64_t blk = (std::lower_bound(b0, b1, Ai[k] / 3) - B.P.Ai.data());
            B.value_pos[k] = 9 * blk + Ai[k] % 3 + 3 * (j % 3);
        }
    }
    return B;
}

Symbolic analyze(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins,
                 int64_t npins, const SymbolicOptions& opt) {
    Symbolic S;
    if (nb <= 0) throw Error(-1, "empty matrix");
    if (nb >= (int64_t(1) << 30)) throw Error(-6, "matrix too large");
    S.n_full_blocks = nb;
    S.nz_full = Ap[nb];
    for (int64_t j = 0; j < nb; ++j) {
        if (Ap[j + 1] <= Ap[j]) throw Error(-1, "column " + std::to_string(j) + " has no diagonal block");
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            if (Ai[k] < 0 || Ai[k] > j) throw Error(-1, "block pattern must be upper triangular");
            if (k > Ap[j] && Ai[k] <= Ai[k - 1]) throw Error(-1, "block rows must be sorted and unique");
        }
        if (Ai[Ap[j + 1] - 1] != j) throw Error(-1, "diagonal block must be last in its column");
    }

    // ---- pins (CatamariFactorizer.cc:187-246): whole-block pins remove the block's row and
    // column; a partially pinned block stays and its pinned components are masked (the
    // reference falls back to a scalar factorization there, :207-216, with the same solution).
    // Padding scalars of a scalar matrix grouped into blocks are masked the same way. ----
    const int64_t nsf = opt.n_scalar < 0 ? 3 * nb : opt.n_scalar;
    if (nsf > 3 * nb || nsf <= 3 * (nb - 1)) throw Error(-1, "scalar size does not match the block count");
    S.n_full = nsf;
    std::vector<uint8_t> bmask(nb, 0);       // per full block: bit c = component c pinned or padding
    for (int64_t v = nsf; v < 3 * nb; ++v) bmask[v / 3] |= (uint8_t)(1u << (v % 3));
    {
        std::vector<char> seen((size_t)nsf, 0);
        for (int64_t q = 0; q < npins; ++q) {
            const int64_t v = pins[q];
            if (v < 0 || v >= nsf) throw Error(-1, "Fixed variable index out of range");
            if (seen[v]) continue;
            seen[v] = 1;
            S.fixed_scalar.push_back(v);
            bmask[v / 3] |= (uint8_t)(1u << (v % 3));
        }
        std::sort(S.fixed_scalar.begin(), S.fixed_scalar.end());
    }
    S.n_free = nsf - (int64_t)S.fixed_scalar.size();
    S.red_of_full.assign(nb, -1);
    int64_t nr = 0;
    bool any_mask = false;
    for (int64_t b = 0; b < nb; ++b)
        if (bmask[b] != 7) {
            S.red_of_full[b] = nr++;
            any_mask |= bmask[b] != 0;
        }
    S.n_red = nr;
    if (nr == 0) throw Error(-1, "every variable is pinned");
    const int32_t n = (int32_t)nr;

    // reduced entries (reduced-CSC order) and the adjacency graph
    struct RE { int32_t i, j; int32_t src; };
    std::vector<RE> red;
    red.reserve(S.nz_full);
    for (int64_t j = 0; j < nb; ++j) {
        if (S.red_of_full[j] < 0) continue;
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            const int64_t i = Ai[k];
            if (S.red_of_full[i] < 0) continue;
            red.push_back({(int32_t)S.red_of_full[i], (int32_t)S.red_of_full[j], (int32_t)k});
        }
    }
    std::vector<int64_t> xadj(n + 1, 0);
    for (const RE& e : red)
        if (e.i != e.j) { xadj[e.i + 1]++; xadj[e.j + 1]++; }
    for (int32_t v = 0; v < n; ++v) xadj[v + 1] += xadj[v];
    std::vector<int32_t> adj(xadj[n]);
    {
        std::vector<int64_t> bk(xadj.begin(), xadj.end() - 1);
        for (const RE& e : red)
            if (e.i != e.j) { adj[bk[e.i]++] = e.j; adj[bk[e.j]++] = e.i; }
    }

    auto build_patterns = [&](const std::vector<int32_t>& pos_, std::vector<int64_t>& Lp,
                              std::vector<int32_t>& Li, std::vector<int64_t>& Up,
                              std::vector<int32_t>& Ui) {
        Lp.assign(n + 1, 0);
        Up.assign(n + 1, 0);
        for (const RE& e : red) {
            if (e.i == e.j) continue;
            const int32_t a = pos_[e.i], b = pos_[e.j];
            Lp[std::min(a, b) + 1]++;
            Up[std::max(a, b) + 1]++;
        }
        for (int32_t v = 0; v < n; ++v) { Lp[v + 1] += Lp[v]; Up[v + 1] += Up[v]; }
        Li.resize(Lp[n]);
        Ui.resize(Up[n]);
        std::vector<int64_t> bl(Lp.begin(), Lp.end() - 1), bu(Up.begin(), Up.end() - 1);
        for (const RE& e : red) {
            if (e.i == e.j) continue;
            const int32_t a = pos_[e.i], b = pos_[e.j];
            Li[bl[std::min(a, b)]++] = std::max(a, b);
            Ui[bu[std::max(a, b)]++] = std::min(a, b);
        }
        for (int32_t v = 0; v < n; ++v) {
            std::sort(Li.begin() + Lp[v], Li.begin() + Lp[v + 1]);
            std::sort(Ui.begin() + Up[v], Ui.begin() + Up[v + 1]);
        }
    };
    // ---- ordering: ND, then etree postorder ----
    std::vector<int32_t> perm;   // perm[new] = old
    if (opt.given_perm) {
        if ((int32_t)opt.given_perm->size() != n) throw Error(-1, "given ordering has the wrong size");
        perm = *opt.given_perm;
        if (opt.given_forest && opt.given_forest->member.size() == (size_t)n) {
            S.nd_forest = *opt.given_forest;
            S.has_nd_forest = true;
        }
    } else {
        NDState local;
        NDState* st = opt.nd_state ? opt.nd_state : &local;
        perm = nested_dissection(n, xadj, adj, opt.nd, st);
        S.nd_forest = st->forest;
        S.has_nd_forest = true;
    }
    S.nd_perm = perm;
    std::vector<int32_t> posv(n);
    for (int32_t k = 0; k < n; ++k) posv[perm[k]] = k;
    std::vector<int64_t> Lp, Up;
    std::vector<int32_t> Li, Ui, parent;
    build_patterns(posv, Lp, Li, Up, Ui);
    etree(n, Up, Ui, parent);
    {
        std::vector<int32_t> post = postorder(n, parent);
        S.iperm.resize(n);
        for (int32_t k = 0; k < n; ++k) S.iperm[k] = perm[post[k]];
        S.pos.resize(n);
        for (int32_t k = 0; k < n; ++k) S.pos[S.iperm[k]] = k;
    }
    build_patterns(S.pos, Lp, Li, Up, Ui);
    etree(n, Up, Ui, parent);
    for (int32_t j = 0; j < n; ++j)
        if (parent[j] != -1 && parent[j] <= j) throw Error(-2, "internal: etree not postordered");
    std::vector<int64_t> cc = column_counts(n, parent, Lp, Li);
    for (int32_t j = 0; j < n; ++j) {
        const double c = 3.0 * cc[j];
        S.flops_unrelaxed += c * c + (c - 1) * (c - 1) + (c - 2) * (c - 2);
    }

    // ---- fundamental supernodes ----
    std::vector<int32_t> nchild(n, 0);
    for (int32_t j = 0; j < n; ++j) if (parent[j] != -1) nchild[parent[j]]++;
    struct SN { int32_t f, w; int64_t d; double nz; };
    std::vector<SN> fund;
    for (int32_t j = 0; j < n; ++j) {
        const bool cont = j > 0 && parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1;
        if (cont) fund.back().w++;
        else fund.push_back({j, 1, 0, 0.0});
    }
    for (SN& s : fund) {
        s.d = cc[s.f] - s.w;
        s.nz = 0.5 * s.w * (s.w + 1.0) + (double)s.w * s.d;
    }

    // ---- relaxed amalgamation (contiguous last child into parent) ----
    std::vector<SN> merged;
    for (const SN& s0 : fund) {
        SN cur = s0;
        while (!merged.empty()) {
            const SN& c = merged.back();
            const int32_t clast = c.f + c.w - 1;
            if (clast != cur.f - 1) break;
            const int32_t pcol = parent[clast];
            if (pcol < cur.f || pcol >= cur.f + cur.w) break;
            const int64_t w = (int64_t)c.w + cur.w;
            const double tot = 0.5 * w * (w + 1.0) + (double)w * cur.d;
            const double z = (tot - (c.nz + cur.nz)) / tot;
            const int64_t ncol = 3 * w;
            const bool ok = ncol <= opt.nrelax0 || (ncol <= opt.nrelax1 && z < opt.zrelax0) ||
                            (ncol <= opt.nrelax2 && z < opt.zrelax1) || z < opt.zrelax2;
            if (!ok) break;
            cur.f = c.f;
            cur.w = (int32_t)w;
            cur.nz = c.nz + cur.nz;
            merged.pop_back();
        }
        merged.push_back(cur);
    }
    const int32_t ns = (int32_t)merged.size();
    S.ns = ns;
    S.sfirst.resize(ns);
    S.swidth.resize(ns);
    S.super_of.resize(n);
    for (int32_t s = 0; s < ns; ++s) {
        S.sfirst[s] = merged[s].f;
        S.swidth[s] = merged[s].w;
        for (int32_t j = merged[s].f; j < merged[s].f + merged[s].w; ++j) S.super_of[j] = s;
    }
    S.sparent.assign(ns, -1);
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t last = S.sfirst[s] + S.swidth[s] - 1;
        if (parent[last] != -1) S.sparent[s] = S.super_of[parent[last]];
    }
    // children lists (ascending supernode id)
    S.child_ptr.assign(ns + 1, 0);
    for (int32_t s = 0; s < ns; ++s) if (S.sparent[s] >= 0) S.child_ptr[S.sparent[s] + 1]++;
    for (int32_t s = 0; s < ns; ++s) S.child_ptr[s + 1] += S.child_ptr[s];
    S.children.resize(S.child_ptr[ns]);
    {
        std::vector<int32_t> bk(S.child_ptr.begin(), S.child_ptr.end() - 1);
        for (int32_t s = 0; s < ns; ++s) if (S.sparent[s] >= 0) S.children[bk[S.sparent[s]]++] = s;
    }

    // ---- supernode row structures ----
    S.sdeg.assign(ns, 0);
    S.srows_off.assign(ns + 1, 0);
    {
        std::vector<int32_t> mark(n, -1), tmp;
        std::vector<std::vector<int32_t>> st(ns);
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t f = S.sfirst[s], last = f + S.swidth[s] - 1;
            tmp.clear();
            for (int32_t j = f; j <= last; ++j)
                for (int64_t p = Lp[j]; p < Lp[j + 1]; ++p) {
                    const int32_t i = Li[p];
                    if (i > last && mark[i] != s) { mark[i] = s; tmp.push_back(i); }
                }
            for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q)
                for (int32_t i : st[S.children[q]])
                    if (i > last && mark[i] != s) { mark[i] = s; tmp.push_back(i); }
            std::sort(tmp.begin(), tmp.end());
            st[s] = tmp;
        }
        for (int32_t s = 0; s < ns; ++s) {
            S.sdeg[s] = (int32_t)st[s].size();
            if (S.sdeg[s] != merged[s].d) throw Error(-2, "internal: supernode structure size mismatch");
            S.srows_off[s + 1] = S.srows_off[s] + S.sdeg[s];
        }
        S.rows.resize(S.srows_off[ns]);
        for (int32_t s = 0; s < ns; ++s)
            std::copy(st[s].begin(), st[s].end(), S.rows.begin() + S.srows_off[s]);
    }

    // ---- levels (height in the assembly tree), sizes, offsets ----
    S.slevel.assign(ns, 0);
    for (int32_t s = 0; s < ns; ++s)
        if (S.sparent[s] >= 0) S.slevel[S.sparent[s]] = std::max(S.slevel[S.sparent[s]], S.slevel[s] + 1);
    S.num_levels = 0;
    for (int32_t s = 0; s < ns; ++s) S.num_levels = std::max(S.num_levels, S.slevel[s] + 1);
    S.panel_off.resize(ns + 1);
    S.panel_off[0] = 0;
    S.factor_nnz = 0;
    S.flops = 0;
    for (int32_t s = 0; s < ns; ++s) {
        const int64_t W = 3 * (int64_t)S.swidth[s], D = 3 * (int64_t)S.sdeg[s];
        S.panel_off[s + 1] = S.panel_off[s] + ((W + D) * W + 31) / 32 * 32;
        S.factor_nnz += W * (W + 1) / 2 + W * D;
        S.flops += (double)W * W * W / 3.0 + (double)D * W * W + (double)D * D * W;
        S.max_width = std::max(S.max_width, W);
        S.max_degree = std::max(S.max_degree, D);
    }
    S.L_size = S.panel_off[ns];

    // relative indices of each child's structure in its parent's row list
    S.rel_off.assign(ns + 1, 0);
    for (int32_t s = 0; s < ns; ++s) S.rel_off[s + 1] = S.rel_off[s] + (S.sparent[s] >= 0 ? S.sdeg[s] : 0);
    S.rel.resize(S.rel_off[ns]);
    for (int32_t c = 0; c < ns; ++c) {
        const int32_t p = S.sparent[c];
        if (p < 0) continue;
        const int32_t fp = S.sfirst[p], wp = S.swidth[p];
        const int32_t* prow = S.rows.data() + S.srows_off[p];
        const int32_t* pend = prow + S.sdeg[p];
        for (int32_t k = 0; k < S.sdeg[c]; ++k) {
            const int32_t r = S.rows[S.srows_off[c] + k];
            int32_t rp;
            if (r < fp) throw Error(-2, "internal: child row above parent");
            if (r < fp + wp) rp = r - fp;
            else {
                const int32_t* it = std::lower_bound(prow, pend, r);
                if (it == pend || *it != r) throw Error(-2, "internal: child row missing in parent");
                rp = wp + (int32_t)(it - prow);
            }
            S.rel[S.rel_off[c] + k] = rp;
        }
    }

    // ---- injection plan ----
    S.plan_src.resize(red.size());
    S.plan_dst.resize(red.size());
    S.plan_row.resize(red.size());
    S.plan_col.resize(red.size());
    S.diag_src.assign(n, -1);
    for (int64_t j = 0; j < nb; ++j)
        if (S.red_of_full[j] >= 0 && Ap[j + 1] > Ap[j] && Ai[Ap[j + 1] - 1] == j)
            S.diag_src[S.pos[S.red_of_full[j]]] = (int32_t)(Ap[j + 1] - 1);   // -1: no stored diagonal
    for (size_t q = 0; q < red.size(); ++q) {
        const RE& e = red[q];
        const int32_t a = S.pos[e.i], b = S.pos[e.j];
        const int32_t col = std::min(a, b), row = std::max(a, b);
        const int32_t s = S.super_of[col];
        const int32_t f = S.sfirst[s], w = S.swidth[s];
        int64_t rp;
        if (row < f + w) rp = row - f;
        else {
            const int32_t* prow = S.rows.data() + S.srows_off[s];
            const int32_t* it = std::lower_bound(prow, prow + S.sdeg[s], row);
            if (it == prow + S.sdeg[s] || *it != row) throw Error(-2, "internal: injection row missing");
            rp = w + (it - prow);
        }
        const int64_t ld = 3 * ((int64_t)w + S.sdeg[s]);
        const int64_t dst = S.panel_off[s] + 3 * (int64_t)(col - f) * ld + 3 * rp;
        const int64_t mode = (a == b) ? 2 : (a > b ? 0 : 1);
        S.plan_src[q] = e.src;
        S.plan_dst[q] = (dst << 2) | mode;
        S.plan_row[q] = row;
        S.plan_col[q] = col;
    }

    // ---- level schedule: U arena, tile tasks, solve tasks ----
    const int T = opt.tile;
    std::vector<std::vector<int32_t>> bylevel(S.num_levels);
    for (int32_t s = 0; s < ns; ++s) bylevel[S.slevel[s]].push_back(s);
    S.u_off.assign(ns, -1);
    Arena arena;
    S.levels.resize(S.num_levels);
    S.pulls_complete = S.panel_pulls_complete = opt.pull_lists;
    std::vector<int32_t> tile_owner;   // per panel pull task: its supernode
    S.u_solve_off.assign(ns, 0);
    {
        int64_t acc = 0;
        for (int32_t s = 0; s < ns; ++s) { S.u_solve_off[s] = acc; acc += 3 * (int64_t)S.sdeg[s]; }
        S.u_solve_size = acc;
    }
    int64_t max_part = 0;
    const int OB = std::max(opt.outer / T, 1) * T;
    // solve column blocks: per supernode, blocks of SB columns with explicit diagonal inverses
    const int SB = std::min(1024 / T, std::max(opt.solve_block / T, 1)) * T;   // kernels hold <= 1024
    S.solve_block = SB;
    // blocks wider than 512 columns only for wide supernodes (they halve the sequential solve steps
    // at the top of the tree; narrower supernodes keep the cheaper 512-column block inverses)
    auto sbw = [&](int32_t s) { return (SB > 512 && 3 * S.swidth[s] < 4096) ? 512 : SB; };
    std::vector<int32_t> tile_base(ns + 1, 0), sblock0(ns + 1, 0);
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t W = 3 * S.swidth[s];
        tile_base[s + 1] = tile_base[s] + (W + T - 1) / T;
        sblock0[s + 1] = sblock0[s] + (W + sbw(s) - 1) / sbw(s);
    }
    S.tile_slot.assign(tile_base[ns], -1);
    S.sblocks.resize(sblock0[ns]);
    {
        int64_t xb = 0;
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t W = 3 * S.swidth[s], ld = W + 3 * S.sdeg[s];
            for (int32_t k = sblock0[s]; k < sblock0[s + 1]; ++k) {
                const int32_t c0 = (k - sblock0[s]) * sbw(s), nbk = std::min(sbw(s), W - c0);
                SolveBlock& b = S.sblocks[k];
                b.l_off = S.panel_off[s] + (int64_t)c0 * ld + c0;
                b.xb_off = xb;
                b.x_off = 3 * (int64_t)S.sfirst[s] + c0;
                b.ld = ld;
                b.nb = nbk;
                b.ts_off = tile_base[s] + c0 / T;
                b.part_begin = b.part_end = 0;
                b.pad = 0;
                xb += (int64_t)nbk * nbk;
                for (int32_t j = 0; j < (nbk + T - 1) / T; ++j) S.inv_tiles.push_back({k, j});
            }
        }
        S.xb_size = std::max<int64_t>(xb, 1);
    }
    const int BT = opt.big_tile;
    if (opt.schur_tile != 128 && opt.schur_tile != 256) throw Error(-1, "the Schur tile must be 128 or 256");
    S.schur_tile = opt.schur_tile;
    // Tiles of the lower part (rows >= cols, in region coordinates) of the GEMM region
    // C(r, c) (+)= -A(r, :) B(c, :)^T, r in [max(c0,rlo), rhi), c in [c0, c1), with A(r,k) at
    // a_base + r + k*lda, B(c,k) at b_base + c + k*ldb and C(r,c) at c_base + r + c*ldc.
    // 128x128 tiles when the region is large in both dimensions, else 64x64.
    auto emit_lower = [&](std::vector<GemmTask>& big, std::vector<GemmTask>& small, int64_t a_base, int lda,
                          int64_t b_base, int ldb, int64_t c_base, int ldc, int c0, int c1, int rhi, int K,
                          int flags, bool allow_big, bool c_is_u = false) {
        (void)c_is_u;
        if (c1 <= c0 || K <= 0) return;
        // the Schur update U = -L21 L21^T (C in the update arena) may use the larger Schur tile;
        // TRSM, in-panel and trailing updates keep 128x128 tiles
        const int BIG = (flags & GT_C_IN_U) ? opt.schur_tile : 128;
        const int BTH = std::max(BT, BIG);
        const bool use_big = allow_big && (c1 - c0) >= BTH && (rhi - c0) >= BTH;
        const int TT = use_big ? BIG : T;     // the large tile of the GEMM kernel (BT is only the threshold)
        // groups of G tile columns, row by row inside a group: a run of G*G consecutive tasks
        // covers a G x G block of tiles that shares G row panels of A and G column panels of B
        // (the kernel hands each XCD a contiguous run of tasks, so a run shares one L2)
        const int G = std::max(1, opt.tile_group);
        for (int cg = c0; cg < c1; cg += G * TT)
            for (int rr = cg; rr < rhi; rr += TT)
                for (int cc = cg; cc < std::min(c1, cg + G * TT) && cc <= rr; cc += TT) {
                GemmTask g{};
                g.a_off = a_base + rr;
                g.lda = lda;
                g.b_off = b_base + cc;
                g.ldb = ldb;
                g.c_off = c_base + (int64_t)cc * ldc + rr;
                g.ldc = ldc;
                g.M = std::min(TT, rhi - rr);
                g.N = std::min(TT, c1 - cc);
                g.K = K;
                g.flags = flags;
                (use_big ? big : small).push_back(g);
            }
    };
    auto push_tasks = [](std::vector<GemmTask>& all, std::vector<GemmTask>& big, std::vector<GemmTask>& small) {
        TaskRange r;
        r.b = (int64_t)all.size();
        all.insert(all.end(), big.begin(), big.end());
        r.m = (int64_t)all.size();
        all.insert(all.end(), small.begin(), small.end());
        r.e = (int64_t)all.size();
        return r;
    };
    // the supernode's size bucket allows tiles larger than 64x64
    auto big_ok = [&](int32_t s) {
        return opt.bucket_tile[front_bucket(3 * ((int64_t)S.swidth[s] + S.sdeg[s]))] > 64;
    };
    auto fused_sn = [&](int32_t s) {
        return opt.fuse_degree > 0 && 3 * S.swidth[s] <= T && 3 * S.sdeg[s] <= opt.fuse_degree;
    };
    for (int32_t lv = 0; lv < S.num_levels; ++lv) {
        LevelPlan& P = S.levels[lv];
        const auto& sl_all = bylevel[lv];
        std::vector<int32_t> sl;   // supernodes of the tiled multi-launch path
        // allocate U for this level
        for (int32_t s : sl_all) {
            const int64_t D = 3 * (int64_t)S.sdeg[s];
            if (D > 0) S.u_off[s] = arena.alloc(D * D);
        }
        // small supernodes: one fused task each (one launch for the level)
        {
            const int64_t f0 = (int64_t)S.fused_tasks.size();
            for (int32_t s : sl_all) {
                if (!fused_sn(s)) { sl.push_back(s); continue; }
                const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s];
                PotrfTask pt;
                pt.diag_off = S.panel_off[s];
                pt.ld = W + D;
                pt.kb = W;
                pt.slot = (int32_t)S.potrf_tasks.size();
                pt.col0 = 3 * S.sfirst[s];
                S.potrf_tasks.push_back(pt);
                S.tile_slot[tile_base[s]] = pt.slot;
                S.flops_potrf += (double)W * W * W / 3.0;
                FusedTask ft;
                ft.panel_off = S.panel_off[s];
                ft.u_off = S.u_off[s];
                ft.W = W;
                ft.D = D;
                ft.child_begin = S.child_ptr[s];
                ft.child_end = S.child_ptr[s + 1];
                ft.slot = pt.slot;
                ft.col0 = pt.col0;
                ft.wblk = S.swidth[s];
                ft.pad = 0;
                S.fused_tasks.push_back(ft);
            }
            // most expensive first, so the kernel's static stride over the tasks balances
            std::stable_sort(S.fused_tasks.begin() + f0, S.fused_tasks.end(), [](const FusedTask& x, const FusedTask& y) {
                const double cx = (double)x.D * x.D * x.W + (double)x.W * x.W * (x.W + x.D);
                const double cy = (double)y.D * y.D * y.W + (double)y.W * y.W * (y.W + y.D);
                return cx > cy;
            });
            P.fused = {f0, (int64_t)S.fused_tasks.size()};
        }
        // extend-add passes by child rank (no two tasks of a pass touch the same parent entry).
        // Column blocks of a child's U that land in the parent's panel are added before the
        // parent is factored; those that land in the parent's U are added after the parent's
        // Schur update, which writes U with beta = 0, so U never needs clearing.
        int32_t maxc = 0;
        for (int32_t s : sl) maxc = std::max(maxc, S.child_ptr[s + 1] - S.child_ptr[s]);
        auto extend_passes = [&](bool to_u, std::vector<std::pair<int64_t,int64_t>>& passes) {
        for (int32_t r = 0; r < maxc; ++r) {
            const int64_t b0 = (int64_t)S.extend_tasks.size();
            for (int32_t p : sl) {
                if (S.child_ptr[p] + r >= S.child_ptr[p + 1]) continue;
                const int32_t c = S.children[S.child_ptr[p] + r];
                if (S.sdeg[c] == 0) continue;
                for (int32_t kj = 0; kj < S.sdeg[c]; ++kj) {
                    if ((S.rel[S.rel_off[c] + kj] >= S.swidth[p]) != to_u) continue;
                    ExtendTask t;
                    t.u_child = S.u_off[c];
                    t.p_panel = S.panel_off[p];
                    t.u_parent = S.u_off[p];
                    t.dc = S.sdeg[c];
                    t.ldp = 3 * (S.swidth[p] + S.sdeg[p]);
                    t.wp = S.swidth[p];
                    t.dp = S.sdeg[p];
                    t.rel_off = S.rel_off[c];
                    t.col_block = kj;
                    // long columns are split into row chunks (disjoint parent entries)
                    for (int32_t r0 = 3 * kj; r0 < 3 * S.sdeg[c]; r0 += kExtendRows) {
                        t.row0 = r0;
                        S.extend_tasks.push_back(t);
                    }
                }
            }
            if ((int64_t)S.extend_tasks.size() > b0) passes.push_back({b0, (int64_t)S.extend_tasks.size()});
        }
        };
        extend_passes(false, P.extend_passes);
        // the same panel part as children lists of the parents' 64 x 64 panel tiles (lower part):
        // per tile and child, the child's block rows and panel block columns that meet the tile
        if (S.panel_pulls_complete) {
            const int64_t t0 = (int64_t)S.panel_pull_tasks.size();
            for (int32_t p : sl) {
                const int32_t Wb = S.swidth[p], W = 3 * Wb, ld = 3 * (Wb + S.sdeg[p]);
                bool any = false;      // a child with a column block in the panel (rel ascending)
                for (int32_t q = S.child_ptr[p]; q < S.child_ptr[p + 1]; ++q) {
                    const int32_t c = S.children[q];
                    if (S.sdeg[c] > 0 && S.rel[S.rel_off[c]] < Wb) any = true;
                }
                if (!any) continue;
                for (int32_t c0 = 0; c0 < W; c0 += 64)
                    for (int32_t r0 = c0; r0 < ld; r0 += 64) {
                        const int32_t M = std::min(64, ld - r0), N = std::min(64, W - c0);
                        const int32_t begin = (int32_t)S.pull_ents.size();
                        for (int32_t q = S.child_ptr[p]; q < S.child_ptr[p + 1]; ++q) {
                            const int32_t c = S.children[q];
                            const int32_t dc = S.sdeg[c];
                            if (dc == 0) continue;
                            const int32_t* rl = S.rel.data() + S.rel_off[c];
                            auto range = [&](int32_t o, int32_t n, int32_t& b0, int32_t& nb) {
                                const int32_t lo = o / 3, hi = (o + n - 1) / 3 + 1;
                                b0 = (int32_t)(std::lower_bound(rl, rl + dc, lo) - rl);
                                nb = (int32_t)(std::lower_bound(rl, rl + dc, hi) - rl) - b0;
                            };
                            PullEnt e{};
                            range(r0, M, e.ib0, e.nib);
                            range(c0, N, e.jb0, e.njb);
                            if (e.nib <= 0 || e.njb <= 0 || e.ib0 + e.nib - 1 < e.jb0) continue;
                            // the kernel addresses the entry's 3 njb child columns with 32-bit offsets
                            if (9 * (int64_t)e.njb * dc * (int64_t)sizeof(double) >= ((int64_t)1 << 31))
                                S.panel_pulls_complete = false;
                            e.u_child = S.u_off[c];
                            e.dc3 = 3 * dc;
                            e.rel_off = (int32_t)S.rel_off[c];
                            e.base_r = r0;
                            e.base_c = c0;
                            S.pull_ents.push_back(e);
                        }
                        const int32_t n = (int32_t)S.pull_ents.size() - begin;
                        if (n > 0) {
                            S.panel_pull_tasks.push_back(
                                {S.panel_off[p] + (int64_t)c0 * ld + r0, ld, M, N, begin, n, 0, 0, 0});
                            tile_owner.push_back(p);
                        }
                    }
            }
            P.panel_pulls = {t0, (int64_t)S.panel_pull_tasks.size()};
        }
        // release children's U (consumed by this level's extend-add; the U passes run at the end
        // of the level and nothing else is allocated in between)
        for (int32_t p : sl_all)
            for (int32_t q = S.child_ptr[p]; q < S.child_ptr[p + 1]; ++q) {
                const int32_t c = S.children[q];
                if (S.u_off[c] >= 0) {
                    const int64_t D = 3 * (int64_t)S.sdeg[c];
                    arena.release(S.u_off[c], D * D);
                }
            }
        // dense steps: per supernode, outer blocks of `outer` columns; inside, `tile`-wide steps
        // (diagonal POTRF + tile inverse, TRSM of all rows below, then the left-looking update of
        // the next column block from the outer block's finished columns); after an outer block,
        // one right-looking trailing update of all later columns with K = outer.
        int32_t maxsteps = 0;
        for (int32_t s : sl) maxsteps = std::max<int32_t>(maxsteps, (3 * S.swidth[s] + T - 1) / T);
        std::vector<int64_t> slot_of_step;   // per (supernode at this level) current step's inverse slot
        for (int32_t t = 0; t < maxsteps; ++t) {
            const int64_t p0 = (int64_t)S.potrf_tasks.size();
            std::vector<std::pair<int32_t,int64_t>> active;   // (sn, inverse slot)
            for (int32_t s : sl) {
                const int32_t W = 3 * S.swidth[s];
                if (t * T >= W) continue;
                const int32_t D = 3 * S.sdeg[s], ld = W + D, k0 = t * T, kb = std::min(T, W - k0);
                PotrfTask pt;
                pt.diag_off = S.panel_off[s] + (int64_t)k0 * ld + k0;
                pt.ld = ld;
                pt.kb = kb;
                pt.slot = (int32_t)S.potrf_tasks.size();
                pt.col0 = 3 * S.sfirst[s] + k0;
                S.potrf_tasks.push_back(pt);
                S.tile_slot[tile_base[s] + t] = pt.slot;
                S.flops_potrf += (double)kb * kb * kb / 3.0;
                active.push_back({s, pt.slot});
            }
            P.potrf_steps.push_back({p0, (int64_t)S.potrf_tasks.size()});
            {   // TRSM: X := X L_kk^{-T} on all rows below the diagonal tile (64x64 tiles, in place)
                std::vector<GemmTask> big, small;
                for (auto [s, slot] : active) {
                    const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s], ld = W + D;
                    const int32_t k0 = t * T, kb = std::min(T, W - k0);
                    const int64_t off = S.panel_off[s];
                    for (int32_t r0 = k0 + kb; r0 < W + D; r0 += T) {
                        GemmTask g{};
                        g.a_off = off + (int64_t)k0 * ld + r0;
                        g.lda = ld;
                        g.b_off = slot * (int64_t)T * T;
                        g.ldb = T;
                        g.c_off = g.a_off;
                        g.ldc = ld;
                        g.M = std::min(T, W + D - r0);
                        g.N = kb;
                        g.K = kb;
                        g.flags = GT_BETA0 | GT_ALPHA_POS | GT_B_IN_INV;
                        small.push_back(g);
                    }
                }
                const TaskRange tr = push_tasks(S.gemm_tasks, big, small);
                P.trsm_steps.push_back(tr);
                // fused diagonal tile + TRSM: one workgroup per row tile (one per POTRF task
                // without rows below), tasks in the order they were emitted above
                const int64_t nrow = tr.e - tr.b, np = (int64_t)active.size();
                const int64_t i0 = (int64_t)S.panel_items.size();
                if (np > 0 && std::max(nrow, np) <= opt.panel_max) {
                    int64_t g = tr.m;
                    for (size_t a = 0; a < active.size(); ++a) {
                        const int32_t s = active[a].first;
                        const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s];
                        const int32_t k1 = t * T + std::min(T, W - t * T);
                        const int32_t nt = (W + D - k1 + T - 1) / T;
                        const int32_t pi = (int32_t)(p0 + (int64_t)a);
                        if (nt == 0) S.panel_items.push_back({pi, -1, 1, 0});
                        for (int32_t q = 0; q < nt; ++q) S.panel_items.push_back({pi, (int32_t)g++, nt, 0});
                    }
                    if (g != tr.e) throw Error(-1, "internal: panel items do not cover the TRSM tasks");
                }
                P.panel_steps.push_back({i0, (int64_t)S.panel_items.size()});
            }
            {   // after TRSM(t): the next step's column block gets the left-looking update from the
                // outer block's finished columns (K = k1 - J0), or, when step t closes its outer
                // block, all later columns get the outer trailing update (K = outer).  With
                // lookahead the trailing update is split: the next outer block's columns stay in
                // the step (they are next on the critical path), the columns after them form the
                // step's rest range, which may run beside the next outer block's steps.
                std::vector<GemmTask> big, small, rbig, rsmall;
                bool trailing = false;
                auto emit_step = [&](bool allow_big) {
                    for (auto [s, slot] : active) {
                        (void)slot;
                        const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s], ld = W + D;
                        const int32_t k0 = t * T, kb = std::min(T, W - k0), k1 = k0 + kb;
                        const int32_t J0 = (k0 / OB) * OB, Jend = std::min(W, J0 + OB);
                        const int64_t off = S.panel_off[s];
                        if (k1 < Jend) {
                            const int32_t kb1 = std::min(T, W - k1);
                            emit_lower(big, small, off + (int64_t)J0 * ld, ld, off + (int64_t)J0 * ld, ld, off,
                                       ld, k1, k1 + kb1, W + D, k1 - J0, 0, allow_big && big_ok(s));
                        } else if (Jend < W) {
                            trailing = true;
                            const int32_t split = opt.lookahead ? std::min(W, Jend + OB) : W;
                            emit_lower(big, small, off + (int64_t)J0 * ld, ld, off + (int64_t)J0 * ld, ld, off,
                                       ld, Jend, split, W + D, Jend - J0, 0, allow_big && big_ok(s));
                            if (split < W)
                                emit_lower(rbig, rsmall, off + (int64_t)J0 * ld, ld, off + (int64_t)J0 * ld, ld,
                                           off, ld, split, W, W + D, Jend - J0, 0, big_ok(s));
                        }
                    }
                };
                emit_step(true);
                // a critical-path launch of few 128x128 tiles leaves CUs idle or gives some CUs
                // two tiles: it then runs on 64x64 tiles (four per 128x128 tile) instead.  The
                // two tile kernels issue the same MFMAs per entry in the same order, so the factor
                // is bitwise the same.  Measured on the 64^3 step (MFS_SPLIT_BIG, one box): below
                // 256 tiles FP64 254.1 -> 253.3 ms, FP32 154.3 -> 153.7 ms; 128 to 1000 all
                // land within 0.2 % of that.
                if (!big.empty() && (int64_t)big.size() < opt.split_big) {
                    big.clear();
                    small.clear();
                    rbig.clear();
                    rsmall.clear();
                    emit_step(false);
                }
                P.update_steps.push_back(push_tasks(S.gemm_tasks, big, small));
                P.rest_steps.push_back(push_tasks(S.gemm_tasks, rbig, rsmall));
                P.trailing_step.push_back(trailing ? 1 : 0);
            }
        }
        // Schur update U -= L21 L21^T (lower tiles), K = W.  Each tile also lists, in child order,
        // the children's update-matrix blocks that land in it (the part of the extend-add into U
        // that the tile's epilogue can add instead of the U extend-add passes)
        {
            std::vector<GemmTask> big, small;
            auto attach_pulls = [&](GemmTask& g, int32_t s) {
                const int64_t D = 3 * (int64_t)S.sdeg[s];
                const int64_t rel0 = g.c_off - S.u_off[s];
                const int32_t rr = (int32_t)(rel0 % D), cc = (int32_t)(rel0 / D);
                const int32_t Wb = S.swidth[s];
                const int32_t begin = (int32_t)S.pull_ents.size();
                if (!S.pulls_complete || g.M > 256 || g.N > 256) { S.pulls_complete = false; return; }
                for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q) {
                    const int32_t c = S.children[q];
                    const int32_t dc = S.sdeg[c];
                    if (dc == 0) continue;
                    const int32_t* rl = S.rel.data() + S.rel_off[c];
                    // parent block rows/columns whose scalars meet the tile: [lo, hi)
                    auto range = [&](int32_t o, int32_t n, int32_t& b0, int32_t& nb) {
                        const int32_t lo = Wb + o / 3, hi = Wb + (o + n - 1) / 3 + 1;
                        b0 = (int32_t)(std::lower_bound(rl, rl + dc, lo) - rl);
                        nb = (int32_t)(std::lower_bound(rl, rl + dc, hi) - rl) - b0;
                    };
                    PullEnt e{};
                    range(rr, g.M, e.ib0, e.nib);
                    range(cc, g.N, e.jb0, e.njb);
                    if (e.nib <= 0 || e.njb <= 0 || e.ib0 + e.nib - 1 < e.jb0) continue;   // nothing lower
                    if (e.nib > 128 || e.njb > 128) throw Error(-1, "internal: Schur tile pull exceeds 128 blocks");
                    // the tile kernel addresses the entry's 3 njb child columns with 32-bit offsets
                    if (9 * (int64_t)e.njb * dc * (int64_t)sizeof(double) >= ((int64_t)1 << 31)) {
                        S.pulls_complete = false;
                        return;
                    }
                    e.u_child = S.u_off[c];
                    e.dc3 = 3 * dc;
                    e.rel_off = (int32_t)S.rel_off[c];
                    e.base_r = 3 * Wb + rr;
                    e.base_c = 3 * Wb + cc;
                    S.pull_ents.push_back(e);
                }
                const int32_t n = (int32_t)S.pull_ents.size() - begin;
                if (n > 0) {
                    g.pad = begin;
                    g.flags |= n << GT_PULL_SHIFT;
                }
            };
            for (int32_t s : sl) {
                const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s], ld = W + D;
                if (D == 0) continue;
                const int64_t off = S.panel_off[s];
                const size_t nb0 = big.size(), ns0 = small.size();
                // region rows/cols are indices into L21 (0..D); A = P[W + r, 0..W), C = U[r, c]
                emit_lower(big, small, off + W, ld, off + W, ld, S.u_off[s], D, 0, D, D, W, GT_C_IN_U | GT_BETA0,
                           big_ok(s), /*c_is_u=*/true);
                for (size_t q = nb0; q < big.size(); ++q) attach_pulls(big[q], s);
                for (size_t q = ns0; q < small.size(); ++q) attach_pulls(small[q], s);
            }
            P.syrk = push_tasks(S.gemm_tasks, big, small);
        }
        extend_passes(true, P.extend_u_passes);
        // solves.  Gather tasks (one per supernode of the level) serve both directions.
        {
            const int64_t g0 = (int64_t)S.gather_tasks.size();
            for (int32_t s : sl_all) {
                SolveGatherTask g{};
                g.sn = s;
                g.child_begin = S.child_ptr[s];
                g.child_end = S.child_ptr[s + 1];
                g.w = 3 * S.swidth[s];
                g.d = 3 * S.sdeg[s];
                g.x_off = 3 * (int64_t)S.sfirst[s];
                g.u_off = S.u_solve_off[s];
                g.rows_off = S.srows_off[s];
                S.gather_tasks.push_back(g);
            }
            P.gather = {g0, (int64_t)S.gather_tasks.size()};
        }
        int32_t maxblk = 0;
        for (int32_t s : sl_all) maxblk = std::max(maxblk, sblock0[s + 1] - sblock0[s]);
        // row ranges [lo, hi) split into chunks of at most `ch` rows that do not straddle W
        auto chunks = [](int32_t lo, int32_t W, int32_t hi, int32_t ch) {
            std::vector<std::pair<int32_t,int32_t>> out;
            for (int32_t r = lo; r < std::min(W, hi); r += ch) out.push_back({r, std::min(r + ch, std::min(W, hi))});
            for (int32_t r = std::max(lo, W); r < hi; r += ch) out.push_back({r, std::min(r + ch, hi)});
            return out;
        };
        // forward: per step k, y_b = X_b x_b for block k of every supernode that has one
        // (one workgroup per 64-row tile), then every row below the block is updated
        for (int32_t k = 0; k < maxblk; ++k) {
            const int64_t d0 = (int64_t)S.fdiag_tiles.size(), u0 = (int64_t)S.fupd_tasks.size();
            for (int32_t s : sl_all) {
                if (k >= sblock0[s + 1] - sblock0[s]) continue;
                const int32_t bi = sblock0[s] + k;
                const SolveBlock& b = S.sblocks[bi];
                const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s], c0 = k * sbw(s);
                for (int32_t t = 0; t < (b.nb + T - 1) / T; ++t) S.fdiag_tiles.push_back({bi, t});
                for (auto [r0, r1] : chunks(c0 + b.nb, W, W + D, T)) {
                    SolveRowTask rt{};
                    rt.l_off = S.panel_off[s] + (int64_t)c0 * b.ld + r0;
                    rt.v_off = b.x_off;
                    rt.mode = r0 < W ? 0 : 1;
                    rt.d_off = r0 < W ? 3 * (int64_t)S.sfirst[s] + r0 : S.u_solve_off[s] + (r0 - W);
                    rt.ld = b.ld;
                    rt.nr = r1 - r0;
                    rt.nc = b.nb;
                    S.fupd_tasks.push_back(rt);
                }
            }
            P.fdiag_steps.push_back({d0, (int64_t)S.fdiag_tiles.size()});
            P.fupd_steps.push_back({u0, (int64_t)S.fupd_tasks.size()});
        }
        // backward: per step (last block first), partial sums of L(rows below, b)^T x over row
        // chunks and 16-column groups, then x_b = X_b^T (y_b - sum of partials) per 64-column tile
        {
            const int32_t RC = std::min(1024, std::max(T, opt.solve_chunk / T * T));
            const int32_t CG = 16;
            for (int32_t st = 0; st < maxblk; ++st) {
                const int64_t p0 = (int64_t)S.bpart_tasks.size(), d0 = (int64_t)S.bdiag_tiles.size();
                int32_t slot = 0;
                for (int32_t s : sl_all) {
                    const int32_t nbs = sblock0[s + 1] - sblock0[s];
                    if (st >= nbs) continue;
                    const int32_t k = nbs - 1 - st, bi = sblock0[s] + k;
                    SolveBlock& b = S.sblocks[bi];
                    const int32_t W = 3 * S.swidth[s], D = 3 * S.sdeg[s], c0 = k * sbw(s);
                    b.part_begin = slot;
                    for (auto [r0, r1] : chunks(c0 + b.nb, W, W + D, RC)) {
                        for (int32_t g = 0; g < b.nb; g += CG) {
                            SolveRowTask rt{};
                            rt.l_off = S.panel_off[s] + (int64_t)(c0 + g) * b.ld + r0;
                            rt.mode = r0 < W ? 0 : 1;
                            rt.v_off = r0 < W ? 3 * (int64_t)S.sfirst[s] + r0 : S.u_solve_off[s] + (r0 - W);
                            rt.d_off = (int64_t)slot * SB + g;
                            rt.ld = b.ld;
                            rt.nr = r1 - r0;
                            rt.nc = std::min(CG, b.nb - g);
                            S.bpart_tasks.push_back(rt);
                        }
                        ++slot;
                    }
                    b.part_end = slot;
                    for (int32_t j = 0; j < (b.nb + T - 1) / T; ++j) S.bdiag_tiles.push_back({bi, j});
                }
                max_part = std::max<int64_t>(max_part, slot);
                P.bpart_steps.push_back({p0, (int64_t)S.bpart_tasks.size()});
                P.bdiag_steps.push_back({d0, (int64_t)S.bdiag_tiles.size()});
            }
        }
    }
    // every task must fit the tile of the kernel that runs it: [b, m) 128x128, [m, e) 64x64
    auto check_range = [&](const TaskRange& r) {
        for (int64_t q = r.b; q < r.e; ++q) {
            const GemmTask& g = S.gemm_tasks[q];
            const int lim = q < r.m ? ((g.flags & GT_C_IN_U) ? S.schur_tile : 128) : T;
            if (g.M < 1 || g.N < 1 || g.M > lim || g.N > lim || g.K < 1)
                throw Error(-2, "internal: GEMM task does not fit its tile kernel");
        }
    };
    for (const LevelPlan& P : S.levels) {
        for (const TaskRange& r : P.trsm_steps) check_range(r);
        for (const TaskRange& r : P.update_steps) check_range(r);
        for (const TaskRange& r : P.rest_steps) check_range(r);
        check_range(P.syrk);
    }
    // fused panel items: each TRSM task solves with its own POTRF task's tile, fits one 64-row
    // tile, and the items of a POTRF task are `count` consecutive entries
    for (size_t q = 0; q < S.panel_items.size();) {
        const PanelItem& it = S.panel_items[q];
        const PotrfTask& pt = S.potrf_tasks.at(it.potrf);
        if (it.count < 1 || q + it.count > S.panel_items.size())
            throw Error(-2, "internal: panel item count");
        for (int32_t e = 0; e < it.count; ++e) {
            const PanelItem& ie = S.panel_items[q + e];
            if (ie.potrf != it.potrf || ie.count != it.count) throw Error(-2, "internal: panel item group");
            if (ie.trsm >= 0) {
                const GemmTask& g = S.gemm_tasks.at(ie.trsm);
                if (g.b_off != (int64_t)pt.slot * T * T || g.N != pt.kb || g.K != pt.kb || g.M < 1 || g.M > T ||
                    g.flags != (GT_BETA0 | GT_ALPHA_POS | GT_B_IN_INV))
                    throw Error(-2, "internal: panel item TRSM task does not match its tile");
            } else if (it.count != 1) {
                throw Error(-2, "internal: panel item without rows must be alone");
            }
        }
        q += it.count;
    }
    S.U_size = std::max<int64_t>(arena.peak, 32);
    S.inv_slots = std::max<int64_t>((int64_t)S.potrf_tasks.size(), 1);
    S.bpart_size = std::max<int64_t>(max_part, 1);

    // ---- injection entries of the panel tiles (every 3x3 block that meets a tile) ----
    if (S.panel_pulls_complete) {
        std::vector<std::vector<int64_t>> plan_of(ns);     // injection entries per supernode
        for (size_t q = 0; q < S.plan_dst.size(); ++q) {
            const int64_t dst = S.plan_dst[q] >> 2;
            const int32_t s = (int32_t)(std::upper_bound(S.panel_off.begin(), S.panel_off.end(), dst) -
                                        S.panel_off.begin()) - 1;
            plan_of[s].push_back((int64_t)q);
        }
        std::vector<std::vector<TileInj>> lists(S.panel_pull_tasks.size());
        std::vector<int32_t> tid;                           // per tile of a panel: task or -1
        for (size_t a = 0; a < S.panel_pull_tasks.size();) {
            const int32_t p = tile_owner[a];
            size_t b = a;
            while (b < S.panel_pull_tasks.size() && tile_owner[b] == p) ++b;
            const int64_t ld = 3 * ((int64_t)S.swidth[p] + S.sdeg[p]);
            const int32_t ntc = (3 * S.swidth[p] + 63) / 64, ntr = (int32_t)((ld + 63) / 64);
            tid.assign((size_t)ntc * ntr, -1);
            for (size_t q = a; q < b; ++q) {
                const int64_t o = S.panel_pull_tasks[q].c_off - S.panel_off[p];
                tid[(size_t)(o / ld / 64) * ntr + (size_t)(o % ld / 64)] = (int32_t)q;
            }
            for (int64_t q : plan_of[p]) {
                const int64_t o = (S.plan_dst[q] >> 2) - S.panel_off[p];
                const int32_t row = (int32_t)(o % ld), col = (int32_t)(o / ld);
                for (int32_t tc = col / 64; tc <= (col + 2) / 64 && tc < ntc; ++tc)
                    for (int32_t tr = row / 64; tr <= (row + 2) / 64 && tr < ntr; ++tr) {
                        const int32_t k = tid[(size_t)tc * ntr + tr];
                        if (k >= 0)
                            lists[k].push_back({(int32_t)q, (int16_t)(row - 64 * tr), (int16_t)(col - 64 * tc)});
                    }
            }
            a = b;
        }
        for (size_t k = 0; k < lists.size(); ++k) {
            S.panel_pull_tasks[k].inj_begin = (int32_t)S.tile_inj.size();
            S.panel_pull_tasks[k].ninj = (int32_t)lists[k].size();
            S.tile_inj.insert(S.tile_inj.end(), lists[k].begin(), lists[k].end());
        }
        // every other 64 x 64 tile of every panel's lower part is cleared (tiles wholly above the
        // diagonal hold only entries above it, which no kernel reads)
        std::vector<int64_t> tfirst(ns, 0), tcount(ns, 0);    // a supernode's tasks are contiguous
        for (size_t a = S.panel_pull_tasks.size(); a-- > 0;) {
            tfirst[tile_owner[a]] = (int64_t)a;
            ++tcount[tile_owner[a]];
        }
        std::vector<char> pulled;
        for (int32_t s = 0; s < ns; ++s) {
            const int64_t ld = 3 * ((int64_t)S.swidth[s] + S.sdeg[s]);
            const int32_t W = 3 * S.swidth[s], ntc = (W + 63) / 64, ntr = (int32_t)((ld + 63) / 64);
            pulled.assign((size_t)ntc * ntr, 0);
            for (int64_t a = tfirst[s]; a < tfirst[s] + tcount[s]; ++a) {
                const int64_t o = S.panel_pull_tasks[a].c_off - S.panel_off[s];
                pulled[(size_t)(o / ld / 64) * ntr + (size_t)(o % ld / 64)] = 1;
            }
            for (int32_t tc = 0; tc < ntc; ++tc)
                for (int32_t tr = tc; tr < ntr; ++tr)
                    if (!pulled[(size_t)tc * ntr + tr])
                        S.fill_tiles.push_back({S.panel_off[s] + (int64_t)64 * tc * ld + 64 * tr, (int32_t)ld,
                                                (int32_t)std::min<int64_t>(64, ld - 64 * tr),
                                                std::min(64, W - 64 * tc), 0});
        }
    }

    // ---- symmetric SpMV map over the full pattern (block-CSR, both triangles) ----
    {
        std::vector<int64_t> cnt(nb + 1, 0);
        for (int64_t j = 0; j < nb; ++j)
            for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
                cnt[j + 1]++;                          // (Ai[k], j) contributes to row j
                if (Ai[k] != j) cnt[Ai[k] + 1]++;      // and to row Ai[k]
            }
        for (int64_t i = 0; i < nb; ++i) cnt[i + 1] += cnt[i];
        S.spmv_ptr = cnt;
        S.spmv_ent.resize(cnt[nb]);
        S.spmv_col.resize(cnt[nb]);
        std::vector<int64_t> bk(cnt.begin(), cnt.end() - 1);
        // row i: first its own column's blocks (rows j <= i, transposed / diagonal), ascending j
        for (int64_t i = 0; i < nb; ++i)
            for (int64_t k = Ap[i]; k < Ap[i + 1]; ++k) {
                const int64_t j = Ai[k];
                S.spmv_ent[bk[i]] = (k << 2) | (j == i ? 2 : 1);
                S.spmv_col[bk[i]++] = (int32_t)j;
            }
        // then the blocks (i, j), j > i, ascending j
        for (int64_t j = 0; j < nb; ++j)
            for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
                const int64_t i = Ai[k];
                if (i == j) continue;
                S.spmv_ent[bk[i]] = (k << 2) | 0;
                S.spmv_col[bk[i]++] = (int32_t)j;
            }
    }

    // ---- pin / permutation maps (scalar) ----
    if (any_mask) {
        S.comp_mask.assign(n, 0);
        for (int64_t b = 0; b < nb; ++b)
            if (S.red_of_full[b] >= 0) S.comp_mask[S.pos[S.red_of_full[b]]] = bmask[b];
    }
    const int64_t ni = 3 * (int64_t)n;
    auto is_masked = [&](int64_t p) { return any_mask && ((S.comp_mask[p / 3] >> (p % 3)) & 1); };
    std::vector<int64_t> int_to_uperm(ni, -1);
    S.uperm_to_int.resize(S.n_free);
    {
        int64_t u = 0;
        for (int64_t p = 0; p < ni; ++p)
            if (!is_masked(p)) {
                if (u >= S.n_free) throw Error(-2, "internal: free scalar count");
                int_to_uperm[p] = u;
                S.uperm_to_int[u++] = p;
            }
        if (u != S.n_free) throw Error(-2, "internal: free scalar count");
    }
    S.full_to_perm.assign(nsf, -1);
    S.full_to_red.assign(nsf, -1);
    S.full_to_uperm.assign(nsf, -1);
    S.red_to_int.resize(S.n_free);
    int64_t ur = 0;
    for (int64_t v = 0; v < nsf; ++v) {
        const int64_t r = S.red_of_full[v / 3];
        if (r < 0 || ((bmask[v / 3] >> (v % 3)) & 1)) continue;
        const int64_t p = 3 * (int64_t)S.pos[r] + v % 3;
        S.full_to_perm[v] = p;
        S.full_to_red[v] = ur;
        S.full_to_uperm[v] = int_to_uperm[p];
        S.red_to_int[ur++] = p;
    }
    // nnz of the factor restricted to the free scalars (masked rows and columns hold the identity)
    S.factor_nnz_free = S.factor_nnz;
    if (any_mask) {
        S.factor_nnz_free = 0;
        for (int32_t sn = 0; sn < S.ns; ++sn) {
            const int64_t W = 3 * (int64_t)S.swidth[sn], D = 3 * (int64_t)S.sdeg[sn];
            int64_t struct_free = 0;
            for (int64_t r = 0; r < D; ++r)
                struct_free += !is_masked(3 * (int64_t)S.rows[S.srows_off[sn] + r / 3] + r % 3);
            const int64_t f0 = 3 * (int64_t)S.sfirst[sn];
            int64_t below = 0;      // free own columns at or after c
            for (int64_t c = W - 1; c >= 0; --c)
                if (!is_masked(f0 + c)) S.factor_nnz_free += ++below + struct_free;
        }
    }
    return S;
}

} // namespace mfs
