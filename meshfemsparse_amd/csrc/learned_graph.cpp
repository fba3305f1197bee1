// Host side of the learned preconditioner: the free-node block graph of A, a greedy multicolor
// ordering (adjacent nodes get different colors, so the factor's triangular solves run one color
// at a time), directed edge lists by destination and by source, and the lower blocks of the
// factor in that ordering.  Built once per (pattern, pins).
#include "learned.h"

#include <algorithm>

namespace mfs {

LPGraph build_lp_graph(const int64_t* Ap, const int64_t* Ai, int64_t nb, const int64_t* pins, int64_t npins) {
    LPGraph g;
    if (nb <= 0) throw Error(-1, "empty matrix");
    if (nb >= (int64_t(1) << 30)) throw Error(-6, "matrix too large");
    g.nb_full = nb;
    std::vector<int8_t> pc(nb, 0);
    {
        std::vector<char> seen((size_t)nb * 3, 0);
        for (int64_t q = 0; q < npins; ++q) {
            const int64_t v = pins[q];
            if (v < 0 || v >= 3 * nb) throw Error(-1, "Fixed variable index out of range");
            if (!seen[v]) { seen[v] = 1; pc[v / 3]++; }
        }
        for (int64_t b = 0; b < nb; ++b)
            if (pc[b] != 0 && pc[b] != 3) throw Error(-6, "Partially-pinned block variables are not supported");
    }
    g.free_of_full.assign(nb, -1);
    for (int64_t b = 0; b < nb; ++b)
        if (!pc[b]) { g.free_of_full[b] = (int64_t)g.full_of.size(); g.full_of.push_back(b); }
    const int32_t n = (int32_t)g.full_of.size();
    if (n == 0) throw Error(-1, "every variable is pinned");
    g.n = n;
    g.diagblk.assign(n, -1);
    // adjacency with block ids
    std::vector<int64_t> deg(n + 1, 0);
    for (int64_t j = 0; j < nb; ++j) {
        const int64_t fj = g.free_of_full[j];
        for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
            const int64_t i = Ai[k], fi = g.free_of_full[i];
            if (fj < 0 || fi < 0) continue;
            if (i == j) { g.diagblk[fj] = k; continue; }
            deg[fi + 1]++;
            deg[fj + 1]++;
        }
    }
    for (int32_t v = 0; v < n; ++v)
        if (g.diagblk[v] < 0) throw Error(-1, "every free block needs a stored diagonal block");
    for (int32_t v = 0; v < n; ++v) deg[v + 1] += deg[v];
    g.dptr = deg;
    const int64_t E = deg[n];
    g.edst.resize(E);
    g.esrc.resize(E);
    g.eblk.resize(E);
    {
        std::vector<int64_t> bk(deg.begin(), deg.end() - 1);
        for (int64_t j = 0; j < nb; ++j) {
            const int64_t fj = g.free_of_full[j];
            for (int64_t k = Ap[j]; k < Ap[j + 1]; ++k) {
                const int64_t i = Ai[k], fi = g.free_of_full[i];
                if (fj < 0 || fi < 0 || i == j) continue;
                // stored block A(i, j), i < j: edge (i <- j) uses it as is, edge (j <- i) transposed
                int64_t q = bk[fi]++;
                g.edst[q] = (int32_t)fi; g.esrc[q] = (int32_t)fj; g.eblk[q] = k << 1;
                q = bk[fj]++;
                g.edst[q] = (int32_t)fj; g.esrc[q] = (int32_t)fi; g.eblk[q] = (k << 1) | 1;
            }
        }
        // sort each destination's edges by source
        std::vector<std::pair<int32_t, int64_t>> tmp;
        for (int32_t v = 0; v < n; ++v) {
            tmp.clear();
            for (int64_t q = deg[v]; q < deg[v + 1]; ++q) tmp.push_back({g.esrc[q], g.eblk[q]});
            std::sort(tmp.begin(), tmp.end());
            for (int64_t q = deg[v]; q < deg[v + 1]; ++q) {
                g.esrc[q] = tmp[q - deg[v]].first;
                g.eblk[q] = tmp[q - deg[v]].second;
            }
        }
    }
    // greedy coloring in node order
    g.color.assign(n, -1);
    {
        std::vector<int32_t> mark;
        for (int32_t v = 0; v < n; ++v) {
            for (int64_t q = g.dptr[v]; q < g.dptr[v + 1]; ++q) {
                const int32_t c = g.color[g.esrc[q]];
                if (c >= 0) {
                    if ((int32_t)mark.size() <= c) mark.resize(c + 1, -1);
                    mark[c] = v;
                }
            }
            int32_t c = 0;
            while (c < (int32_t)mark.size() && mark[c] == v) ++c;
            g.color[v] = c;
            g.ncolors = std::max(g.ncolors, c + 1);
        }
    }
    g.cptr.assign(g.ncolors + 1, 0);
    for (int32_t v = 0; v < n; ++v) g.cptr[g.color[v] + 1]++;
    for (int32_t c = 0; c < g.ncolors; ++c) g.cptr[c + 1] += g.cptr[c];
    g.order.resize(n);
    g.pos.resize(n);
    {
        std::vector<int32_t> bk(g.cptr.begin(), g.cptr.end() - 1);
        for (int32_t v = 0; v < n; ++v) {
            g.pos[v] = bk[g.color[v]]++;
            g.order[g.pos[v]] = v;
        }
    }
    // by source
    g.sptr.assign(n + 1, 0);
    for (int64_t q = 0; q < E; ++q) g.sptr[g.esrc[q] + 1]++;
    for (int32_t v = 0; v < n; ++v) g.sptr[v + 1] += g.sptr[v];
    g.sedge.resize(E);
    {
        std::vector<int64_t> bk(g.sptr.begin(), g.sptr.end() - 1);
        for (int64_t q = 0; q < E; ++q) g.sedge[bk[g.esrc[q]]++] = (int32_t)q;
    }
    // lower blocks in (row, column) order
    g.lrptr.assign(n + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        for (int64_t q = g.dptr[i]; q < g.dptr[i + 1]; ++q)
            if (g.pos[i] > g.pos[g.esrc[q]]) {
                g.ledge.push_back((int32_t)q);
                g.ldst.push_back(i);
                g.lsrc.push_back(g.esrc[q]);
            }
        g.lrptr[i + 1] = (int64_t)g.ledge.size();
    }
    const int64_t nl = (int64_t)g.ledge.size();
    g.lcptr.assign(n + 1, 0);
    for (int64_t q = 0; q < nl; ++q) g.lcptr[g.lsrc[q] + 1]++;
    for (int32_t v = 0; v < n; ++v) g.lcptr[v + 1] += g.lcptr[v];
    g.lcol.resize(nl);
    {
        std::vector<int64_t> bk(g.lcptr.begin(), g.lcptr.end() - 1);
        for (int64_t q = 0; q < nl; ++q) g.lcol[bk[g.lsrc[q]]++] = (int32_t)q;
    }
    g.scal.resize(3 * (size_t)n);
    for (int32_t v = 0; v < n; ++v)
        for (int c = 0; c < 3; ++c) g.scal[3 * v + c] = 3 * g.full_of[v] + c;
    return g;
}

} // namespace mfs
