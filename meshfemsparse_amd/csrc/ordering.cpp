// Nested-dissection ordering for the reduced block graph (host, once per sparsity pattern).
//
// The reference orders with CHOLMOD/METIS nested dissection (Solvers/cholmod_ordering.hh:71-96,
// ParallelNestedDissection/), which is unavailable here; any fill-reducing ND ordering is
// acceptable for the path (the factor is unique for a fixed ordering).  This is a multilevel
// bisection: heavy-edge-matching coarsening, greedy graph growing on the coarsest graph,
// boundary Fiduccia-Mattheyses refinement while uncoarsening, then a minimum vertex cover of
// the cut edges (Koenig's theorem on the bipartite boundary graph) as the vertex separator.
// Parts are ordered first, the separator last; small parts are leaves in their given order.
#include "mfs_internal.h"

#include <algorithm>
#include <climits>
#include <atomic>
#include <chrono>
#include <exception>
#include <numeric>
#include <queue>
#include <thread>
#include <tuple>

namespace mfs {

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int32_t below(int32_t n) { return (int32_t)(next() % (uint64_t)n); }
};

struct G {
    int32_t n = 0;
    std::vector<int64_t> xadj;
    std::vector<int32_t> adj, ew, vw;
    int64_t total_vw() const { return std::accumulate(vw.begin(), vw.end(), int64_t(0)); }
};

G coarsen(const G& g, std::vector<int32_t>& cmap, Rng& rng, int32_t maxvw) {
    const int32_t n = g.n;
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    for (int32_t i = n - 1; i > 0; --i) std::swap(order[i], order[rng.below(i + 1)]);
    std::vector<int32_t> match(n, -1);
    for (int32_t v : order) {
        if (match[v] != -1) continue;
        int32_t best = -1, bw = -1;
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int32_t u = g.adj[e];
            if (match[u] == -1 && u != v && g.vw[v] + g.vw[u] <= maxvw && g.ew[e] > bw) {
                best = u;
                bw = g.ew[e];
            }
        }
        if (best >= 0) { match[v] = best; match[best] = v; }
        else match[v] = v;
    }
    cmap.assign(n, -1);
    std::vector<int32_t> rep;
    rep.reserve(n / 2 + 1);
    int32_t cn = 0;
    for (int32_t v = 0; v < n; ++v)
        if (cmap[v] == -1) {
            cmap[v] = cn;
            cmap[match[v]] = cn;
            rep.push_back(v);
            ++cn;
        }
    G c;
    c.n = cn;
    c.xadj.assign(cn + 1, 0);
    c.vw.assign(cn, 0);
    c.adj.reserve(g.adj.size() / 2);
    c.ew.reserve(g.adj.size() / 2);
    std::vector<int64_t> marker(cn, -1);
    for (int32_t ci = 0; ci < cn; ++ci) {
        const int64_t start = (int64_t)c.adj.size();
        const int32_t v0 = rep[ci], v1 = match[v0];
        for (int k = 0; k < 2; ++k) {
            const int32_t u = k == 0 ? v0 : v1;
            if (k == 1 && v1 == v0) break;
            c.vw[ci] += g.vw[u];
            for (int64_t e = g.xadj[u]; e < g.xadj[u + 1]; ++e) {
                const int32_t cu = cmap[g.adj[e]];
                if (cu == ci) continue;
                const int64_t m = marker[cu];
                if (m >= start && c.adj[m] == cu) c.ew[m] += g.ew[e];
                else {
                    marker[cu] = (int64_t)c.adj.size();
                    c.adj.push_back(cu);
                    c.ew.push_back(g.ew[e]);
                }
            }
        }
        c.xadj[ci + 1] = (int64_t)c.adj.size();
    }
    return c;
}

int64_t edge_cut(const G& g, const std::vector<int8_t>& where) {
    int64_t cut = 0;
    for (int32_t v = 0; v < g.n; ++v)
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
            if (where[g.adj[e]] != where[v]) cut += g.ew[e];
    return cut / 2;
}

// Boundary FM refinement of a 2-way edge cut under a balance constraint.
void fm_refine(const G& g, std::vector<int8_t>& where, double ub, int npasses) {
    const int32_t n = g.n;
    const int64_t tw = g.total_vw();
    int64_t maxvw = 0;
    for (int32_t v = 0; v < n; ++v) maxvw = std::max<int64_t>(maxvw, g.vw[v]);
    const int64_t maxw = std::max<int64_t>((int64_t)((1.0 + ub) * 0.5 * tw), (tw + 1) / 2 + maxvw);
    std::vector<int64_t> id(n), ed(n);
    std::vector<char> locked(n);
    int64_t pw[2] = {0, 0};
    for (int32_t v = 0; v < n; ++v) pw[where[v]] += g.vw[v];
    for (int pass = 0; pass < npasses; ++pass) {
        for (int32_t v = 0; v < n; ++v) {
            id[v] = ed[v] = 0;
            locked[v] = 0;
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
                (where[g.adj[e]] == where[v] ? id[v] : ed[v]) += g.ew[e];
        }
        using Ent = std::pair<int64_t, int32_t>;
        std::priority_queue<Ent> heap[2];
        for (int32_t v = 0; v < n; ++v)
            if (ed[v] > 0) heap[where[v]].push({ed[v] - id[v], v});
        int64_t cut = edge_cut(g, where);
        const int64_t cut0 = cut;
        int64_t best_cut = cut;
        int64_t best_imb = std::abs(pw[0] - pw[1]);
        bool best_ok = pw[0] <= maxw && pw[1] <= maxw;
        std::vector<int32_t> moves;
        size_t best_len = 0;
        const size_t limit = (size_t)std::max<int32_t>(40, std::min<int32_t>(n / 20, 400));
        while (true) {
            int32_t cand[2] = {-1, -1};
            int64_t cg[2] = {0, 0};
            for (int s = 0; s < 2; ++s) {
                while (!heap[s].empty()) {
                    const auto [gn, v] = heap[s].top();
                    if (locked[v] || where[v] != s || gn != ed[v] - id[v]) { heap[s].pop(); continue; }
                    cand[s] = v;
                    cg[s] = gn;
                    break;
                }
            }
            int side = -1;
            const bool ok0 = cand[0] >= 0 && pw[1] + g.vw[cand[0]] <= maxw;
            const bool ok1 = cand[1] >= 0 && pw[0] + g.vw[cand[1]] <= maxw;
            if (pw[0] > maxw && cand[0] >= 0) side = 0;
            else if (pw[1] > maxw && cand[1] >= 0) side = 1;
            else if (ok0 && ok1) side = (cg[0] > cg[1] || (cg[0] == cg[1] && pw[0] >= pw[1])) ? 0 : 1;
            else if (ok0) side = 0;
            else if (ok1) side = 1;
            if (side < 0) break;
            const int32_t v = cand[side];
            heap[side].pop();
            locked[v] = 1;
            where[v] = (int8_t)(1 - side);
            pw[side] -= g.vw[v];
            pw[1 - side] += g.vw[v];
            cut -= ed[v] - id[v];
            std::swap(id[v], ed[v]);
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int32_t u = g.adj[e];
                if (where[u] == where[v]) { id[u] += g.ew[e]; ed[u] -= g.ew[e]; }
                else { id[u] -= g.ew[e]; ed[u] += g.ew[e]; }
                if (!locked[u] && ed[u] > 0) heap[where[u]].push({ed[u] - id[u], u});
            }
            moves.push_back(v);
            const bool ok = pw[0] <= maxw && pw[1] <= maxw;
            const int64_t imb = std::abs(pw[0] - pw[1]);
            if ((ok && !best_ok) || (ok == best_ok && (cut < best_cut || (cut == best_cut && imb < best_imb)))) {
                best_cut = cut;
                best_imb = imb;
                best_ok = ok;
                best_len = moves.size();
            }
            if (moves.size() - best_len > limit) break;
        }
        for (size_t k = moves.size(); k > best_len; --k) {
            const int32_t v = moves[k - 1];
            const int s = where[v];
            where[v] = (int8_t)(1 - s);
            pw[s] -= g.vw[v];
            pw[1 - s] += g.vw[v];
        }
        if (best_cut >= cut0 && best_len == 0) break;
    }
}

// Greedy graph growing bisection of a (small) coarse graph; best of several trials.
std::vector<int8_t> initial_bisection(const G& g, Rng& rng, double ub) {
    const int32_t n = g.n;
    const int64_t tw = g.total_vw();
    std::vector<int8_t> best;
    int64_t best_cut = -1;
    const int trials = 8;
    for (int t = 0; t < trials; ++t) {
        std::vector<int8_t> where(n, 1);
        std::vector<int64_t> conn0(n, 0);   // edge weight to part 0
        std::vector<int64_t> deg(n, 0);
        for (int32_t v = 0; v < n; ++v)
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) deg[v] += g.ew[e];
        int64_t w0 = 0;
        auto move = [&](int32_t v) {
            where[v] = 0;
            w0 += g.vw[v];
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) conn0[g.adj[e]] += g.ew[e];
        };
        move(rng.below(n));
        while (2 * w0 < tw) {
            int32_t bv = -1;
            int64_t bg = INT64_MIN;
            for (int32_t v = 0; v < n; ++v) {
                if (where[v] != 1 || conn0[v] == 0) continue;
                const int64_t gain = 2 * conn0[v] - deg[v];
                if (gain > bg) { bg = gain; bv = v; }
            }
            if (bv < 0) {   // disconnected: restart growth from a random vertex of part 1
                std::vector<int32_t> rest;
                for (int32_t v = 0; v < n; ++v) if (where[v] == 1) rest.push_back(v);
                if (rest.empty()) break;
                bv = rest[rng.below((int32_t)rest.size())];
            }
            move(bv);
        }
        fm_refine(g, where, ub, 4);
        const int64_t cut = edge_cut(g, where);
        if (best_cut < 0 || cut < best_cut) { best_cut = cut; best = where; }
    }
    return best;
}

// Minimum vertex cover of the cut edges -> separator (where = 2).
void edge_to_vertex_separator(const G& g, std::vector<int8_t>& where) {
    const int32_t n = g.n;
    std::vector<int32_t> bidx(n, -1);
    std::vector<int32_t> L, R;
    for (int32_t v = 0; v < n; ++v) {
        bool bnd = false;
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1] && !bnd; ++e) bnd = where[g.adj[e]] != where[v];
        if (!bnd) continue;
        if (where[v] == 0) { bidx[v] = (int32_t)L.size(); L.push_back(v); }
        else { bidx[v] = (int32_t)R.size(); R.push_back(v); }
    }
    if (L.empty()) return;
    // Maximum bipartite matching L-R over cut edges (greedy + augmenting DFS).
    std::vector<int32_t> mL(L.size(), -1), mR(R.size(), -1);
    for (size_t i = 0; i < L.size(); ++i) {
        const int32_t v = L[i];
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int32_t u = g.adj[e];
            if (where[u] == 1 && mR[bidx[u]] == -1) { mL[i] = bidx[u]; mR[bidx[u]] = (int32_t)i; break; }
        }
    }
    std::vector<int32_t> stamp(L.size(), -1);
    std::vector<std::pair<int32_t, int64_t>> stack;
    for (size_t i0 = 0; i0 < L.size(); ++i0) {
        if (mL[i0] != -1) continue;
        // iterative DFS for an augmenting path from L[i0]
        stack.clear();
        stack.push_back({(int32_t)i0, g.xadj[L[i0]]});
        stamp[i0] = (int32_t)i0;
        std::vector<int32_t> pathR;
        bool found = false;
        while (!stack.empty() && !found) {
            auto& [li, ep] = stack.back();
            const int32_t v = L[li];
            bool advanced = false;
            while (ep < g.xadj[v + 1]) {
                const int32_t u = g.adj[ep++];
                if (where[u] != 1) continue;
                const int32_t ri = bidx[u];
                const int32_t lj = mR[ri];
                if (lj == -1) {
                    pathR.push_back(ri);
                    found = true;
                    advanced = true;
                    break;
                }
                if (stamp[lj] == (int32_t)i0) continue;
                stamp[lj] = (int32_t)i0;
                pathR.push_back(ri);
                stack.push_back({lj, g.xadj[L[lj]]});
                advanced = true;
                break;
            }
            if (!advanced) {
                stack.pop_back();
                if (!pathR.empty() && !stack.empty()) pathR.pop_back();
            }
        }
        if (found) {
            // stack holds L vertices of the path, pathR the R vertices (same length)
            for (size_t k = 0; k < stack.size(); ++k) {
                const int32_t li = stack[k].first, ri = pathR[k];
                mL[li] = ri;
                mR[ri] = li;
            }
        }
    }
    // Koenig: Z = reachable from unmatched L via alternating paths; cover = (L \ Z) u (R n Z)
    std::vector<char> zL(L.size(), 0), zR(R.size(), 0);
    std::vector<int32_t> q;
    for (size_t i = 0; i < L.size(); ++i)
        if (mL[i] == -1) { zL[i] = 1; q.push_back((int32_t)i); }
    for (size_t h = 0; h < q.size(); ++h) {
        const int32_t v = L[q[h]];
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
            const int32_t u = g.adj[e];
            if (where[u] != 1) continue;
            const int32_t ri = bidx[u];
            if (zR[ri]) continue;
            zR[ri] = 1;
            const int32_t lj = mR[ri];
            if (lj >= 0 && !zL[lj]) { zL[lj] = 1; q.push_back(lj); }
        }
    }
    for (size_t i = 0; i < L.size(); ++i) if (!zL[i]) where[L[i]] = 2;
    for (size_t i = 0; i < R.size(); ++i) if (zR[i]) where[R[i]] = 2;
}

// Node-based FM refinement of a vertex separator (where: 0/1 sides, 2 separator).  A move takes a
// separator vertex into side s and pulls its neighbours on the other side into the separator
// (gain = own weight - pulled weight), so every state is a valid separator.  Moves go best gain
// first, each vertex at most once per pass, hill-climbing through non-improving moves; the pass
// then rolls back to the smallest separator it saw (ties: better balance).  Sides stay within
// the bisection's balance bound.
void node_fm_refine(const G& g, std::vector<int8_t>& where, double ub, int npasses) {
    const int32_t n = g.n;
    const int64_t tw = g.total_vw();
    int64_t maxvw = 0;
    for (int32_t v = 0; v < n; ++v) maxvw = std::max<int64_t>(maxvw, g.vw[v]);
    const int64_t maxw = std::max<int64_t>((int64_t)((1.0 + ub) * 0.5 * tw), (tw + 1) / 2 + maxvw);
    const int limit = (int)std::min<int64_t>(300, std::max<int64_t>(25, n / 50));
    auto gain = [&](int32_t v, int s) {
        int64_t pulled = 0;
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
            if (where[g.adj[e]] == 1 - s) pulled += g.vw[g.adj[e]];
        return (int64_t)g.vw[v] - pulled;
    };
    std::vector<char> locked(n);
    std::vector<std::pair<int32_t, int8_t>> log;   // (vertex, previous side)
    for (int pass = 0; pass < npasses; ++pass) {
        int64_t pw[3] = {0, 0, 0};
        for (int32_t v = 0; v < n; ++v) pw[where[v]] += g.vw[v];
        const int64_t sep0 = pw[2];
        int64_t best_sep = pw[2], best_bal = std::abs(pw[0] - pw[1]);
        size_t best_len = 0;
        std::fill(locked.begin(), locked.end(), 0);
        log.clear();
        using Ent = std::tuple<int64_t, int32_t, int8_t>;   // gain, vertex, side
        std::priority_queue<Ent> pq;
        auto push = [&](int32_t v) {
            pq.push({gain(v, 0), v, 0});
            pq.push({gain(v, 1), v, 1});
        };
        for (int32_t v = 0; v < n; ++v)
            if (where[v] == 2) push(v);
        int nobetter = 0;
        while (!pq.empty() && nobetter < limit) {
            const auto [gn, v, sd] = pq.top();
            pq.pop();
            if (where[v] != 2 || locked[v]) continue;
            const int s = sd;
            const int64_t gv = gain(v, s);
            if (gv != gn) { pq.push({gv, v, sd}); continue; }     // stale entry
            if (pw[s] + g.vw[v] > maxw) continue;
            log.push_back({v, 2});
            where[v] = (int8_t)s;
            locked[v] = 1;
            pw[s] += g.vw[v];
            pw[2] -= g.vw[v];
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int32_t u = g.adj[e];
                if (where[u] == 1 - s) {
                    log.push_back({u, (int8_t)(1 - s)});
                    where[u] = 2;
                    pw[1 - s] -= g.vw[u];
                    pw[2] += g.vw[u];
                }
            }
            // gains of separator vertices next to v or to the vertices it pulled have changed
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
                const int32_t u = g.adj[e];
                if (where[u] != 2 || locked[u]) continue;
                push(u);
                for (int64_t f = g.xadj[u]; f < g.xadj[u + 1]; ++f) {
                    const int32_t w = g.adj[f];
                    if (where[w] == 2 && !locked[w]) push(w);
                }
            }
            const int64_t bal = std::abs(pw[0] - pw[1]);
            if (pw[2] < best_sep || (pw[2] == best_sep && bal < best_bal)) {
                best_sep = pw[2];
                best_bal = bal;
                best_len = log.size();
                nobetter = 0;
            } else {
                ++nobetter;
            }
        }
        for (size_t i = log.size(); i > best_len; --i) where[log[i - 1].first] = log[i - 1].second;
        if (best_sep >= sep0) break;
    }
}

std::vector<int8_t> bisect(const G& g0, Rng& rng, const NDOptions& opt) {
    std::vector<G> graphs;
    std::vector<std::vector<int32_t>> cmaps;
    graphs.push_back(g0);
    const int64_t tw = g0.total_vw();
    const int32_t maxvw = (int32_t)std::max<int64_t>(1, (int64_t)(1.5 * tw / opt.coarse_size));
    while (graphs.back().n > opt.coarse_size) {
        std::vector<int32_t> cmap;
        G c = coarsen(graphs.back(), cmap, rng, maxvw);
        if (c.n > 0.92 * graphs.back().n) break;
        cmaps.push_back(std::move(cmap));
        graphs.push_back(std::move(c));
    }
    std::vector<int8_t> where = initial_bisection(graphs.back(), rng, opt.imbalance);
    for (int lvl = (int)graphs.size() - 2; lvl >= 0; --lvl) {
        const auto& cm = cmaps[lvl];
        std::vector<int8_t> fw(graphs[lvl].n);
        for (int32_t v = 0; v < graphs[lvl].n; ++v) fw[v] = where[cm[v]];
        where.swap(fw);
        fm_refine(graphs[lvl], where, opt.imbalance, 6);
    }
    edge_to_vertex_separator(g0, where);
    if (opt.node_passes > 0) node_fm_refine(g0, where, opt.imbalance, opt.node_passes);
    return where;
}

G induced(const G& g, const std::vector<int8_t>& where, int8_t part, std::vector<int32_t>& local,
          std::vector<int32_t>& members) {
    members.clear();
    for (int32_t v = 0; v < g.n; ++v)
        if (where[v] == part) { local[v] = (int32_t)members.size(); members.push_back(v); }
    G s;
    s.n = (int32_t)members.size();
    s.xadj.assign(s.n + 1, 0);
    s.vw.assign(s.n, 1);
    for (int32_t i = 0; i < s.n; ++i) {
        const int32_t v = members[i];
        for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
            if (where[g.adj[e]] == part) { s.adj.push_back(local[g.adj[e]]); s.ew.push_back(1); }
        s.xadj[i + 1] = (int64_t)s.adj.size();
    }
    return s;
}

// Connected components (label per vertex, components numbered by their smallest vertex).
int32_t components(const G& g, std::vector<int32_t>& comp) {
    comp.assign(g.n, -1);
    int32_t nc = 0;
    std::vector<int32_t> stack;
    for (int32_t v0 = 0; v0 < g.n; ++v0) {
        if (comp[v0] != -1) continue;
        comp[v0] = nc;
        stack.push_back(v0);
        while (!stack.empty()) {
            const int32_t v = stack.back();
            stack.pop_back();
            for (int64_t e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
                if (comp[g.adj[e]] == -1) { comp[g.adj[e]] = nc; stack.push_back(g.adj[e]); }
        }
        ++nc;
    }
    return nc;
}

// Separator forest under construction: components get ids as they are closed, so children
// always precede their parent (postorder).  A component owns either the vertices of a leaf or
// the vertices of a separator; disconnected parts add no component of their own.
struct ForestBuilder {
    std::vector<int32_t> parent;
    std::vector<std::vector<int32_t>> children, verts;
    int32_t add(std::vector<int32_t> own, std::vector<int32_t> kids) {
        const int32_t id = (int32_t)parent.size();
        for (int32_t c : kids) parent[c] = id;
        parent.push_back(-1);
        children.push_back(std::move(kids));
        verts.push_back(std::move(own));
        return id;
    }
};

std::vector<int32_t> nd_recurse(const G& g, const std::vector<int32_t>& gid, Rng& rng, const NDOptions& opt,
                                ForestBuilder& fb);

// Components are ordered one after another (each by its own dissection): no separator is
// needed between parts that share no edge.  Returns false for a connected graph.
bool nd_components(const G& g, const std::vector<int32_t>& gid, Rng& rng, const NDOptions& opt,
                   ForestBuilder& fb, std::vector<int32_t>& roots) {
    std::vector<int32_t> comp;
    const int32_t nc = components(g, comp);
    if (nc <= 1) return false;
    std::vector<int32_t> local(g.n, -1), members;
    std::vector<int8_t> where(g.n);
    for (int32_t c = 0; c < nc; ++c) {
        for (int32_t v = 0; v < g.n; ++v) where[v] = comp[v] == c ? 0 : 1;
        G s = induced(g, where, 0, local, members);
        std::vector<int32_t> sg(members.size());
        for (size_t i = 0; i < members.size(); ++i) sg[i] = gid[members[i]];
        const std::vector<int32_t> r = nd_recurse(s, sg, rng, opt, fb);
        roots.insert(roots.end(), r.begin(), r.end());
    }
    return true;
}

// Dissects g (vertex v is gid[v] globally); returns the roots of its separator forest in order.
std::vector<int32_t> nd_recurse(const G& g, const std::vector<int32_t>& gid, Rng& rng, const NDOptions& opt,
                                ForestBuilder& fb) {
    if (g.n == 0) return {};
    if (g.n <= opt.leaf_size) return {fb.add(gid, {})};
    std::vector<int32_t> roots;
    if (nd_components(g, gid, rng, opt, fb, roots)) return roots;
    // opt.nseps bisections; the smallest vertex separator that splits the graph is kept
    auto sep_size = [](const std::vector<int8_t>& w) {
        int64_t c[3] = {0, 0, 0};
        for (int8_t x : w) c[x]++;
        return (c[0] == 0 || c[1] == 0) ? INT64_MAX : c[2];
    };
    std::vector<int8_t> where = bisect(g, rng, opt);
    int64_t best = sep_size(where);
    for (int t = 1; t < opt.nseps; ++t) {
        std::vector<int8_t> w = bisect(g, rng, opt);
        const int64_t sz = sep_size(w);
        if (sz < best) { best = sz; where.swap(w); }
    }
    if (best == INT64_MAX) return {fb.add(gid, {})};    // could not split: leaf
    std::vector<int32_t> local(g.n, -1), members, kids;
    for (int8_t part = 0; part < 2; ++part) {
        G s = induced(g, where, part, local, members);
        std::vector<int32_t> sg(members.size());
        for (size_t i = 0; i < members.size(); ++i) sg[i] = gid[members[i]];
        const std::vector<int32_t> r = nd_recurse(s, sg, rng, opt, fb);
        kids.insert(kids.end(), r.begin(), r.end());
    }
    std::vector<int32_t> sep;
    for (int32_t v = 0; v < g.n; ++v)
        if (where[v] == 2) sep.push_back(gid[v]);
    if (sep.empty()) return kids;
    return {fb.add(std::move(sep), std::move(kids))};
}

// Dissection of the subgraph induced by `verts` (global ids, in this order) as a forest whose
// vertex lists hold global ids.  Every call starts from the same seed.  `local` is scratch of n
// entries, -1 for every vertex of the region and its neighbours; the call sets the region's
// entries and restores them to -1 (calls on disjoint regions whose vertices have no neighbour in
// another running call's region may share it).
ForestBuilder dissect_region(const std::vector<int64_t>& xadj, const std::vector<int32_t>& adj,
                             const std::vector<int32_t>& verts, const NDOptions& opt,
                             std::vector<int32_t>& roots, std::vector<int32_t>& local) {
    for (size_t i = 0; i < verts.size(); ++i) local[verts[i]] = (int32_t)i;
    G s;
    s.n = (int32_t)verts.size();
    s.xadj.assign(s.n + 1, 0);
    s.vw.assign(s.n, 1);
    for (int32_t i = 0; i < s.n; ++i) {
        const int32_t v = verts[i];
        for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e)
            if (local[adj[e]] >= 0 && adj[e] != v) { s.adj.push_back(local[adj[e]]); s.ew.push_back(1); }
        s.xadj[i + 1] = (int64_t)s.adj.size();
    }
    for (int32_t v : verts) local[v] = -1;
    ForestBuilder fb;
    Rng rng{opt.seed};
    roots = nd_recurse(s, verts, rng, opt, fb);
    return fb;
}

// Appends the ordering of the subtree of c (children first, in order, then c's own vertices).
template <class Forest>
void flatten(const Forest& f, int32_t c, std::vector<int32_t>& out) {
    std::vector<std::pair<int32_t, size_t>> st{{c, 0}};
    while (!st.empty()) {
        auto& [node, k] = st.back();
        if (k < f.children[node].size()) {
            const int32_t ch = f.children[node][k++];
            st.push_back({ch, 0});
        } else {
            out.insert(out.end(), f.verts[node].begin(), f.verts[node].end());
            st.pop_back();
        }
    }
}

// Renumbers the live components reachable from the roots in postorder and rebuilds member.
SepForest compact(const SepForest& f, int32_t n) {
    SepForest o;
    std::vector<int32_t> id(f.parent.size(), -1);
    std::vector<std::pair<int32_t, size_t>> st;
    for (int32_t r : f.roots) {
        st.push_back({r, 0});
        while (!st.empty()) {
            auto& [node, k] = st.back();
            if (k < f.children[node].size()) {
                const int32_t ch = f.children[node][k++];
                st.push_back({ch, 0});
            } else {
                id[node] = (int32_t)o.parent.size();
                o.parent.push_back(-1);
                o.verts.push_back(f.verts[node]);
                std::vector<int32_t> kids;
                for (int32_t ch : f.children[node]) kids.push_back(id[ch]);
                for (int32_t ch : kids) o.parent[ch] = id[node];
                o.children.push_back(std::move(kids));
                st.pop_back();
            }
        }
        o.roots.push_back(id[r]);
    }
    o.member.assign(n, -1);
    for (size_t c = 0; c < o.verts.size(); ++c)
        for (int32_t v : o.verts[c]) o.member[v] = (int32_t)c;
    for (int32_t v = 0; v < n; ++v)
        if (o.member[v] < 0) throw Error(-1, "separator forest does not cover every vertex");
    return o;
}

// Appends the builder's components to f (ids offset by the current size); returns the offset.
int32_t append(SepForest& f, const ForestBuilder& fb) {
    const int32_t off = (int32_t)f.parent.size();
    for (size_t c = 0; c < fb.parent.size(); ++c) {
        f.parent.push_back(fb.parent[c] < 0 ? -1 : fb.parent[c] + off);
        std::vector<int32_t> kids = fb.children[c];
        for (int32_t& k : kids) k += off;
        f.children.push_back(std::move(kids));
        f.verts.push_back(fb.verts[c]);
    }
    return off;
}

// Options of dissection trial t (trial 0: the options themselves).
NDOptions trial_options(const NDOptions& opt, int t) {
    NDOptions o = opt;
    if (t & 1) o.leaf_size = opt.leaf_alt;
    if (t & 2) o.imbalance = opt.imbalance_alt;
    o.seed = opt.seed + (uint64_t)(t / 4);
    return o;
}

// Full dissection: every connected component dissected opt.trials times (trial options above),
// keeping the trial with the fewest predicted factor flops (the first on ties); the (component,
// trial) pairs run in parallel and the kept components are concatenated in component order.
SepForest full_dissection(int32_t n, const std::vector<int64_t>& xadj, const std::vector<int32_t>& adj,
                          const NDOptions& opt) {
    G g;
    g.n = n;
    g.xadj = xadj;
    g.adj = adj;
    std::vector<int32_t> comp;
    const int32_t nc = components(g, comp);
    std::vector<std::vector<int32_t>> gids(nc);
    for (int32_t v = 0; v < n; ++v) gids[comp[v]].push_back(v);
    const int K = std::max(1, opt.trials);
    const int32_t small = std::min(opt.leaf_size, K > 1 ? opt.leaf_alt : opt.leaf_size);
    std::vector<std::pair<int32_t, int>> items;     // (component, trial)
    std::vector<int> ntrial(nc);
    for (int32_t c = 0; c < nc; ++c) {
        ntrial[c] = (int32_t)gids[c].size() > small ? K : 1;   // a single leaf needs no trials
        for (int t = 0; t < ntrial[c]; ++t) items.push_back({c, t});
    }
    // largest components first, so that no large dissection starts last (results do not depend
    // on the order: every item is independent)
    std::stable_sort(items.begin(), items.end(), [&](const auto& a, const auto& b) {
        return gids[a.first].size() > gids[b.first].size();
    });
    std::vector<std::vector<ForestBuilder>> fbs(nc);
    std::vector<std::vector<std::vector<int32_t>>> roots(nc);
    std::vector<std::vector<double>> flops(nc);
    for (int32_t c = 0; c < nc; ++c) {
        fbs[c].resize(ntrial[c]);
        roots[c].resize(ntrial[c]);
        flops[c].assign(ntrial[c], 0.0);
    }
    const int nt = std::max(1, std::min<int>({(int)std::thread::hardware_concurrency(), 16, (int)items.size()}));
    std::vector<std::vector<int32_t>> scratch(nt);  // per-thread vertex maps (trials of one
                                                    // component overlap in their vertices)
    auto work = [&](int tid, size_t q) {
        std::vector<int32_t>& loc = scratch[tid];
        if (loc.empty()) loc.assign(n, -1);
        const auto [c, t] = items[q];
        fbs[c][t] = dissect_region(xadj, adj, gids[c], trial_options(opt, t), roots[c][t], loc);
        if (ntrial[c] > 1) {
            std::vector<int32_t> ord;
            for (int32_t r : roots[c][t]) flatten(fbs[c][t], r, ord);
            flops[c][t] = order_flops(xadj, adj, ord, loc);
        }
    };
    if (nt > 1) {
        std::atomic<size_t> next{0};
        std::vector<std::exception_ptr> err(nt);      // a worker's first exception, rethrown here
        std::vector<std::thread> th;
        for (int tid = 0; tid < nt; ++tid)
            th.emplace_back([&, tid] {
                try {
                    for (size_t q; (q = next.fetch_add(1)) < items.size();) work(tid, q);
                } catch (...) {
                    err[tid] = std::current_exception();
                    next.store(items.size());           // the other workers stop early
                }
            });
        for (auto& t : th) t.join();
        for (const auto& e : err)
            if (e) std::rethrow_exception(e);
    } else {
        for (size_t q = 0; q < items.size(); ++q) work(0, q);
    }
    SepForest f;
    for (int32_t c = 0; c < nc; ++c) {
        int best = 0;
        for (int t = 1; t < ntrial[c]; ++t)
            if (flops[c][t] < flops[c][best]) best = t;
        const int32_t off = append(f, fbs[c][best]);
        for (int32_t r : roots[c][best]) f.roots.push_back(r + off);
    }
    return compact(f, n);
}

} // namespace

std::vector<int32_t> forest_order(const SepForest& f) {
    std::vector<int32_t> out;
    for (int32_t r : f.roots) flatten(f, r, out);
    return out;
}

// Temporal reuse (ParallelNestedDissection/persistent_separator_tree.hh:81-113 and
// cholmod_nesdis_parallel.cc:1391-1506): an edge whose end components are not on one root path
// violates the separator of their lowest common ancestor; the outermost violated components are
// re-dissected on the current graph and spliced in place, everything else keeps its ordering.
// An edge between two trees of the forest, a violation at the sole root, no (or a mismatching)
// previous forest, or `period` incremental analyses since the last full one: full dissection.
std::vector<int32_t> nested_dissection(int32_t n, const std::vector<int64_t>& xadj,
                                       const std::vector<int32_t>& adj, const NDOptions& opt,
                                       NDState* state) {
    const auto t0 = std::chrono::steady_clock::now();
    NDReuseStats st;
    st.graph_vertices = n;
    SepForest next;
    bool full = !state || state->period <= 0 || state->forest.member.size() != (size_t)n ||
                state->since_rebuild >= state->period || n == 0;
    std::vector<int32_t> dirty_roots;
    if (!full) {
        const SepForest& f = state->forest;
        const int32_t nc = (int32_t)f.parent.size();
        std::vector<int32_t> depth(nc, 0);
        for (int32_t c = nc - 1; c >= 0; --c)          // parents have larger ids
            depth[c] = f.parent[c] < 0 ? 0 : depth[f.parent[c]] + 1;
        auto lca = [&](int32_t a, int32_t b) {
            while (depth[a] > depth[b]) a = f.parent[a];
            while (depth[b] > depth[a]) b = f.parent[b];
            while (a != b && a >= 0 && b >= 0) { a = f.parent[a]; b = f.parent[b]; }
            return a == b ? a : -1;
        };
        std::vector<char> dirty(nc, 0);
        for (int32_t v = 0; v < n && !full; ++v)
            for (int64_t e = xadj[v]; e < xadj[v + 1]; ++e) {
                const int32_t u = adj[e];
                if (u >= v) continue;
                const int32_t a = f.member[v], b = f.member[u], r = lca(a, b);
                if (r == a || r == b) continue;
                if (r < 0 || (f.parent[r] < 0 && f.roots.size() == 1)) { full = true; break; }
                dirty[r] = 1;
            }
        if (!full)
            for (int32_t c = 0; c < nc; ++c)
                if (dirty[c]) {
                    int32_t p = f.parent[c];
                    while (p >= 0 && !dirty[p]) p = f.parent[p];
                    if (p < 0) dirty_roots.push_back(c);
                }
    }
    if (full) {
        next = full_dissection(n, xadj, adj, opt);
        st.full_rebuild = true;
        st.repartitioned_vertices = n;
    } else if (dirty_roots.empty()) {
        next = state->forest;
    } else {
        SepForest f = state->forest;
        std::vector<int32_t> local(n, -1);
        for (int32_t d : dirty_roots) {
            std::vector<int32_t> region;
            flatten(f, d, region);                 // the region's previous ordering
            st.repartitioned_vertices += (int64_t)region.size();
            std::vector<int32_t> roots;
            const ForestBuilder fb = dissect_region(xadj, adj, region, opt, roots, local);
            const int32_t off = append(f, fb);
            for (int32_t& r : roots) r += off;
            std::vector<int32_t>& slot = f.parent[d] < 0 ? f.roots : f.children[f.parent[d]];
            const auto it = std::find(slot.begin(), slot.end(), d);
            const size_t pos = (size_t)(it - slot.begin());
            slot.erase(it);
            slot.insert(slot.begin() + pos, roots.begin(), roots.end());
            for (int32_t r : roots) f.parent[r] = f.parent[d];
        }
        next = compact(f, n);                       // drops the replaced subtrees
        st.recomputed_subtrees = (int64_t)dirty_roots.size();
    }
    std::vector<int32_t> perm = forest_order(next);
    st.nd_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (state) {
        state->since_rebuild = st.full_rebuild ? 0 : state->since_rebuild + 1;
        state->forest = std::move(next);
        state->stats = st;
    }
    return perm;
}

std::vector<int32_t> nested_dissection(int32_t n, const std::vector<int64_t>& xadj,
                                       const std::vector<int32_t>& adj, const NDOptions& opt,
                                       SepForest* forest) {
    NDState s;
    std::vector<int32_t> perm = nested_dissection(n, xadj, adj, opt, &s);
    if (forest) *forest = std::move(s.forest);
    return perm;
}

} // namespace mfs
