// Block sparsity pattern construction and the assembly slot map.
//
// Pattern: the reference (SystemAssembler.hh:251-465) sorts each element's block variables,
// appends rows bvars[0..i] to the bucket of column bvars[i], deduplicates and sorts each bucket.
// The pattern is a set, so any construction yielding the sorted, deduplicated upper-triangle
// rows per column gives identical Ap/Ai; this one counts, fills and sort-uniques per column.
//
// Slot map: the reference assembler (SystemAssembler.hh:60-97) visits, per element, the pairs of
// local block variables in sorted global order and adds He(lbi,lbj) (or He(lbj,lbi)^T when the
// local order is reversed) to the slot of global block (bi,bj), bi < bj, plus the upper triangle
// of He(lbj,lbj) to the diagonal slot.  For a local pair (a,b), a<b, that is always the local-
// upper block (a,b) of He, transposed iff nodes[a] > nodes[b].  The map lists these
// contributions per destination slot in ascending element order, which is the order a serial
// run of the reference applies them in.
#include "mfs_internal.h"

#include <algorithm>

namespace mfs {

Pattern build_block_pattern(const int32_t* en, int64_t ne, int npe, int64_t nb) {
    if (npe <= 0 || npe > 16) throw Error(-1, "nodes_per_elem must be in 1..16");
    Pattern P;
    P.n = nb;
    std::vector<int64_t> cnt(nb + 1, 0);
    int32_t v[16];
    for (int64_t e = 0; e < ne; ++e) {
        for (int a = 0; a < npe; ++a) {
            v[a] = en[npe * e + a];
            if (v[a] < 0 || v[a] >= nb) throw Error(-1, "element node index out of range");
        }
        std::sort(v, v + npe);
        for (int i = 0; i < npe; ++i) cnt[v[i] + 1] += i + 1;
    }
    for (int64_t j = 0; j < nb; ++j) cnt[j + 1] += cnt[j];
    std::vector<int32_t> bucket(cnt[nb]);
    std::vector<int64_t> back(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e) {
        for (int a = 0; a < npe; ++a) v[a] = en[npe * e + a];
        std::sort(v, v + npe);
        for (int i = 0; i < npe; ++i)
            for (int k = 0; k <= i; ++k) bucket[back[v[i]]++] = v[k];
    }
    P.Ap.assign(nb + 1, 0);
    std::vector<int64_t> newsz(nb);
    for (int64_t j = 0; j < nb; ++j) {
        auto b0 = bucket.begin() + cnt[j], b1 = bucket.begin() + cnt[j + 1];
        std::sort(b0, b1);
        auto e1 = std::unique(b0, b1);
        newsz[j] = e1 - b0;
    }
    for (int64_t j = 0; j < nb; ++j) P.Ap[j + 1] = P.Ap[j] + newsz[j];
    P.Ai.resize(P.Ap[nb]);
    for (int64_t j = 0; j < nb; ++j)
        std::copy(bucket.begin() + cnt[j], bucket.begin() + cnt[j] + newsz[j], P.Ai.begin() + P.Ap[j]);
    return P;
}

namespace {
inline int64_t find_slot(const Pattern& P, int64_t bi, int64_t bj) {
    const int64_t* b = P.Ai.data() + P.Ap[bj];
    const int64_t* e = P.Ai.data() + P.Ap[bj + 1];
    const int64_t* it = std::lower_bound(b, e, bi);
    if (it == e || *it != bi)
        throw Error(-1, "element block pair (" + std::to_string(bi) + "," + std::to_string(bj) +
                            ") missing from the sparsity pattern");
    return it - P.Ai.data();
}
const int kPairA[6] = {0, 0, 0, 1, 1, 2};
const int kPairB[6] = {1, 2, 3, 2, 3, 3};
} // namespace

SlotMap build_slot_map(const Pattern& P, const int32_t* en, int64_t ne, int npe) {
    if (npe != 4) throw Error(-6, "assembly slot maps are built for 4-node (tet) elements");
    if (ne >= (int64_t(1) << 27)) throw Error(-6, "too many elements for the slot-map encoding");
    SlotMap S;
    S.nslots = P.nz();
    if (S.nslots >= (int64_t(1) << 31)) throw Error(-6, "pattern too large for 32-bit slot map");
    // processing order of the slots
    std::vector<uint8_t> diag_of_slot(S.nslots, 0);
    for (int64_t j = 0; j < P.n; ++j)
        for (int64_t k = P.Ap[j]; k < P.Ap[j + 1]; ++k)
            if (P.Ai[k] == j) diag_of_slot[k] = 1;
    S.slot_of_pos.resize(S.nslots);
    std::vector<int32_t> pos_of_slot(S.nslots);
    {
        int64_t q = 0;
        for (int64_t g = 0; g < P.n; g += 256) {
            const int64_t g1 = std::min<int64_t>(P.n, g + 256);
            for (int d = 0; d < 2; ++d)
                for (int64_t j = g; j < g1; ++j) {
                    for (int64_t k = P.Ap[j]; k < P.Ap[j + 1]; ++k)
                        if (diag_of_slot[k] == d) {
                            pos_of_slot[k] = (int32_t)q;
                            S.slot_of_pos[q++] = (int32_t)k;
                        }
                }
        }
    }
    std::vector<int64_t> slot_of((size_t)ne * 10);
    std::vector<int32_t> cnt(S.nslots + 1, 0);
    for (int64_t e = 0; e < ne; ++e) {
        const int32_t* v = en + 4 * e;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b)
                if (v[a] == v[b]) throw Error(-6, "element with a repeated node");
        for (int p = 0; p < 6; ++p) {
            const int64_t na = v[kPairA[p]], nbb = v[kPairB[p]];
            const int64_t s = find_slot(P, std::min(na, nbb), std::max(na, nbb));
            slot_of[10 * e + p] = s;
            cnt[pos_of_slot[s] + 1]++;
        }
        for (int a = 0; a < 4; ++a) {
            const int64_t s = find_slot(P, v[a], v[a]);
            slot_of[10 * e + 6 + a] = s;
            cnt[pos_of_slot[s] + 1]++;
        }
    }
    for (int64_t s = 0; s < S.nslots; ++s) cnt[s + 1] += cnt[s];
    S.seg = cnt;
    S.contrib.resize(cnt[S.nslots]);
    std::vector<int32_t> back(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; ++e) {
        const int32_t* v = en + 4 * e;
        for (int p = 0; p < 10; ++p) {
            const int64_t s = slot_of[10 * e + p];
            uint32_t tr = 0;
            if (p < 6) tr = v[kPairA[p]] > v[kPairB[p]] ? 1u : 0u;
            S.contrib[back[pos_of_slot[s]]++] = ((uint32_t)e << 5) | ((uint32_t)p << 1) | tr;
        }
    }
    S.is_diag.resize(S.nslots);
    for (int64_t q = 0; q < S.nslots; ++q) S.is_diag[q] = diag_of_slot[S.slot_of_pos[q]];
    return S;
}

} // namespace mfs
