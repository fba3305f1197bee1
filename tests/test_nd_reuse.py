"""Nested-dissection separator tree and its temporal reuse (host, native library): the forest is a
valid separator tree of the graph it ordered and the ordering is its postorder flattening; reuse
decisions, repaired regions and kept orderings follow the reference's rules
(ParallelNestedDissection/persistent_separator_tree.hh:63-113, cholmod_nesdis_parallel.cc:
1391-1506) as restated in tests/nd_oracle.py."""
import numpy as np
import pytest
import scipy.sparse as sp

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl
from tests import nd_oracle as orc


def cube_graph(N, extra=()):
    tets, X = wl.cube_mesh(N)
    H = mfs.SystemAssembler(len(X)).blockSparsityPattern(tets)
    return block_graph(H.n, H.Ap, H.Ai, extra)


def block_graph(n, Ap, Ai, extra=()):
    cols = np.repeat(np.arange(n), np.diff(Ap))
    rows = np.asarray(Ai)
    off = rows != cols
    r = np.concatenate([rows[off], cols[off], [a for a, b in extra], [b for a, b in extra]]).astype(np.int64)
    c = np.concatenate([cols[off], rows[off], [b for a, b in extra], [a for a, b in extra]]).astype(np.int64)
    G = sp.csc_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    G.sum_duplicates()
    G.sort_indices()
    return G.indptr.astype(np.int64), G.indices.astype(np.int64)


def check_forest(nd, perm, p, i):
    n = len(p) - 1
    o = nd.ndOrdering()
    par, mem = o.CParent.tolist(), o.CMember.tolist()
    assert sorted(perm.tolist()) == list(range(n))
    assert all(par[c] == -1 or par[c] > c for c in range(len(par)))          # postordered
    assert set(mem) == set(range(len(par)))                                   # no empty component
    assert (np.diff(o.CMember[perm]) >= 0).all()                              # ordering = flattening
    assert orc.is_separator_tree(par, mem, orc.edges_of(p, i))
    return par, mem


@pytest.mark.parametrize("N", [3, 6, 10])
def test_forest_is_a_valid_separator_tree(N):
    p, i = cube_graph(N)
    nd = mfs.NestedDissection()
    perm = nd.order(p, i)
    par, mem = check_forest(nd, perm, p, i)
    st = nd.temporalReuseStatistics()
    assert st.full_rebuild and st.graph_vertices == len(p) - 1 and st.repartitioned_vertices == len(p) - 1
    assert sum(1 for c in par if c < 0) == 1                                  # one connected mesh


def test_disconnected_graph_is_a_forest_of_its_components():
    p1, i1 = cube_graph(5)
    n1 = len(p1) - 1
    A = sp.csc_matrix((np.ones(len(i1)), i1, p1), shape=(n1, n1))
    G = sp.block_diag([A, A, sp.csc_matrix((3, 3))], format="csc")          # 3 isolated vertices
    G.sort_indices()
    p, i = G.indptr.astype(np.int64), G.indices.astype(np.int64)
    nd = mfs.NestedDissection()
    perm = nd.order(p, i)
    par, mem = check_forest(nd, perm, p, i)
    roots = [c for c in range(len(par)) if par[c] < 0]
    assert len(roots) == 5                    # two meshes and three isolated vertices
    # each mesh is ordered exactly as it is on its own
    alone = mfs.NestedDissection().order(p1, i1)
    assert (perm[:n1] == alone).all() and (perm[n1:2 * n1] == alone + n1).all()


def test_empty_and_tiny_graphs():
    nd = mfs.NestedDissection(3)
    assert len(nd.order(np.zeros(1, np.int64), np.zeros(0, np.int64))) == 0
    perm = nd.order(np.array([0, 0], np.int64), np.zeros(0, np.int64))
    assert perm.tolist() == [0] and nd.ndOrdering().CParent.tolist() == [-1]
    with pytest.raises(ValueError):
        nd.order(np.array([0, 1], np.int64), np.array([5], np.int64))      # row index out of range


def test_unchanged_pattern_reuses_the_tree():
    p, i = cube_graph(8)
    nd = mfs.NestedDissection(4)
    perm = nd.order(p, i)
    o = nd.ndOrdering()
    perm2 = nd.order(p, i)
    st = nd.temporalReuseStatistics()
    assert not st.full_rebuild and st.recomputed_subtrees == 0 and st.repartitioned_vertices == 0
    assert (perm2 == perm).all()
    assert (nd.ndOrdering().CParent == o.CParent).all() and (nd.ndOrdering().CMember == o.CMember).all()


def deep_violation(par, mem, rng):
    """A vertex pair whose components are not on one root path and whose common ancestor is not
    the root: the contact a local repair handles."""
    d = orc.depths(par)
    n = len(mem)
    for _ in range(100000):
        u, v = (int(x) for x in rng.integers(0, n, 2))
        a, b = mem[u], mem[v]
        r = orc.lca(par, d, a, b)
        if r not in (a, b) and r >= 0 and par[r] >= 0 and d[r] >= 2:
            return u, v, r
    raise AssertionError("no deep violating pair found")


def test_local_contact_repairs_one_subtree_and_keeps_the_rest():
    p, i = cube_graph(10)
    nd = mfs.NestedDissection(8)
    perm = nd.order(p, i)
    par, mem = check_forest(nd, perm, p, i)
    u, v, r = deep_violation(par, mem, np.random.default_rng(1))
    p2, i2 = cube_graph(10, [(u, v)])
    full, roots = orc.dirty_subtrees(par, mem, orc.edges_of(p2, i2))
    assert not full and roots == [r]
    region = orc.subtree_vertices(par, mem, r)
    perm2 = nd.order(p2, i2)
    st = nd.temporalReuseStatistics()
    assert not st.full_rebuild and st.recomputed_subtrees == 1 and st.repartitioned_vertices == len(region)
    check_forest(nd, perm2, p2, i2)
    # the repaired subtree keeps its contiguous range; every position outside it is unchanged
    pos = np.sort(np.nonzero(np.isin(perm, region))[0])
    lo, hi = pos[0], pos[-1] + 1
    assert hi - lo == len(region)
    assert (perm2[:lo] == perm[:lo]).all() and (perm2[hi:] == perm[hi:]).all()
    assert sorted(perm2[lo:hi].tolist()) == sorted(region)


def test_violating_the_root_separator_forces_a_full_rebuild():
    p, i = cube_graph(8)
    nd = mfs.NestedDissection(8)
    perm = nd.order(p, i)
    par, mem = check_forest(nd, perm, p, i)
    root = len(par) - 1
    kids = [c for c in range(len(par)) if par[c] == root]
    u = orc.subtree_vertices(par, mem, kids[0])[0]
    v = orc.subtree_vertices(par, mem, kids[1])[0]
    p2, i2 = cube_graph(8, [(u, v)])
    assert orc.dirty_subtrees(par, mem, orc.edges_of(p2, i2))[0]
    perm2 = nd.order(p2, i2)
    assert nd.temporalReuseStatistics().full_rebuild
    check_forest(nd, perm2, p2, i2)


def test_period_expiry_reset_and_period_change():
    p, i = cube_graph(6)
    nd = mfs.NestedDissection(2)
    flags = []
    for _ in range(5):
        nd.order(p, i)
        flags.append(nd.temporalReuseStatistics().full_rebuild)
    assert flags == [True, False, False, True, False]          # P incremental calls, then rebuild
    nd.resetTemporalReuse()
    nd.order(p, i)
    assert nd.temporalReuseStatistics().full_rebuild
    nd.order(p, i)
    assert not nd.temporalReuseStatistics().full_rebuild
    nd.temporalReusePeriod = 3                                   # a new period restarts the history
    nd.order(p, i)
    assert nd.temporalReuseStatistics().full_rebuild
    nd.temporalReusePeriod = 0                                   # reuse off: always from scratch
    nd.order(p, i)
    nd.order(p, i)
    assert nd.temporalReuseStatistics().full_rebuild
    q, j = cube_graph(5)                                         # other vertex set: from scratch
    nd.temporalReusePeriod = 4
    nd.order(p, i)
    nd.order(q, j)
    assert nd.temporalReuseStatistics().full_rebuild


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_contact_sequences_follow_the_oracle(seed):
    rng = np.random.default_rng(seed)
    N = 9
    tets, X = wl.cube_mesh(N)
    H = mfs.SystemAssembler(len(X)).blockSparsityPattern(tets)
    n = H.n
    nd = mfs.NestedDissection(6)
    p, i = block_graph(n, H.Ap, H.Ai)
    perm = nd.order(p, i)
    pool = [tuple(sorted(int(x) for x in rng.integers(0, n, 2))) for _ in range(30)]
    pool = [e for e in pool if e[0] != e[1]]
    on = rng.random(len(pool)) < 0.1
    since = 0
    for step in range(16):
        o = nd.ndOrdering()
        par, mem = o.CParent.tolist(), o.CMember.tolist()
        on ^= rng.random(len(pool)) < 0.15
        extra = [e for e, k in zip(pool, on) if k]
        p, i = block_graph(n, H.Ap, H.Ai, extra)
        full, roots = orc.dirty_subtrees(par, mem, orc.edges_of(p, i))
        expect_full = full or since >= 6
        perm = nd.order(p, i)
        st = nd.temporalReuseStatistics()
        assert st.full_rebuild == expect_full, step
        if not expect_full:
            assert st.recomputed_subtrees == len(roots), step
            assert st.repartitioned_vertices == sum(len(orc.subtree_vertices(par, mem, r)) for r in roots)
        since = 0 if expect_full else since + 1
        check_forest(nd, perm, p, i)
