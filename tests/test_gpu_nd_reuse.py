"""GPU factorizer: the separator tree of its analysis (ndOrdering, CholeskyFactorizerBase.hh:51-59)
and temporal reuse of that tree across symbolic analyses of changing (contact) patterns
(CatamariFactorizer.hh:84-90, provider CatamariNesdisReuse -> GpuSupernodalReuse): the tree is a
valid separator tree of the reduced block graph, incremental analyses repair it, and every
factorization still solves its own system."""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl
from tests import nd_oracle as orc

pytestmark = pytest.mark.gpu


def contact_hessian(dev, tets, Xr, N, extra, seed):
    """Cube Hessian plus SPD springs between the block pairs in `extra`."""
    nb = len(Xr)
    base = mfs.SystemAssembler(nb).blockSparsityPattern(tets)
    cols = [set(base.Ai[base.Ap[j]:base.Ap[j + 1]].tolist()) for j in range(nb)]
    for (i, j) in extra:
        cols[max(i, j)].add(min(i, j))
    Ap = np.zeros(nb + 1, np.int64)
    Ap[1:] = np.cumsum([len(c) for c in cols])
    Ai = np.array([r for j in range(nb) for r in sorted(cols[j])], np.int64)
    A = np.zeros((3 * nb, 3 * nb))
    He, _, _ = wl.neohookean(Xr, wl.perturb(N, Xr, seed), tets)
    base.setZero()
    mfs.SystemAssembler(nb).assembleHessian(base, tets, torch.from_numpy(He.ravel()).to(dev))
    A += base.toDense()
    for (i, j) in extra:
        for d in range(3):
            a, b = 3 * i + d, 3 * j + d
            A[a, a] += 1e2
            A[b, b] += 1e2
            A[a, b] -= 1e2
            A[b, a] -= 1e2
    H = mfs.BlockCSCHessian(Ap, Ai, dev)
    Ax = np.zeros(9 * len(Ai))
    for j in range(nb):
        for k in range(Ap[j], Ap[j + 1]):
            i = Ai[k]
            Ax[9 * k:9 * k + 9] = A[3 * i:3 * i + 3, 3 * j:3 * j + 3].ravel(order="F")
    H.Ax = torch.from_numpy(Ax).to(dev)
    return H, A


def reduced_block_graph(H, pins):
    pinned_blocks = sorted(set(int(p) // 3 for p in pins))
    keep = np.setdiff1d(np.arange(H.n), pinned_blocks)
    red = -np.ones(H.n, np.int64)
    red[keep] = np.arange(len(keep))
    edges = set()
    for j in range(H.n):
        for k in range(H.Ap[j], H.Ap[j + 1]):
            i = int(H.Ai[k])
            if i != j and red[i] >= 0 and red[j] >= 0:
                edges.add((int(min(red[i], red[j])), int(max(red[i], red[j]))))
    return len(keep), edges


def test_nd_ordering_of_the_analysis(dev):
    N = 6
    tets, Xr = wl.cube_mesh(N)
    pins = wl.z0_pins(N)
    f = mfs.make_cholesky_factorizer()
    assert f.ndOrdering() is None
    assert f.provider() == mfs.CholeskyProvider.GpuSupernodal and f.temporalReusePeriod() == 0
    H = mfs.SystemAssembler(len(Xr)).blockSparsityPattern(tets)
    f.factorizeSymbolic(H, pins)
    o = f.ndOrdering()
    nv, edges = reduced_block_graph(H, pins)
    assert o.blockSize == 3 and len(o.CMember) == nv == f.n_reduced() // 3
    par, mem = o.CParent.tolist(), o.CMember.tolist()
    assert all(par[c] == -1 or par[c] > c for c in range(len(par)))
    assert set(mem) == set(range(len(par)))
    assert orc.is_separator_tree(par, mem, edges)


def deep_contacts(o, keep, rng, count):
    """Block pairs whose components are not on one root path and whose common ancestor is at
    depth >= 2 of the analysis's tree: contacts that a local repair handles."""
    par, mem = o.CParent.tolist(), o.CMember.tolist()
    d = orc.depths(par)
    out = []
    while len(out) < count:
        u, v = (int(x) for x in rng.integers(0, len(mem), 2))
        a, b = mem[u], mem[v]
        r = orc.lca(par, d, a, b)
        if r not in (a, b) and r >= 0 and d[r] >= 2:
            out.append((int(keep[u]), int(keep[v])))
    return out


def test_temporal_reuse_across_changing_contacts(dev):
    N = 8
    tets, Xr = wl.cube_mesh(N)
    nb = len(Xr)
    pins = wl.z0_pins(N)
    keep = np.setdiff1d(np.arange(nb), sorted(set(int(p) // 3 for p in pins)))
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodalReuse)
    assert f.temporalReusePeriod() == 32 and f.provider() == mfs.CholeskyProvider.GpuSupernodalReuse
    rng = np.random.default_rng(3)
    free = np.setdiff1d(np.arange(3 * nb), pins)
    stats, pool = [], None
    schedule = [[], [0], [0, 1], [1, 2], []]
    for step, sel in enumerate(schedule):
        extra = [pool[q] for q in sel] if sel else []
        H, A = contact_hessian(dev, tets, Xr, N, extra, step)
        f.factorizeSymbolic(H, pins)
        stats.append(f.temporalReuseStatistics())
        o = f.ndOrdering()
        if pool is None:
            pool = deep_contacts(o, keep, rng, 3)
        _, edges = reduced_block_graph(H, pins)
        assert orc.is_separator_tree(o.CParent.tolist(), o.CMember.tolist(), edges), step
        f.factorizeNumeric(H)
        b = np.zeros(3 * nb)
        b[free] = rng.standard_normal(len(free))
        x = f.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
        xe = np.zeros(3 * nb)
        xe[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
        assert np.linalg.norm(x - xe) <= 1e-9 * np.linalg.norm(xe), step
    assert stats[0].full_rebuild
    assert not stats[1].full_rebuild and stats[1].recomputed_subtrees >= 1      # a local repair
    assert 0 < stats[1].repartitioned_vertices < len(keep)
    # different pins restart the history
    f.factorizeSymbolic(H, pins[3:])
    assert f.temporalReuseStatistics().full_rebuild
    f.factorizeSymbolic(H, pins[3:])
    assert not f.temporalReuseStatistics().full_rebuild
    f.resetSymbolicFactorizationReuse()
    f.factorizeSymbolic(H, pins[3:])
    assert f.temporalReuseStatistics().full_rebuild
