"""Host side of the learned preconditioner: the multicolor block graph, and the CPU oracle of the
network (pinned by autograd finite-difference checks and by the closed form of its initial
factor)."""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl, learned
from tests import learned_oracle as lo


def mesh_pattern(N):
    tets, X = wl.cube_mesh(N)
    H = mfs.SystemAssembler(len(X)).blockSparsityPattern(tets)
    return H, wl.z0_pins(N), tets, X


@pytest.mark.parametrize("N", [2, 5])
def test_graph_coloring_and_lower_blocks(N):
    H, pins, _, _ = mesh_pattern(N)
    g = learned.graph_host(H, pins)
    d, s, c = g["edge_dst"], g["edge_src"], g["color"]
    # edges: both directions of every off-diagonal free block, grouped by destination
    free = np.ones(H.n, bool)
    free[np.asarray(pins) // 3] = False
    idx = -np.ones(H.n, np.int64)
    idx[free] = np.arange(free.sum())
    pairs = set()
    for j in range(H.n):
        for k in range(H.Ap[j], H.Ap[j + 1]):
            i = H.Ai[k]
            if i != j and free[i] and free[j]:
                pairs.add((idx[i], idx[j]))
                pairs.add((idx[j], idx[i]))
    assert set(zip(d.tolist(), s.tolist())) == pairs and len(d) == len(pairs)
    assert (np.diff(d) >= 0).all()
    assert (c[d] != c[s]).all()                                # a proper coloring
    # lower blocks: one per undirected pair, the later color is the row
    ld, ls, le = g["lower_dst"], g["lower_src"], g["lower_edge"]
    assert len(ld) * 2 == len(d)
    assert (c[ld] > c[ls]).all()
    assert (d[le] == ld).all() and (s[le] == ls).all()


def tiny_problem(seed=0, n_nodes=5, edges=((0, 1), (1, 2), (2, 3), (3, 4), (0, 4), (1, 3))):
    rng = np.random.default_rng(seed)
    n = n_nodes
    und = list(edges)
    color = np.zeros(n, np.int32)
    for v in range(n):        # greedy
        used = {color[u] for (a, b) in und for u in ((b,) if a == v else (a,) if b == v else ()) if u < v}
        color[v] = min(set(range(n)) - used)
    ed, es = [], []
    for v in range(n):
        for (a, b) in sorted(und):
            if a == v:
                ed.append(a); es.append(b)
            elif b == v:
                ed.append(b); es.append(a)
    order = np.lexsort((np.arange(n), color))
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    ed, es = np.array(ed, np.int32), np.array(es, np.int32)
    o = np.lexsort((es, ed))
    ed, es = ed[o], es[o]
    low = np.nonzero(pos[ed] > pos[es])[0]
    g = dict(edge_dst=ed, edge_src=es, lower_edge=low.astype(np.int32), lower_dst=ed[low], lower_src=es[low])
    # a symmetric positive definite block matrix, Jacobi scaled
    A = rng.standard_normal((3 * n, 3 * n)) * 0.1
    A = A + A.T
    mask = np.zeros((n, n), bool)
    for a, b in und:
        mask[a, b] = mask[b, a] = True
    np.fill_diagonal(mask, True)
    A = A * np.kron(mask, np.ones((3, 3)))
    A += np.diag(np.abs(A).sum(1) + 1.0)
    s = 1.0 / np.sqrt(np.diag(A))
    Ah = s[:, None] * A * s[None, :]
    blk = lambda i, j: Ah[3 * i:3 * i + 3, 3 * j:3 * j + 3]
    feats = dict(node=rng.standard_normal((n, 6)).astype(np.float32),
                 edge=np.array([blk(i, j).T.ravel() for i, j in zip(ed, es)], np.float32),
                 diag=np.array([blk(i, i).T.ravel() for i in range(n)], np.float32))
    return g, feats, Ah, n


def oracle_params(seed=1, out_scale=1.0):
    rng = np.random.default_rng(seed)
    flat, offs = [], [0]
    for k in lo.NAMES:
        sz = int(np.prod(lo.SHAPES[k]))
        v = rng.standard_normal(sz) * (0.2 if k.startswith("W") else 0.05)
        if k in ("Wd2", "bd2", "Wdd", "bdd"):
            v = v * out_scale
        flat.append(v)
        offs.append(offs[-1] + sz)
    return np.concatenate(flat).astype(np.float32), np.array(offs)


def test_oracle_initial_factor_closed_form():
    """With zero output layers the factor is L = I + strict-lower(A^) in the color order, so the
    oracle's L L^T z equals the dense product of that matrix."""
    g, feats, Ah, n = tiny_problem()
    flat, offs = oracle_params(out_scale=0.0)
    z = np.random.default_rng(3).standard_normal((1, 3 * n))
    loss, _ = lo.loss_and_grads(flat, offs, g, feats, z)
    color = np.zeros(n)
    pos = np.zeros(n, np.int64)
    # rebuild the order from the lower blocks: row later than column
    Ahf = Ah.astype(np.float32).astype(np.float64)
    Ld = np.zeros((3 * n, 3 * n))
    for i in range(n):
        B = Ahf[3 * i:3 * i + 3, 3 * i:3 * i + 3]
        Ld[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.tril(B, -1) + np.eye(3)
    for i, j in zip(g["lower_dst"], g["lower_src"]):
        Ld[3 * i:3 * i + 3, 3 * j:3 * j + 3] = Ahf[3 * i:3 * i + 3, 3 * j:3 * j + 3]
    zz = z[0]
    w = Ld @ (Ld.T @ zz)
    a = Ahf @ zz
    assert abs(loss - np.sum((w - a) ** 2) / np.sum(a ** 2)) <= 1e-6 * loss
    del color, pos


def test_oracle_gradients_pass_finite_differences():
    """The oracle's autograd gradient against central differences of its own loss (float64,
    unrounded weights), per parameter tensor along a random direction."""
    g, feats, Ah, n = tiny_problem(seed=2)
    flat, offs = oracle_params(seed=4, out_scale=0.3)
    flat = flat.astype(np.float64) * 0.25
    z = np.random.default_rng(5).standard_normal((2, 3 * n))
    loss, grads = lo.loss_and_grads(flat, offs, g, feats, z, round_weights=False)
    assert 1e-3 < loss < 1e3
    rng = np.random.default_rng(6)
    for k in lo.NAMES:
        i = lo.NAMES.index(k)
        d = np.zeros_like(flat)
        d[offs[i]:offs[i + 1]] = rng.standard_normal(offs[i + 1] - offs[i])
        h = 1e-6
        fp = lo.loss_and_grads(flat + h * d, offs, g, feats, z, round_weights=False)[0]
        fm = lo.loss_and_grads(flat - h * d, offs, g, feats, z, round_weights=False)[0]
        fd = (fp - fm) / (2 * h)
        an = float(np.dot(grads, d))
        assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-6) + 1e-9, (k, fd, an)
