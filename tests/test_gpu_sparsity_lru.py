"""GPU: a changing (contact-like) pattern through SparsityLRU as the reference's replay uses it
(tests/benchmark_sparsity_lru.cc): one symbolic factorization of the cached pattern serves every
step whose pattern the cache covers; each step's numeric matrix is expanded into the cached
pattern and factored; solutions match dense solves of the step's own matrix."""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl
from meshfemsparse_amd.sparsity_lru import EXPIRED, SparsityLRU, expand_values

pytestmark = pytest.mark.gpu


def step_matrix(dev, tets, Xr, N, extra, seed):
    """Cube Hessian plus small SPD 'contact' couplings between the node pairs in `extra`."""
    nb = len(Xr)
    cols = {}
    Hs = mfs.SystemAssembler(nb).blockSparsityPattern(tets)
    for j in range(nb):
        cols[j] = set(Hs.Ai[Hs.Ap[j]:Hs.Ap[j + 1]].tolist())
    for (i, j) in extra:
        cols[j].add(i)
    Ap = np.zeros(nb + 1, np.int64)
    Ap[1:] = np.cumsum([len(cols[j]) for j in range(nb)])
    Ai = np.array([r for j in range(nb) for r in sorted(cols[j])], np.int64)
    H = mfs.BlockCSCHessian(Ap, Ai, dev)
    H.setZero()
    X = wl.perturb(N, Xr, seed)
    He, _, _ = wl.neohookean(Xr, X, tets)
    Hb = mfs.SystemAssembler(nb).blockSparsityPattern(tets)
    Hb.setZero()
    mfs.SystemAssembler(nb).assembleHessian(Hb, tets, torch.from_numpy(He.ravel()).to(dev))
    Hb = expand_values(Hb, Ap, Ai)
    Ax = Hb.Ax.clone()
    k = 1e2
    for (i, j) in extra:             # spring k (e e^T) between i and j, e = x axis: SPD addition
        def slot(r, c):
            p = np.searchsorted(Ai[Ap[c]:Ap[c + 1]], r) + Ap[c]
            return 9 * int(p)
        Ax[slot(i, i)] += k
        Ax[slot(j, j)] += k
        Ax[slot(i, j)] -= k
    H.Ax = Ax
    return H


def test_cached_pattern_serves_changing_contacts(dev):
    N = 3
    tets, Xr = wl.cube_mesh(N)
    nb = len(Xr)
    pins = wl.z0_pins(N)
    rng = np.random.default_rng(0)
    pool = [(int(i), int(j)) for i, j in zip(rng.integers(0, nb // 2, 12), rng.integers(nb // 2, nb, 12)) if i < j]
    base = step_matrix(dev, tets, Xr, N, [], 0)
    lru = SparsityLRU(base)
    lru.entryCacheBudgetRatio = 0.05
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(mfs.BlockCSCHessian(lru.Ap, lru.Ai, dev), pins)
    symbolic = 1
    for step in range(8):
        extra = [pool[q] for q in range(len(pool)) if (q + step) % 3 == 0]
        H = step_matrix(dev, tets, Xr, N, extra, step + 1)
        r = lru.update(H)
        if r != 0:
            f.factorizeSymbolic(mfs.BlockCSCHessian(lru.Ap, lru.Ai, dev), pins)
            symbolic += 1
        Hc = expand_values(H, lru.Ap, lru.Ai)
        f.factorizeNumeric(Hc)
        A = H.toDense()
        free = np.setdiff1d(np.arange(3 * nb), pins)
        b = np.zeros(3 * nb)
        b[free] = rng.standard_normal(len(free))
        x = f.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
        xe = np.zeros(3 * nb)
        xe[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
        assert np.linalg.norm(x - xe) <= 1e-9 * np.linalg.norm(xe)
        assert r in (0, EXPIRED) or r > 0
    assert symbolic < 9          # the cache let some steps reuse the symbolic factorization
