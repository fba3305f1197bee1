"""GPU: bordered Hessians (BorderedSparseHessian.hh / .cc) solved through the GPU supernodal
factorization of the sparse block, against dense solves of the full bordered system: dense
padding variables, a low-rank term (Sherman-Morrison-Woodbury through the -I block), pinned
sparse variables, either variable order, and the projection of an indefinite Schur block."""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl

pytestmark = pytest.mark.gpu


def sparse_block(dev, N=4, seed=0):
    tets, Xr = wl.cube_mesh(N)
    X = wl.perturb(N, Xr, seed)
    He, _, _ = wl.neohookean(Xr, X, tets)
    asm = mfs.SystemAssembler(len(Xr))
    H = asm.blockSparsityPattern(tets)
    H.setZero()
    asm.assembleHessian(H, tets, torch.from_numpy(He.ravel()).to(dev))
    return H, wl.z0_pins(N)


def dense_full(Hb):
    """The bordered matrix [H_ss H_sd; H_ds H_dd] + V V^T in the Hessian's variable order."""
    A = Hb.H_ss.toDense()
    nsv, ndv = Hb.numSparseVars(), Hb.numDenseVars()
    M = np.zeros((nsv + ndv, nsv + ndv))
    s = slice(0, nsv) if Hb.sparse_first else slice(ndv, ndv + nsv)
    d = slice(nsv, nsv + ndv) if Hb.sparse_first else slice(0, ndv)
    M[s, s] = A
    M[s, d] = Hb.H_sd.cpu().numpy()
    M[d, s] = Hb.H_sd.cpu().numpy().T
    M[d, d] = Hb.H_dd.cpu().numpy()
    V = np.zeros((nsv + ndv, Hb.low_rank_rank()))
    V[s] = Hb.V_s.cpu().numpy()
    V[d] = Hb.V_d.cpu().numpy()
    return M + V @ V.T


def reduced_solve(M, b, fixed):
    n = M.shape[0]
    free = np.setdiff1d(np.arange(n), fixed)
    x = np.zeros(n)
    x[free] = np.linalg.solve(M[np.ix_(free, free)], b[free])
    return x


def bordered(dev, ndv, rank, sparse_first=True, seed=0, hdd_scale=None):
    H, pins = sparse_block(dev, seed=seed)
    rng = np.random.default_rng(seed + 10)
    Hb = mfs.BorderedSparseHessian(H, ndv, sparse_first)
    nsv = Hb.numSparseVars()
    diag = float(np.abs(H.toDense().diagonal()).mean())
    Hb.H_sd = torch.from_numpy(0.1 * diag * rng.standard_normal((nsv, ndv))).to(dev)
    G = rng.standard_normal((ndv, ndv))
    Hb.H_dd = torch.from_numpy((hdd_scale if hdd_scale is not None else 1e4 * nsv * diag) *
                               (G @ G.T / ndv + np.eye(ndv)) if ndv else np.zeros((0, 0))).to(dev)
    if rank:
        Hb.addLowRank(torch.from_numpy(np.sqrt(diag) * rng.standard_normal((nsv + ndv, rank))).to(dev), w=0.5)
    off = 0 if sparse_first else ndv
    fixed = [off + p for p in pins]
    return Hb, fixed


@pytest.mark.parametrize("ndv,rank", [(0, 0), (3, 0), (0, 2), (4, 3)])
@pytest.mark.parametrize("sparse_first", [True, False])
def test_bordered_solve_matches_dense(dev, ndv, rank, sparse_first):
    Hb, fixed = bordered(dev, ndv, rank, sparse_first)
    F = mfs.BorderedSparseFactorization(Hb, fixed)
    assert not F.indefinite
    rng = np.random.default_rng(5)
    b = rng.standard_normal(Hb.numVars())
    b[fixed] = 0
    x = F.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
    xe = reduced_solve(dense_full(Hb), b, fixed)
    assert np.linalg.norm(x - xe) <= 1e-9 * np.linalg.norm(xe)
    assert np.abs(x[fixed]).max(initial=0.0) == 0.0


def test_indefinite_schur_block_is_projected(dev):
    Hb, fixed = bordered(dev, 3, 2, hdd_scale=0.0)
    Hb.H_dd = -torch.eye(3, dtype=torch.float64, device=dev) * 1e3
    F = mfs.BorderedSparseFactorization(Hb, fixed)
    assert F.indefinite
    # the true Schur complement of the reduced system, by dense algebra
    M = dense_full(Hb)
    nsv = Hb.numSparseVars()
    free = np.setdiff1d(np.arange(nsv), fixed)
    A = Hb.H_ss.toDense()[np.ix_(free, free)]
    B = np.hstack([Hb.H_sd.cpu().numpy(), Hb.V_s.cpu().numpy()])[free]
    D = np.block([[Hb.H_dd.cpu().numpy(), Hb.V_d.cpu().numpy()],
                  [Hb.V_d.cpu().numpy().T, -np.eye(2)]])
    S = D - B.T @ np.linalg.solve(A, B)
    lam = np.linalg.eigvalsh(0.5 * (S[:3, :3] + S[:3, :3].T))
    got = np.linalg.eigvalsh(F.S[:3, :3])
    assert np.allclose(np.sort(got), np.sort(np.maximum(np.abs(lam), 1e-10)), rtol=1e-8)
    assert np.allclose(F.S[3:, 3:], S[3:, 3:], rtol=1e-9, atol=1e-9 * np.abs(S).max())
    del M


def test_numeric_update_and_errors(dev):
    Hb, fixed = bordered(dev, 2, 1)
    F = mfs.BorderedSparseFactorization(Hb, fixed)
    b = np.random.default_rng(1).standard_normal(Hb.numVars())
    b[fixed] = 0
    Hb.H_dd.mul_(2.0)
    Hb.addDiag(1.0)
    F.updateNumericFactorization(Hb)
    x = F.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
    xe = reduced_solve(dense_full(Hb), b, fixed)
    assert np.linalg.norm(x - xe) <= 1e-9 * np.linalg.norm(xe)
    with pytest.raises(RuntimeError):
        mfs.BorderedSparseFactorization(Hb, fixed + [Hb.numSparseVars()])    # a dense fixed variable
    with pytest.raises(RuntimeError):
        mfs.BorderedSparseFactorization(Hb, [Hb.numVars() + 5])
    Hb.H_sd = Hb.H_sd[:-1]
    with pytest.raises(RuntimeError):
        F.updateNumericFactorization(Hb)
