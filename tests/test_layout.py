"""CPU: the non-contiguous block layout (BlockCSCHessian<..., ContiguousBlocks=false>).

Follows the reference's tests/test_block_csc_hessian.cc:52-110: values shuffled into the other
layout equal the scalar upper-triangle CSC of the same matrix (here built independently by
scipy from the dense matrix), shuffling back restores the block values bit for bit (diagonal
blocks' strict lower triangles zeroed), and the offset helpers (scalarOffsetForColumn,
scalarColStride, locForBlock) agree with the scalar CSC's column pointers and entry positions.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl


def random_hessian(N, seed=0):
    tets, _ = wl.cube_mesh(N)
    H = mfs.SystemAssembler((N + 1) ** 3).blockSparsityPattern(tets)
    H = mfs.BlockCSCHessian(H.Ap, H.Ai, device="cpu")
    rng = np.random.default_rng(seed)
    Ax = rng.standard_normal(9 * H.nz)
    d = 9 * (H.Ap[1:] - 1)
    for c in range(3):
        for r in range(c + 1, 3):
            Ax[d + 3 * c + r] = 0.0          # strict lower triangle of the diagonal blocks
    H.Ax = torch.from_numpy(Ax)
    return H


@pytest.mark.parametrize("N", [1, 3])
def test_noncontiguous_values_are_the_scalar_upper_csc(N):
    H = random_hessian(N)
    A = H.toDense()
    U = sp.csc_matrix(np.triu(A))
    U.sort_indices()
    nc = H.cloneWithLayout(False)
    assert not nc.hasContiguousBlocks() and nc.scalarNNZ() == U.nnz == 9 * H.nz - 3 * H.n
    assert np.array_equal(nc.Ax.numpy(), U.data)
    S = nc.toScalar()
    assert np.array_equal(S.Ap, U.indptr) and np.array_equal(S.Ai, U.indices)
    assert S.Ax is nc.Ax                      # shared, not copied
    back = nc.cloneWithLayout(True)
    assert np.array_equal(back.Ax.numpy(), H.Ax.numpy())
    assert np.array_equal(nc.toDense(), A)
    assert nc.trace() == pytest.approx(np.trace(A), rel=1e-14)
    assert np.array_equal(nc.Ax.numpy()[nc.diagEntryPositions()], np.diag(A))
    for bj in range(H.n):
        assert nc.scalarOffsetForColumn(bj) == U.indptr[3 * bj]
        assert nc.scalarColStride(bj) == U.indptr[3 * bj + 1] - U.indptr[3 * bj]
        for k in range(H.Ap[bj], H.Ap[bj + 1]):
            bi = H.Ai[k]
            col = U.indices[U.indptr[3 * bj]:U.indptr[3 * bj + 1]]
            assert nc.locForBlock(bi, bj) == U.indptr[3 * bj] + np.searchsorted(col, 3 * bi)
            assert H.locForBlock(bi, bj) == 9 * k


def test_noncontiguous_layout_needs_diagonal_blocks():
    with pytest.raises(ValueError):
        mfs.BlockCSCHessian(np.array([0, 1, 2]), np.array([0, 0]), device="cpu", contiguous=False)


def test_noncontiguous_binary_round_trip(tmp_path):
    """The binary format is the non-contiguous layout on disk (BlockCSCHessian.cc:11-40); a
    non-contiguous matrix writes and reads back bit for bit."""
    nc = random_hessian(2, seed=3).cloneWithLayout(False)
    p = tmp_path / "h.bin"
    nc.dumpBinaryToFile(p)
    r = mfs.BlockCSCHessian.constructFromBinaryFile(p, device="cpu", contiguous=False)
    assert not r.hasContiguousBlocks()
    assert np.array_equal(r.Ap, nc.Ap) and np.array_equal(r.Ai, nc.Ai)
    assert np.array_equal(r.Ax.numpy(), nc.Ax.numpy())
