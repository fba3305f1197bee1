"""GPU: the non-contiguous block layout (BlockCSCHessian<..., ContiguousBlocks=false>) in memory.

* Assembly straight into the non-contiguous layout equals the contiguous assembly shuffled into
  it, bit for bit (same element order per entry), for both element-Hessian layouts, and
  accumulates like the contiguous one (SystemAssembler.hh:559-565).
* The reference runs checkSolves over both layouts (tests/test_catamari_factorizer.cc:130):
  here on the chain and mesh problems with whole-block and partial pins, in FP64 and FP32.
* The factor of the non-contiguous matrix is the contiguous matrix's factor, bit for bit, under
  the same ordering.
"""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl
from meshfemsparse_amd._lib import MFS_HE_FULL12, MFS_HE_UPPER78
from tests import oracle_lib as orc
from tests.test_gpu_assembly import pack78
from tests.test_gpu_factor import chain_problem, check_solves, mesh_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", [MFS_HE_FULL12, MFS_HE_UPPER78])
def test_assembly_into_noncontiguous_layout(dev, layout):
    N = 5
    tets, Xr = wl.cube_mesh(N)
    X = wl.perturb(N, Xr, 1)
    He, _, _ = wl.neohookean(Xr, X, tets)
    if layout == MFS_HE_UPPER78:
        He = pack78(He)
    He = torch.from_numpy(np.ascontiguousarray(He).ravel()).to(dev)
    asm = mfs.SystemAssembler(len(Xr))
    Hc = asm.blockSparsityPattern(tets)
    Hc.setZero()
    asm.assembleHessian(Hc, tets, He, layout)
    Hn = mfs.BlockCSCHessian(Hc.Ap, Hc.Ai, dev, contiguous=False)
    Hn.setZero()
    assert Hn.Ax.numel() == 9 * Hc.nz - 3 * Hc.n
    asm.assembleHessian(Hn, tets, He, layout)
    assert torch.equal(Hn.Ax, Hc.cloneWithLayout(False).Ax)
    asm.assembleHessian(Hc, tets, He, layout)      # accumulate a second time
    asm.assembleHessian(Hn, tets, He, layout)
    assert torch.equal(Hn.Ax, Hc.cloneWithLayout(False).Ax)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("problem", ["chain", "mesh", "mesh_partial"])
def test_check_solves_noncontiguous(dev, problem, single):
    if problem == "chain":
        H, pins = chain_problem(dev), [0, 1, 2]
    else:
        H, pins, _, _, _ = mesh_problem(dev, 4, seed=3)
        if problem == "mesh_partial":
            pins = sorted(set(list(pins) + [3 * 40 + 0, 3 * 41 + 2, 3 * 57 + 1]))
    check_solves(dev, H.cloneWithLayout(False), pins, single)


@pytest.mark.parametrize("single", [False, True])
def test_noncontiguous_factor_is_the_contiguous_factor(dev, single):
    H, pins, g, _, _ = mesh_problem(dev, 6, seed=2)
    Hn = H.cloneWithLayout(False)
    fc = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
    fn = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
    fc.factorize(H, pins)
    fn.factorize(Hn, pins)
    assert np.array_equal(fc.getInversePermutation(), fn.getInversePermutation())
    assert fc.getFactorNNZ() == fn.getFactorNNZ()
    Lc = orc.dense_L(*fc.exportFactorCSC())
    assert np.array_equal(orc.dense_L(*fn.exportFactorCSC()), Lc)
    b = torch.from_numpy(-g).to(dev)
    xc, _, _ = fc.solveRefined(H, b, tol=1e-10, maxIters=10)
    xn, _, rr = fn.solveRefined(Hn, b, tol=1e-10, maxIters=10)
    assert rr <= 1e-10
    assert torch.linalg.norm(xn - xc) <= 1e-9 * torch.linalg.norm(xc)
    with pytest.raises(ValueError):
        fn.factorizeNumeric(H)                       # layout differs from the analysis
