"""GPU: the allocation-free solve (mfs_solve_noalloc) against mfs_solve, through the C ABI.

The allocation-free entry point runs the same kernels as mfs_solve with every scratch vector
carved out of the library's static device array, so its solution must equal mfs_solve's bit for
bit.  Two calls in a row on one GPU share that array: the second (a smaller problem, with a
partially pinned block) must not see what the first left there.  A fresh factorizer's
allocation-free solve leaves the device's free memory unchanged, and a solve whose scratch does
not fit is refused with MFS_ERR_ARG.
"""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import _lib
from meshfemsparse_amd.sparse import _dptr, _stream
from tests.test_gpu_factor import mesh_problem

pytestmark = pytest.mark.gpu


def c_solve(name, f, b, nrhs):
    x = torch.empty_like(b)
    _lib.call(name, f._h, _dptr(b), _dptr(x), nrhs, _stream())
    return x


@pytest.mark.parametrize("single", [False, True])
def test_noalloc_solve_equals_solve(dev, single):
    rng = np.random.default_rng(5)
    for k, (N, nrhs, extra) in enumerate(((10, 3, []), (6, 1, [3 * 60 + 1]))):
        H, pins, g, _, _ = mesh_problem(dev, N, seed=N)
        pins = sorted(set(list(pins) + extra))
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
        f.factorizeSymbolic(H, pins)
        f.factorizeNumeric(H)
        scale = 1e6 if k == 0 else 1.0         # the first call leaves large values in the scratch
        b = torch.from_numpy(scale * rng.standard_normal(nrhs * 3 * H.n)).to(dev)
        x1 = torch.empty_like(b)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        _lib.call("mfs_solve_noalloc", f._h, _dptr(b), _dptr(x1), nrhs, _stream())
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        if k == 1:
            # a fresh factorizer that never solved before: nothing was allocated for the solve
            assert free1 == free0, (free0, free1)
        x0 = c_solve("mfs_solve", f, b, nrhs)
        torch.cuda.synchronize()
        print(f"N={N} nrhs={nrhs} single={single}: max |x_noalloc - x_solve| = "
              f"{float((x1 - x0).abs().max()):.1e}, free memory {free0} -> {free1}")
        assert torch.equal(x1, x0)
        # and the solution is a solution (pinned entries zero)
        assert float(x1.view(nrhs, -1)[:, torch.from_numpy(np.array(pins)).to(dev)].abs().max()) == 0.0


def test_noalloc_solve_rejects_oversized_scratch(dev):
    H, pins, g, _, _ = mesh_problem(dev, 16, seed=1)
    f = mfs.make_cholesky_factorizer()
    f.factorize(H, pins)
    nrhs = 150                                   # about 1 MB of scratch per right-hand side
    b = torch.zeros(nrhs * 3 * H.n, dtype=torch.float64, device=dev)
    with pytest.raises(_lib.InvalidArgumentError, match="scratch"):
        c_solve("mfs_solve_noalloc", f, b, nrhs)
    bb = torch.from_numpy(-g).to(dev)
    assert torch.equal(c_solve("mfs_solve_noalloc", f, bb, 1), c_solve("mfs_solve", f, bb, 1))
