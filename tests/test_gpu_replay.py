"""GPU: a recorded Newton sequence replayed through the GPU provider the way the reference's
benchmark_linear_systems does (tests/benchmark_linear_systems.cc:216-342): symbolic factorization
per recorded pattern, numeric factorization per recorded matrix, and the backward error of a solve
with a random solution zero on the pins (the driver reports errors above 5e-5)."""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl
from meshfemsparse_amd.recorder import MatrixRecorder, replay

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_shift", [False, True])
def test_recorded_sequence_replays(dev, tmp_path, use_shift):
    N = 5
    tets, Xr = wl.cube_mesh(N)
    asm = mfs.SystemAssembler(len(Xr))
    H = asm.blockSparsityPattern(tets)
    rec = MatrixRecorder(str(tmp_path))
    rec.countSparsityUpdateCall()
    rec.recordSymbolic(H, wl.z0_pins(N))
    for s in range(3):
        He, _, _ = wl.neohookean(Xr, wl.perturb(N, Xr, s), tets, gradient=False)
        H.setZero()
        asm.assembleHessian(H, tets, torch.from_numpy(He.ravel()).to(dev))
        rec.recordNumeric(H)
    f = mfs.make_cholesky_factorizer()
    out = replay(str(tmp_path), f, repeats=2, use_shift=use_shift, device=dev)
    assert out["symbolic"] == 1 and out["posdef"] == 3 and out["indef"] == 0
    assert len(out["numeric_s"]) == 6
    # Unshifted: a direct solve.  Shifted: the solve is of H + s I with s = 1e-8 tr(H)/n while the
    # backward error is measured against H, so it is s |x| / |b| (~1e-8 here); the reference's
    # driver flags errors above 5e-5.
    assert max(out["backward_error"]) <= (5e-5 if use_shift else 1e-12)
    if use_shift:
        assert min(out["backward_error"]) > 1e-12


def test_indefinite_recorded_matrix_is_counted(dev, tmp_path):
    N = 3
    tets, Xr = wl.cube_mesh(N)
    asm = mfs.SystemAssembler(len(Xr))
    H = asm.blockSparsityPattern(tets)
    rec = MatrixRecorder(str(tmp_path))
    rec.countSparsityUpdateCall()
    rec.recordSymbolic(H, wl.z0_pins(N))
    He, _, _ = wl.neohookean(Xr, wl.perturb(N, Xr, 0), tets, gradient=False)
    H.setZero()
    asm.assembleHessian(H, tets, torch.from_numpy(-He.ravel()).to(dev))
    rec.recordNumeric(H)
    out = replay(str(tmp_path), mfs.make_cholesky_factorizer(), device=dev)
    assert out["indef"] == 1 and out["posdef"] == 0
