"""Recorded Newton-step matrix sequences: the reference's MatrixRecorder directory layout
(src/lib/MeshFEMSparse/Solvers/MatrixRecorder.hh:36-72) and a replay of its
benchmark_linear_systems driver (tests/benchmark_linear_systems.cc:216-342) through the GPU
factorizer.

A directory holds, in one id sequence, symbolic_mat_XXXX_from_update_K.bin (a pattern, factor
symbolically with the pins of pinned_vars_XXXX.txt) and numeric_mat_XXXX.bin (values, factor
numerically, then verify the backward error of a solve with a random solution that is zero on
the pins).
"""
import glob
import os
import time

import numpy as np
import torch

from .sparse import BlockCSCHessian


def _id(i):
    return f"{i:04d}"


class MatrixRecorder:
    """Writes matrices in the reference's recording layout."""

    def __init__(self, directory):
        if not directory:
            raise RuntimeError("MatrixRecorder: Must pass a nonempty path")
        os.makedirs(directory, exist_ok=True)
        self.dir = directory
        self.matrixId = 0
        self.numSparsityUpdateCalls = 0

    @staticmethod
    def symbolicMatrixFileName(i, suffix=""):
        return f"symbolic_mat_{_id(i)}{suffix}.bin"

    @staticmethod
    def numericMatrixFileName(i):
        return f"numeric_mat_{_id(i)}.bin"

    @staticmethod
    def pinnedVarsFileName(i):
        return f"pinned_vars_{_id(i)}.txt"

    def countSparsityUpdateCall(self):
        self.numSparsityUpdateCalls += 1

    def recordSymbolic(self, H, pinnedVars):
        i = self.matrixId
        self.matrixId += 1
        suffix = f"_from_update_{self.numSparsityUpdateCalls - 1}"
        Ax, H.Ax = H.Ax, None            # the pattern only
        try:
            H.dumpBinaryToFile(os.path.join(self.dir, self.symbolicMatrixFileName(i, suffix)))
        finally:
            H.Ax = Ax
        with open(os.path.join(self.dir, self.pinnedVarsFileName(i)), "w") as fh:
            for v in pinnedVars:
                fh.write(f"{int(v)}\n")

    def recordNumeric(self, H):
        i = self.matrixId
        self.matrixId += 1
        H.dumpBinaryToFile(os.path.join(self.dir, self.numericMatrixFileName(i)))


def replay(directory, factorizer, repeats=1, use_shift=False, device="cuda", seed=0):
    """benchmark_linear_systems' loop: returns dict(posdef, indef, numeric_s (per factorization),
    backward errors).  The random solution is zero on the pinned variables, as in the driver."""
    rng = np.random.default_rng(seed)
    out = dict(posdef=0, indef=0, numeric_s=[], backward_error=[], symbolic=0)
    x_gt = None
    counter = 0
    while True:
        pat = os.path.join(directory, MatrixRecorder.symbolicMatrixFileName(counter, "_from_update_*"))
        sym = sorted(glob.glob(pat))
        if sym:
            pins = np.loadtxt(os.path.join(directory, MatrixRecorder.pinnedVarsFileName(counter)), dtype=np.int64,
                              ndmin=1)
            Hs = BlockCSCHessian.constructFromBinaryFile(sym[0], device)
            factorizer.factorizeSymbolic(Hs, pins)
            out["symbolic"] += 1
            x_gt = rng.uniform(-1, 1, 3 * Hs.n)
            x_gt[np.asarray(factorizer.getFixedVars(), np.int64)] = 0.0
            counter += 1
            continue
        num = os.path.join(directory, MatrixRecorder.numericMatrixFileName(counter))
        if os.path.exists(num):
            if not factorizer.hasFactorization(1):
                raise RuntimeError("Numeric matrix encountered before symbolic matrix")
            H = BlockCSCHessian.constructFromBinaryFile(num, device)
            for r in range(repeats):
                try:
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    if use_shift:
                        factorizer.factorizeNumericWithShift(H, 1e-8 * H.trace() / H.numScalarCols())
                    else:
                        factorizer.factorizeNumeric(H)
                    torch.cuda.synchronize()
                    out["numeric_s"].append(time.perf_counter() - t0)
                    if not factorizer.checkPosDef():
                        raise RuntimeError("Non-positive definite matrix detected by checkPosDef")
                    if r > 0:
                        continue
                    out["posdef"] += 1
                    xg = torch.from_numpy(x_gt).to(device)
                    b = factorizer.apply(H, xg)
                    x = factorizer.solve(b)
                    err = float(torch.linalg.norm(b - factorizer.apply(H, x)) / torch.linalg.norm(b))
                    out["backward_error"].append(err)
                except RuntimeError:
                    if r == 0:
                        out["indef"] += 1
            counter += 1
            continue
        break
    return out
