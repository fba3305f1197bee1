"""Bordered sparse Hessians and their factorization (BorderedSparseHessian.hh:94-331,
BorderedSparseHessian.cc:1-208): a block-sparse part H_ss padded with dense variables
(H_sd, H_dd) and a symmetric low-rank term V V^T, V = [V_s; V_d].  The system

    [H_ss  B] [x]   [b]        B = [H_sd  V_s],   D = [H_dd   V_d]
    [B^T   D] [y] = [c],                              [V_d^T  -I ]

is solved by block elimination: the GPU supernodal factorization of H_ss, one multi-RHS device
solve for H_ss^-1 B, the small Schur complement S = D - B^T H_ss^-1 B (its H_dd block projected
to positive definite by |eigenvalues| clamped at 1e-10, as the reference does), then
y = S^-1 (c - B^T H_ss^-1 b) by QR and x = H_ss^-1 (b - B y).

Variables: the sparse ones come first (scalar indices 0..nsv-1), the dense ones after them, or
the other way round with sparse_first=False (the reference's var structure allows either).
Dense blocks are float64 tensors on the Hessian's device; the Schur complement work is k x k
(k = dense variables + rank) on the host.
"""
import numpy as np
import torch

from .solvers import CholeskyProvider, make_cholesky_factorizer


class BorderedSparseHessian:
    """H_ss (BlockCSCHessian) bordered by num_dense dense variables and a low-rank term."""

    def __init__(self, H_ss, num_dense=0, sparse_first=True):
        self.H_ss = H_ss
        self.sparse_first = bool(sparse_first)
        dev = H_ss.Ax.device if H_ss is not None and H_ss.Ax is not None else torch.device("cuda")
        nsv = self.numSparseVars()
        ndv = int(num_dense)
        self.H_sd = torch.zeros(nsv, ndv, dtype=torch.float64, device=dev)
        self.H_dd = torch.zeros(ndv, ndv, dtype=torch.float64, device=dev)
        self.V_s = torch.zeros(nsv, 0, dtype=torch.float64, device=dev)
        self.V_d = torch.zeros(ndv, 0, dtype=torch.float64, device=dev)

    def numSparseVars(self):
        return 0 if self.H_ss is None else self.H_ss.numScalarCols()

    def numDenseVars(self):
        return self.H_dd.shape[0]

    def numVars(self):
        return self.numSparseVars() + self.numDenseVars()

    def low_rank_rank(self):
        return self.V_s.shape[1]

    def sparseVars(self, x):
        return x[: self.numSparseVars()] if self.sparse_first else x[self.numDenseVars():]

    def denseVars(self, x):
        return x[self.numSparseVars():] if self.sparse_first else x[: self.numDenseVars()]

    def isSparseVar(self, i):
        nsv, ndv = self.numSparseVars(), self.numDenseVars()
        return i < nsv if self.sparse_first else ndv <= i < ndv + nsv

    def sparseVarOffset(self):
        return 0 if self.sparse_first else self.numDenseVars()

    def denseVarOffset(self):
        return self.numSparseVars() if self.sparse_first else 0

    def validate(self):
        """BorderedSparseHessian::validate (BorderedSparseHessian.cc:110-134)."""
        nsv, ndv = self.numSparseVars(), self.numDenseVars()
        if nsv > 0 and self.H_ss is None:
            raise RuntimeError("H_ss is null")
        if ndv > 0 and nsv > 0:
            if self.H_sd.shape[0] != nsv:
                raise RuntimeError("H_sd has the wrong number of rows")
            if self.H_sd.shape[1] != ndv:
                raise RuntimeError("H_sd has the wrong number of columns")
            if tuple(self.H_dd.shape) != (ndv, ndv):
                raise RuntimeError("H_dd is the wrong shape")
        if self.V_s.numel() or self.V_d.numel():
            if self.V_s.shape[0] != nsv:
                raise RuntimeError("V_s has the wrong number of columns")
            if self.V_d.shape[0] != ndv:
                raise RuntimeError("V_d has the wrong number of columns")

    def setZero(self):
        if self.H_ss is not None:
            self.H_ss.setZero()
        self.H_sd.zero_()
        self.H_dd.zero_()
        self.V_s = self.V_s[:, :0]
        self.V_d = self.V_d[:, :0]

    def trace(self):
        """tr(H_ss) + tr(H_dd) + |V_d|^2 (BorderedSparseHessian.hh:203-206)."""
        if self.H_ss is None:
            raise RuntimeError("Hessian not initialized")
        return self.H_ss.trace() + float(torch.trace(self.H_dd)) + float((self.V_d ** 2).sum())

    def addDiag(self, d):
        if self.H_ss is not None:
            ax = self.H_ss.Ax
            ax[torch.as_tensor(self.H_ss.diagEntryPositions(), device=ax.device)] += d
        self.H_dd.diagonal().add_(d)

    def addLowRank(self, V, w=1.0):
        """Append the columns of sqrt(w) V (numVars x k) to the low-rank term."""
        V = torch.as_tensor(V, dtype=torch.float64, device=self.H_dd.device)
        if V.dim() == 1:
            V = V[:, None]
        if V.shape[0] != self.numVars():
            raise RuntimeError("Low-rank term must have same number of rows as the Hessian")
        if w != 1.0:
            V = V * np.sqrt(w)
        self.V_s = torch.cat([self.V_s, self.sparseVars(V)], dim=1)
        self.V_d = torch.cat([self.V_d, self.denseVars(V)], dim=1)

    def addNZ(self, i, j, val):
        """Scalar upper-triangle entry (i <= j) into the dense blocks (the sparse block takes its
        values through assembly)."""
        assert i <= j
        si, sj = self.isSparseVar(i), self.isSparseVar(j)
        so, do = self.sparseVarOffset(), self.denseVarOffset()
        if si and sj:
            raise RuntimeError("sparse-sparse entries are assembled into H_ss")
        if si:
            self.H_sd[i - so, j - do] += val
        elif sj:
            self.H_sd[j - so, i - do] += val
        else:
            self.H_dd[i - do, j - do] += val


class BorderedSparseFactorization:
    """BorderedSparseFactorization (BorderedSparseHessian.hh:371-431, .cc:5-208) over the GPU
    supernodal provider."""

    def __init__(self, H, fixedVars=(), factorizer=CholeskyProvider.GpuSupernodal, singlePrecision=False):
        self._nsv, self._ndv = H.numSparseVars(), H.numDenseVars()
        self._sparse_first = H.sparse_first
        self._rank = H.low_rank_rank()
        self.sparseFixedVars, self.denseFixedVars = [], []
        for i in fixedVars:
            if H.isSparseVar(i):
                self.sparseFixedVars.append(int(i) - H.sparseVarOffset())
            elif 0 <= i < H.numVars():
                self.denseFixedVars.append(int(i) - H.denseVarOffset())
            else:
                raise RuntimeError(f"Variable {i} is not a sparse or dense variable.")
        self._solver = make_cholesky_factorizer(factorizer, singlePrecision)
        if self._nsv > 0:
            self._solver.factorizeSymbolic(H.H_ss, self.sparseFixedVars)
        self.B = self.H_ss_inv_B = self.S = None
        self.indefinite = False
        self.updateNumericFactorization(H)

    def solver(self):
        return self._solver

    def exists(self):
        return self._solver is not None and (self._nsv == 0 or self._solver.hasNumericFactorization())

    def updateNumericFactorization(self, H):
        H.validate()
        if self._nsv > 0:
            self._solver.factorizeNumeric(H.H_ss)
        self._update_dense(H)

    def _update_dense(self, H):
        nsv, ndv, r = self._nsv, self._ndv, H.low_rank_rank()
        self._rank = r
        k = ndv + r
        if k == 0:
            self.B = self.H_ss_inv_B = self.S = None
            return False
        if self.denseFixedVars:
            raise RuntimeError("Dense fixed variables are not supported yet.")
        dev = H.H_dd.device
        B = torch.empty(nsv, k, dtype=torch.float64, device=dev)
        D = torch.zeros(k, k, dtype=torch.float64, device=dev)
        if ndv > 0:
            B[:, :ndv] = H.H_sd
            D[:ndv, :ndv] = H.H_dd
        if r > 0:
            B[:, ndv:] = H.V_s
            if ndv > 0:
                D[:ndv, ndv:] = H.V_d
                D[ndv:, :ndv] = H.V_d.T
            D[ndv:, ndv:] = -torch.eye(r, dtype=torch.float64, device=dev)
        self.B = B
        self.H_ss_inv_B = self._solver.solveMultiRHS(B.contiguous())
        S = (D - B.T @ self.H_ss_inv_B).cpu().numpy()
        S = 0.5 * (S + S.T)
        # project the H_dd block of the Schur complement (never the SMW block of the low-rank term)
        self.indefinite = False
        if ndv > 0:
            lam, Q = np.linalg.eigh(S[:ndv, :ndv])
            self.indefinite = bool((lam < 0).any())
            S[:ndv, :ndv] = (Q * np.maximum(np.abs(lam), 1e-10)) @ Q.T
        self.S = S
        return self.indefinite

    def solve(self, b):
        """x with [H_ss B; B^T D] [x_s; x_d; lambda] = [b_s; b_d; 0] (the multipliers dropped)."""
        if not self.exists():
            raise RuntimeError("Factorization doesn't exist.")
        b = torch.as_tensor(b, dtype=torch.float64)
        nsv, ndv = self._nsv, self._ndv
        if self.B is None:
            return self._solver.solve(b)
        dev = self.B.device
        b = b.to(dev)
        b_s = b[:nsv] if self._sparse_first else b[ndv:]
        b_d = b[nsv:] if self._sparse_first else b[:ndv]
        c = np.concatenate([b_d.cpu().numpy(), np.zeros(self._rank)])
        rhs = c - (self.H_ss_inv_B.T @ b_s).cpu().numpy()
        Qm, Rm = np.linalg.qr(self.S)
        y = np.linalg.solve(Rm, Qm.T @ rhs)
        y_d = torch.from_numpy(y).to(dev)
        x_s = self._solver.solve((b_s - self.B @ y_d).contiguous())
        x = torch.empty(nsv + ndv, dtype=torch.float64, device=dev)
        if self._sparse_first:
            x[:nsv], x[nsv:] = x_s, y_d[:ndv]
        else:
            x[:ndv], x[ndv:] = y_d[:ndv], x_s
        return x
