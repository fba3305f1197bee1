"""Host mirror of the reference's factorizer interface over the GPU supernodal factorizer.

CholeskyFactorizerBase (src/lib/MeshFEMSparse/Solvers/CholeskyFactorizerBase.hh:50-443) and the
factory make_cholesky_factorizer(provider, singlePrecision)
(src/lib/MeshFEMSparse/Solvers/make_cholesky_factorizer.hh:8-10) with a new provider value,
CholeskyProvider.GpuSupernodal.  Method names, argument meaning and failure behaviour follow the
reference; vectors are float64 CUDA tensors (column-major n x nrhs for multi-RHS calls).
Failures raise the same kinds of errors: a missing factorization or a non-SPD matrix raises
RuntimeError (MFSError subclasses), bad sizes raise ValueError.
"""
import enum

import numpy as np
import torch

from . import _lib
from ._lib import MFSError, NotPositiveDefiniteError, c_vp, ptr
from .ordering import NDOrdering, TemporalReuseStatistics, _stats
from .sparse import BlockCSCHessian, SuiteSparseMatrix, _dptr, _stream


def _as_scalar(H):
    """The scalar upper-CSC view a non-contiguous block matrix is factored through (its value
    array is already that layout, BlockCSCHessian.hh:52-97); other matrices unchanged."""
    if isinstance(H, BlockCSCHessian) and not H.hasContiguousBlocks():
        return H.toScalar()
    return H


def _pattern_key(H):
    nc = isinstance(H, BlockCSCHessian) and not H.hasContiguousBlocks()
    return ("noncontiguous" if nc else "scalar" if isinstance(H, SuiteSparseMatrix) else "block", H.n, H.nz)


class CholeskyProvider(enum.Enum):
    """CholeskyFactorizerBase.hh:19-24 plus the MI355X provider."""
    CHOLMOD = 0
    Catamari = 1
    CatamariNesdis = 2
    CatamariNesdisParallel = 3
    CatamariMetis = 4
    CatamariLegacy = 5
    CatamariAMD = 6
    CatamariScotch = 7
    CatamariAdaptive = 8
    PARDISO = 9
    Accelerate = 10
    CatamariNesdisReuse = 11
    CatamariNative = 12
    CHOLMODNative = 13
    GpuSupernodal = 100
    GpuSupernodalReuse = 101      # GpuSupernodal with temporal ND reuse, period 32 (as CatamariNesdisReuse)


class CholeskySys(enum.Enum):
    A = 0
    L = 1
    Lt = 2
    P = 3
    Pt = 4


class FactorizationType(enum.IntEnum):
    None_ = 0
    Symbolic = 1
    Numeric = 2


class GpuSupernodalFactorizer:
    """Supernodal LL^T on one MI355X: host ND ordering + symbolic analysis (once per pattern),
    device value injection, numeric factorization and solves (per Newton step)."""

    SCHUR_PRECISIONS = {None: _lib.MFS_PREC_FP32, "fp32": _lib.MFS_PREC_FP32, "bf16": _lib.MFS_PREC_BF16,
                        "bf16x3": _lib.MFS_PREC_BF16X3}

    def __init__(self, singlePrecision=False, schurUpdates=None):
        """singlePrecision: FP32 factor (after symmetric diagonal scaling) refined in FP64.
        schurUpdates (single precision only): "bf16" runs the Schur-complement updates on
        bfloat16 MFMA with FP32 accumulation, "bf16x3" with a hi + lo operand split."""
        self.singlePrecision = bool(singlePrecision)
        if schurUpdates not in self.SCHUR_PRECISIONS:
            raise ValueError(f"unknown Schur-update precision {schurUpdates!r}")
        if schurUpdates not in (None, "fp32") and not singlePrecision:
            raise ValueError("bfloat16 Schur updates need singlePrecision=True")
        self.schurUpdates = schurUpdates
        self._h = c_vp()
        prec = self.SCHUR_PRECISIONS[schurUpdates] if singlePrecision else _lib.MFS_PREC_FP64
        _lib.call("mfs_factorizer_create", prec, self._h)
        self._fixed = []
        self._n_full = 0
        self._n_red = 0
        self._Hpattern = None

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mfs_factorizer_destroy(self._h)
            self._h = None

    # -- state ----------------------------------------------------------------------------------
    def _ftype(self):
        t = _lib.c_i32()
        _lib.call("mfs_factorization_type", self._h, t)
        return t.value

    def hasFactorization(self, t=FactorizationType.Numeric):
        if isinstance(t, CholeskySys):
            t = FactorizationType.Numeric if t in (CholeskySys.A, CholeskySys.L, CholeskySys.Lt) \
                else FactorizationType.Symbolic
        return self._ftype() >= int(t)

    def hasNumericFactorization(self):
        return self.hasFactorization(FactorizationType.Numeric)

    def checkPosDef(self):
        return self._ftype() == FactorizationType.Numeric

    def assertFactorization(self, t=FactorizationType.Numeric):
        if not self.hasFactorization(t):
            raise RuntimeError(("Numeric" if t == FactorizationType.Numeric else "Symbolic") +
                               " factorization does not exist")

    def provider(self):
        # CatamariFactorizer.hh:138: the reuse provider whenever a reuse period is set
        return CholeskyProvider.GpuSupernodalReuse if self.temporalReusePeriod() else CholeskyProvider.GpuSupernodal

    def preferInPlaceSolve(self):
        return True

    def supportsPrePermutation(self):
        return True

    def m(self):
        return self._n_full

    def n(self):
        return self._n_full

    def m_reduced(self):
        self.assertFactorization(FactorizationType.Symbolic)
        return self._n_red

    def n_reduced(self):
        return self.m_reduced()

    def hasFixedVars(self):
        return len(self._fixed) > 0

    def getFixedVars(self):
        return list(self._fixed)

    def getFactorNNZ(self):
        v = _lib.c_i64()
        _lib.call("mfs_factor_nnz", self._h, v)
        return v.value

    def getFlopEstimate(self):
        v = _lib.c_double()
        _lib.call("mfs_flop_estimate", self._h, v)
        return v.value

    def flopBreakdown(self):
        a = np.zeros(3)
        _lib.call("mfs_flop_breakdown", self._h, ptr(a, _lib.c_double))
        return dict(total=a[0], potrf=a[1], unrelaxed=a[2])

    def getInversePermutation(self):
        out = np.empty(self.n_reduced(), dtype=np.int64)
        _lib.call("mfs_get_inverse_permutation", self._h, ptr(out, _lib.c_i64))
        return out

    def supernodeStats(self):
        v = [_lib.c_i64() for _ in range(4)]
        _lib.call("mfs_supernode_stats", self._h, *v)
        return dict(num_supernodes=v[0].value, num_levels=v[1].value, max_width=v[2].value,
                    max_degree=v[3].value)

    def numSuccessfulPivots(self):
        v = _lib.c_i64()
        _lib.call("mfs_num_successful_pivots", self._h, v)
        return v.value

    def lastTimings(self):
        a = np.zeros(4)
        _lib.call("mfs_last_timings", self._h, ptr(a, _lib.c_double))
        return dict(inject_ms=a[0], levels_ms=a[2], total_ms=a[3])

    # -- nested-dissection ordering and its temporal reuse ----------------------------------------
    def ndOrdering(self):
        """Separator tree of the most recent analysis (CholeskyFactorizerBase.hh:51-59): CParent of
        the postordered components, CMember of every unpermuted reduced block, blockSize 3; None
        when there is no analysis (or its ordering was not a dissection)."""
        nc, nv, bs = _lib.c_i64(), _lib.c_i64(), _lib.c_i64()
        if _lib.load().mfs_nd_ordering(self._h, nc, nv, bs, None, None) != _lib.MFS_OK:
            return None
        parent = np.empty(nc.value, np.int64)
        member = np.empty(nv.value, np.int64)
        _lib.call("mfs_nd_ordering", self._h, nc, nv, bs, ptr(parent, _lib.c_i64), ptr(member, _lib.c_i64))
        return NDOrdering(parent, member, bs.value)

    def setTemporalReusePeriod(self, period):
        """CatamariFactorizer.hh:84: P > 0 lets the next P symbolic analyses repair only the
        separator subtrees the new pattern violates; 0 disables reuse."""
        _lib.call("mfs_set_temporal_reuse_period", self._h, int(period))

    def temporalReusePeriod(self):
        v = _lib.c_i64()
        _lib.call("mfs_get_temporal_reuse_period", self._h, v)
        return v.value

    def resetTemporalReuse(self):
        _lib.call("mfs_reset_temporal_reuse", self._h)

    def resetSymbolicFactorizationReuse(self):
        """CholeskyFactorizerBase.hh:246 (CatamariFactorizer.hh:87): e.g. after SparsityLRU
        reports EXPIRED, the next analysis dissects from scratch."""
        self.resetTemporalReuse()

    def temporalReuseStatistics(self):
        return _stats("mfs_temporal_reuse_statistics", self._h)

    # -- factorization --------------------------------------------------------------------------
    def factorizeSymbolic(self, H, pinnedVars=()):
        """BlockCSCHessian (3x3 blocks, contiguous or not) or SuiteSparseMatrix (scalar, upper
        triangle) overloads.  Pins may split blocks: a partially pinned block keeps its free
        components.  A non-contiguous block matrix is analysed as the scalar matrix its values
        already form; the factor and its ordering are those of the same matrix in blocks."""
        pins = np.unique(np.asarray(pinnedVars, dtype=np.int64)) if len(pinnedVars) else np.zeros(0, np.int64)
        key = _pattern_key(H)
        S = _as_scalar(H)
        scalar = isinstance(S, SuiteSparseMatrix)
        if scalar:
            _lib.call("mfs_factorize_symbolic_scalar", self._h, ptr(S.Ap, _lib.c_i64), ptr(S.Ai, _lib.c_i64),
                      S.n, ptr(pins, _lib.c_i64), len(pins))
        else:
            _lib.call("mfs_factorize_symbolic", self._h, ptr(H.Ap, _lib.c_i64), ptr(H.Ai, _lib.c_i64),
                      H.n, 3, ptr(pins, _lib.c_i64), len(pins))
        self._fixed = [int(p) for p in pins]
        self._n_full = S.n if scalar else 3 * H.n
        self._n_red = self._n_full - len(self._fixed)
        self._Hpattern = key

    def _check_pattern(self, H):
        if self._Hpattern is not None and self._Hpattern != _pattern_key(H):
            raise ValueError("matrix pattern or value layout differs from the symbolic factorization")
        if H.isSparsityOnly():
            raise ValueError("cannot factorize a sparsity-only matrix")

    def _numeric(self, A, sigma, Bp):
        scalar = isinstance(_as_scalar(A), SuiteSparseMatrix)
        name = "mfs_factorize_numeric_scalar" if scalar else "mfs_factorize_numeric"
        _lib.call(name, self._h, _dptr(A.Ax), float(sigma), Bp, _stream())

    def factorizeNumeric(self, H, isInTryCatch=False):
        self.assertFactorization(FactorizationType.Symbolic)
        self._check_pattern(H)
        self._numeric(H, 0.0, None)

    def factorizeNumericWithShift(self, A, sigma, B=None, isInTryCatch=False):
        self.assertFactorization(FactorizationType.Symbolic)
        self._check_pattern(A)
        Bp = None
        if B is not None:
            if (_pattern_key(B) != _pattern_key(A) or not np.array_equal(B.Ap, A.Ap)
                    or not np.array_equal(B.Ai, A.Ai)):
                raise ValueError("Shift matrix must have the same sparsity and value layout as A")
            Bp = _dptr(B.Ax)
        self._numeric(A, sigma, Bp)

    def factorize(self, H, fixedVars=()):
        self.factorizeSymbolic(H, fixedVars)
        self.factorizeNumeric(H)

    def clearFactors(self):
        _lib.call("mfs_clear_factors", self._h)

    def stashFactorization(self):
        _lib.call("mfs_stash_factorization", self._h)

    def hasStashedFactorization(self):
        v = _lib.c_i32()
        _lib.call("mfs_has_stashed_factorization", self._h, v)
        return bool(v.value)

    def swapStashedFactorization(self):
        _lib.call("mfs_swap_stashed_factorization", self._h)

    def clearStashedFactorization(self):
        _lib.call("mfs_clear_stashed_factorization", self._h)

    # -- solves ---------------------------------------------------------------------------------
    def _colmajor(self, B, rows):
        if B.dim() == 1:
            if B.numel() != rows:
                raise RuntimeError("CholeskyFactorizerBase::solve: incorrect b size")
            return B.contiguous(), 1
        if B.shape[0] != rows:
            raise RuntimeError("Incorrect RHS size")
        if B.shape[1] < 1:
            raise RuntimeError("Must specify at least one rhs.")
        return B.t().contiguous(), B.shape[1]

    def solve(self, b, sys=CholeskySys.A):
        if sys != CholeskySys.A:
            raise RuntimeError("Alternative CholeskySys not yet wrapped")
        self.assertFactorization(FactorizationType.Numeric)
        bc, nrhs = self._colmajor(b, self.m())
        x = torch.empty_like(bc)
        _lib.call("mfs_solve", self._h, _dptr(bc), _dptr(x), nrhs, _stream())
        return x if b.dim() == 1 else x.t()

    def solveMultiRHS(self, B):
        self.assertFactorization(FactorizationType.Numeric)
        if B.dim() != 2:
            raise RuntimeError("Incorrect RHS size")
        return self.solve(B)

    def solveRawReduced(self, b, x, sys=CholeskySys.A, alreadyPermuted=False):
        if sys != CholeskySys.A:
            raise RuntimeError("Alternative CholeskySys not yet wrapped")
        self.assertFactorization(FactorizationType.Numeric)
        nrhs = 1 if b.dim() == 1 else b.shape[0]
        _lib.call("mfs_solve_raw_reduced", self._h, _dptr(b), _dptr(x), nrhs, int(bool(alreadyPermuted)),
                  _stream())

    def solveRawReducedInPlace(self, bx, sys=CholeskySys.A, alreadyPermuted=False):
        self.solveRawReduced(bx, bx, sys, alreadyPermuted)

    def removeFixedEntries(self, x, permute=False):
        self.assertFactorization(FactorizationType.Symbolic)
        if x.numel() != self.n():
            raise RuntimeError("Invalid x size")
        xr = torch.empty(self.n_reduced(), dtype=torch.float64, device=x.device)
        _lib.call("mfs_remove_fixed_entries", self._h, _dptr(x.contiguous()), _dptr(xr), 1,
                  int(bool(permute)), _stream())
        return xr

    def extractFullSolution(self, xr, permute=False):
        self.assertFactorization(FactorizationType.Symbolic)
        if xr.numel() != self.n_reduced():
            raise RuntimeError("Invalid xReduced size")
        x = torch.empty(self.n(), dtype=torch.float64, device=xr.device)
        _lib.call("mfs_extract_full_solution", self._h, _dptr(xr.contiguous()), _dptr(x), 1,
                  int(bool(permute)), _stream())
        return x

    # -- extras: SpMV, refinement, factor export ------------------------------------------------
    def apply(self, H, x):
        self._check_pattern(H)
        y = torch.empty_like(x)
        _lib.call("mfs_spmv", self._h, _dptr(H.Ax), _dptr(x.contiguous()), _dptr(y), _stream())
        return y

    def solveRefined(self, H, b, tol=1e-10, maxIters=10):
        self._check_pattern(H)
        x = torch.empty_like(b)
        it, rr = _lib.c_i32(), _lib.c_double()
        _lib.call("mfs_solve_refined", self._h, _dptr(H.Ax), _dptr(b.contiguous()), _dptr(x), float(tol),
                  int(maxIters), it, rr, _stream())
        return x, it.value, rr.value

    def solveRefinedSegments(self, H, b, segments, tol=1e-10, maxIters=10):
        """Refinement until every segment [segments[s], segments[s+1]) of the full-length vector
        (one mesh of a batch) reaches ||b_s - (A x)_s|| / ||b_s|| <= tol.
        Returns (x, iterations, max relative residual, per-segment relative residuals)."""
        self._check_pattern(H)
        seg = np.ascontiguousarray(segments, np.int64)
        x = torch.empty_like(b)
        it, rr = _lib.c_i32(), _lib.c_double()
        per = np.zeros(len(seg) - 1)
        _lib.call("mfs_solve_refined_segments", self._h, _dptr(H.Ax), _dptr(b.contiguous()), _dptr(x), len(seg) - 1,
                  ptr(seg, _lib.c_i64), float(tol), int(maxIters), it, rr, ptr(per, _lib.c_double), _stream())
        return x, it.value, rr.value, per

    def setGraphCapture(self, on=True):
        """Capture the numeric factorization in a hipGraph and replay it while the value array,
        shift and shift matrix stay the same."""
        _lib.call("mfs_factorizer_set_graph", self._h, int(bool(on)))

    def scheduleStats(self):
        v = np.zeros(9, np.int64)
        _lib.call("mfs_schedule_stats", self._h, ptr(v, _lib.c_i64))
        keys = ["fused_tasks", "fused_launches", "tiled_supernodes", "potrf_launches", "gemm_launches",
                "extend_launches", "levels", "graph_nodes", "panel_launches"]
        return {k: int(x) for k, x in zip(keys, v)}

    def autotune(self, H, pinnedVars=(), fuse=(192, 0, 384), tiles=(128, 64), reps=3):
        """Choose the schedule of H's numeric factorization by timing on one ordering: first the
        fused level kernel's bound among `fuse` (persistent kernel for supernodes one tile wide
        with at most that many structure rows; 0: none), then per supernode size bucket (front
        W + D < 512, < 2048, larger) the largest GEMM tile among `tiles` (128 or 64).  The first
        candidates are the start and a candidate must be 2 % faster to replace the incumbent.
        Leaves a symbolic factorization with the chosen schedule.  H has contiguous blocks."""
        if not isinstance(H, BlockCSCHessian) or not H.hasContiguousBlocks():
            raise ValueError("autotune takes a block matrix with contiguous blocks")
        pins = np.unique(np.asarray(pinnedVars, dtype=np.int64)) if len(pinnedVars) else np.zeros(0, np.int64)
        fz = np.ascontiguousarray(fuse, np.int32)
        tz = np.ascontiguousarray(tiles, np.int32)
        ms = np.zeros(len(fz) + 3 * len(tz))
        best = np.zeros(4, np.int32)
        _lib.call("mfs_autotune", self._h, ptr(H.Ap, _lib.c_i64), ptr(H.Ai, _lib.c_i64), H.n, 3,
                  ptr(pins, _lib.c_i64), len(pins), _dptr(H.Ax), ptr(fz, _lib.c_i32), len(fz), ptr(tz, _lib.c_i32),
                  len(tz), int(reps), ptr(ms, _lib.c_double), ptr(best, _lib.c_i32), _stream())
        self._fixed = [int(p) for p in pins]
        self._n_full = 3 * H.n
        self._n_red = 3 * H.n - len(self._fixed)
        self._Hpattern = _pattern_key(H)
        nf = len(fz)
        table = {"fuse": {int(a): float(ms[i]) for i, a in enumerate(fz)},
                 "bucket_tiles": [{int(t): float(ms[nf + b * len(tz) + j]) for j, t in enumerate(tz)} for b in range(3)]}
        return {"fuse_degree": int(best[0]), "bucket_tiles": [int(x) for x in best[1:]], "ms": table}

    def exportFactorCSC(self):
        n = self.n_reduced()
        nnz = self.getFactorNNZ()
        Lp = np.empty(n + 1, np.int64)
        Li = np.empty(nnz, np.int64)
        Lx = np.empty(nnz, np.float64)
        _lib.call("mfs_factor_export_csc", self._h, ptr(Lp, _lib.c_i64), ptr(Li, _lib.c_i64),
                  ptr(Lx, _lib.c_double))
        return Lp, Li, Lx


def make_cholesky_factorizer(provider=CholeskyProvider.GpuSupernodal, singlePrecision=False, schurUpdates=None):
    """make_cholesky_factorizer.cc:14-87 with the GPU provider; other providers are not built.
    schurUpdates selects the BF16 Schur-update variants of the single-precision factor."""
    if provider == CholeskyProvider.CatamariLegacy and singlePrecision:
        raise ValueError("Single precision is unavailable for legacy Catamari")
    if provider not in (CholeskyProvider.GpuSupernodal, CholeskyProvider.GpuSupernodalReuse):
        raise RuntimeError(f"provider {provider.name} is not available in this build")
    f = GpuSupernodalFactorizer(singlePrecision, schurUpdates)
    if provider == CholeskyProvider.GpuSupernodalReuse:
        f.setTemporalReusePeriod(32)       # make_cholesky_factorizer.cc:68-70
    return f


__all__ = ["CholeskyProvider", "CholeskySys", "FactorizationType", "GpuSupernodalFactorizer",
           "make_cholesky_factorizer", "MFSError", "NotPositiveDefiniteError", "NDOrdering",
           "TemporalReuseStatistics"]
