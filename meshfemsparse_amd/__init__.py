"""MI355X-native per-Newton-step linear-solve pipeline of MeshFEMSparse: block-CSC Hessian
assembly, supernodal Cholesky numeric factorization and supernodal solves, as hand-written gfx950
HIP kernels behind a C ABI (include/meshfemsparse_amd.h), with host mirrors of the reference's
SystemAssembler / BlockCSCHessian / CholeskyFactorizerBase interfaces."""
from . import _lib
from ._lib import MFSError, NotPositiveDefiniteError, UnsupportedError, load
from .sparse import BlockCSCHessian, SuiteSparseMatrix, SystemAssembler, DeviceAssembler
from .solvers import (CholeskyProvider, CholeskySys, FactorizationType, GpuSupernodalFactorizer,
                      make_cholesky_factorizer)
from .bordered import BorderedSparseHessian, BorderedSparseFactorization
from .ordering import NDOrdering, NestedDissection, TemporalReuseStatistics

__all__ = ["BlockCSCHessian", "SuiteSparseMatrix", "SystemAssembler", "DeviceAssembler", "CholeskyProvider", "CholeskySys",
           "FactorizationType", "GpuSupernodalFactorizer", "make_cholesky_factorizer", "BorderedSparseHessian",
           "BorderedSparseFactorization", "NDOrdering", "NestedDissection", "TemporalReuseStatistics", "MFSError",
           "NotPositiveDefiniteError", "UnsupportedError", "load"]
