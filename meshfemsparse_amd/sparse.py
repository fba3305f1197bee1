"""Host mirror of the reference's block-CSC Hessian and assembler interfaces.

BlockCSCHessian  <- BlockCSCHessian<OptimizationVarStructure<3>, ContiguousBlocks=true>
                    (src/lib/MeshFEMSparse/BlockCSCHessian.hh:560-571, 798-875): upper block
                    pattern Ap/Ai (int64, host) and the value array Ax (9 doubles per block,
                    column-major 3x3, diagonal block last per column, upper triangle meaningful),
                    held in device memory.
SystemAssembler  <- SystemAssembler<3> (src/lib/MeshFEMSparse/SystemAssembler.hh):
                    blockSparsityPattern (:251-465) and assembleHessian (:559-565), which
                    accumulates into H.Ax and allocates it when H is sparsity-only.
"""
import numpy as np
import torch

from . import _lib
from ._lib import MFS_HE_FULL12, MFS_HE_UPPER78, P_i32, P_i64, c_i64, c_vp, ptr


def _stream():
    return c_vp(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA (HIP) torch tensor")
    if t.dtype != torch.float64:
        raise TypeError("expected a float64 tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return c_vp(t.data_ptr())


class BlockCSCHessian:
    """Uniform 3x3-block upper-triangle CSC matrix with contiguous blocks."""

    blockSize = 3

    def __init__(self, Ap, Ai, device="cuda"):
        self.Ap = np.ascontiguousarray(Ap, dtype=np.int64)
        self.Ai = np.ascontiguousarray(Ai, dtype=np.int64)
        self.n = len(self.Ap) - 1
        self.m = self.n
        self.nz = int(self.Ap[-1])
        self.Ax = None
        self.device = device

    # -- BlockCSCHessianBase interface --------------------------------------------------------
    def numScalarCols(self):
        return 3 * self.n

    def scalarNNZ(self):
        return 9 * self.nz

    def isSparsityOnly(self):
        return self.Ax is None

    def maxBlockSize(self):
        return 3

    def uniformBlockSize(self):
        return True

    def setZero(self):
        """BlockCSCHessian.hh:638-649: allocate Ax on first use, otherwise zero it."""
        if self.Ax is None or self.Ax.numel() != self.scalarNNZ():
            self.Ax = torch.zeros(self.scalarNNZ(), dtype=torch.float64, device=self.device)
        else:
            self.Ax.zero_()

    def data(self):
        return self.Ax

    def diagBlockLoc(self, j):
        return 9 * (int(self.Ap[j + 1]) - 1)

    def trace(self):
        """Sum of the diagonal scalar entries (BlockCSCHessianBase::trace)."""
        Ax = self.Ax.detach().cpu().numpy() if isinstance(self.Ax, torch.Tensor) else np.asarray(self.Ax)
        d = 9 * (self.Ap[1:] - 1)
        return float(Ax[d].sum() + Ax[d + 4].sum() + Ax[d + 8].sum())

    def dumpBinaryToFile(self, path):
        """BlockCSCHessianBase::dumpBinaryToStream (BlockCSCHessian.cc:11-40): the reference's
        binary format, non-contiguous layout on disk; a sparsity-only matrix writes no values."""
        Ax = None
        if self.Ax is not None:
            Ax = np.ascontiguousarray(self.Ax.detach().cpu().numpy() if isinstance(self.Ax, torch.Tensor)
                                      else self.Ax, np.float64)
        _lib.call("mfs_blockcsc_write", str(path).encode(), ptr(self.Ap, _lib.c_i64), ptr(self.Ai, _lib.c_i64),
                  self.n, None if Ax is None else ptr(Ax, _lib.c_double))

    @classmethod
    def constructFromBinaryFile(cls, path, device="cuda"):
        """BlockCSCHessianBase::constructFromBinaryStream (BlockCSCHessian.cc:88-103), returning the
        contiguous-block layout with values on `device` (None for a sparsity-only file)."""
        nb, nz, snz = _lib.c_i64(), _lib.c_i64(), _lib.c_i64()
        p = str(path).encode()
        _lib.call("mfs_blockcsc_read", p, nb, nz, snz, None, None, None)
        Ap = np.zeros(nb.value + 1, np.int64)
        Ai = np.zeros(nz.value, np.int64)
        Ax = np.zeros(9 * nz.value) if snz.value else None
        _lib.call("mfs_blockcsc_read", p, nb, nz, snz, ptr(Ap, _lib.c_i64), ptr(Ai, _lib.c_i64),
                  None if Ax is None else ptr(Ax, _lib.c_double))
        H = cls(Ap, Ai, device)
        if Ax is not None:
            H.Ax = torch.from_numpy(Ax).to(device)
        return H

    def toDense(self):
        """Dense symmetric scalar matrix (host, float64); for small tests."""
        Ax = self.Ax.detach().cpu().numpy() if isinstance(self.Ax, torch.Tensor) else np.asarray(self.Ax)
        A = np.zeros((3 * self.n, 3 * self.n))
        for j in range(self.n):
            for k in range(self.Ap[j], self.Ap[j + 1]):
                i = self.Ai[k]
                B = Ax[9 * k:9 * k + 9].reshape(3, 3).T    # column-major block
                if i == j:
                    U = np.triu(B)
                    A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = U + np.triu(U, 1).T
                else:
                    A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = B
                    A[3 * j:3 * j + 3, 3 * i:3 * i + 3] = B.T
        return A


class SuiteSparseMatrix:
    """Scalar symmetric matrix stored as its upper triangle in CSC: the reference's
    SuiteSparseMatrix with UPPER_TRIANGLE symmetry, the argument type of the scalar
    CholeskyFactorizerBase overloads (CholeskyFactorizerBase.hh:95-113).  Ap/Ai are int64 host
    arrays (rows sorted, row <= column); Ax holds the nz values on the device."""

    blockSize = 1

    def __init__(self, n, Ap, Ai, Ax=None, device="cuda"):
        self.n = self.m = int(n)
        self.Ap = np.ascontiguousarray(Ap, dtype=np.int64)
        self.Ai = np.ascontiguousarray(Ai, dtype=np.int64)
        if len(self.Ap) != self.n + 1:
            raise ValueError("Ap must have n + 1 entries")
        self.nz = int(self.Ap[-1])
        self.Ax = Ax
        self.device = device

    @classmethod
    def fromTriplets(cls, n, rows, cols, vals, device="cuda"):
        """TripletMatrix -> SuiteSparseMatrix(UPPER_TRIANGLE): entries with row > col are
        mirrored into the upper triangle, duplicates summed."""
        import torch
        r = np.asarray(rows, np.int64)
        c = np.asarray(cols, np.int64)
        i, j = np.minimum(r, c), np.maximum(r, c)
        key = j * n + i
        uk, inv = np.unique(key, return_inverse=True)
        v = np.zeros(len(uk))
        np.add.at(v, inv, np.asarray(vals, np.float64))
        jj, ii = uk // n, uk % n
        Ap = np.zeros(n + 1, np.int64)
        np.add.at(Ap, jj + 1, 1)
        Ap = np.cumsum(Ap)
        return cls(n, Ap, ii, torch.from_numpy(v).to(device), device)

    def isSparsityOnly(self):
        return self.Ax is None

    def toDense(self):
        A = np.zeros((self.n, self.n))
        Ax = self.Ax.cpu().numpy()
        for j in range(self.n):
            for k in range(self.Ap[j], self.Ap[j + 1]):
                A[self.Ai[k], j] = A[j, self.Ai[k]] = Ax[k]
        return A


class SystemAssembler:
    """SystemAssembler<3> over `numBlockVars` 3-vector variables."""

    def __init__(self, numBlockVars):
        self.nb = int(numBlockVars)
        self._asm = {}

    def numBlockVars(self):
        return self.nb

    def numVars(self):
        return 3 * self.nb

    def blockSparsityPattern(self, elem_nodes):
        """SystemAssembler<3>::blockSparsityPattern (SystemAssembler.hh:251-465).  A CUDA int32
        tensor of element nodes is processed on the device (mfs_pattern_build_device), anything
        else on the host; both give the same pattern."""
        lib = _lib.load()
        h = c_vp()
        if isinstance(elem_nodes, torch.Tensor) and elem_nodes.is_cuda:
            en_d = elem_nodes.to(torch.int32).contiguous()
            ne, npe = en_d.shape
            s = c_vp(torch.cuda.current_stream(en_d.device).cuda_stream)
            _lib.check(lib.mfs_pattern_build_device(c_vp(en_d.data_ptr()), ne, npe, self.nb, s, h))
        else:
            en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
            ne, npe = en.shape
            _lib.check(lib.mfs_pattern_build(ptr(en, _lib.c_i32), ne, npe, self.nb, h))
        try:
            nb, nz = c_i64(), c_i64()
            Ap, Ai = P_i64(), P_i64()
            _lib.check(lib.mfs_pattern_get(h, nb, nz, Ap, Ai))
            Apn = np.ctypeslib.as_array(Ap, shape=(self.nb + 1,)).copy()
            Ain = (np.ctypeslib.as_array(Ai, shape=(nz.value,)).copy() if nz.value > 0
                   else np.zeros(0, np.int64))
        finally:
            lib.mfs_pattern_destroy(h)
        return BlockCSCHessian(Apn, Ain)

    def _assembler(self, H, en):
        key = (id(H), H.nz, en.ctypes.data, en.shape)
        a = self._asm.get(key)
        if a is None:
            a = DeviceAssembler(H, en)
            self._asm = {key: a}
        return a

    def assembleHessian(self, H, elem_nodes, He, layout=MFS_HE_FULL12):
        """Accumulate element Hessians He (device, ne x 144 or ne x 78 float64) into H.Ax."""
        if H is None:
            raise RuntimeError("Attempted to assemble into a BorderedSparseHessian with no block sparsity pattern.")
        if H.isSparsityOnly():
            H.setZero()
        en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        self._assembler(H, en).assemble(He, H.Ax, layout)


class DeviceAssembler:
    """Slot map of (pattern, connectivity), resident on the device."""

    def __init__(self, H, elem_nodes):
        en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        self.ne, npe = en.shape
        self._h = c_vp()
        _lib.call("mfs_assembler_create", ptr(H.Ap, _lib.c_i64), ptr(H.Ai, _lib.c_i64), H.n,
                  ptr(en, _lib.c_i32), self.ne, npe, self._h)

    def assemble(self, He, Ax, layout=MFS_HE_FULL12):
        per = 144 if layout == MFS_HE_FULL12 else 78
        if He.numel() != per * self.ne:
            raise ValueError(f"He must hold {per} doubles per element")
        _lib.call("mfs_assemble_hessian", self._h, _dptr(He), layout, _dptr(Ax), _stream())

    def zero(self, Ax):
        _lib.call("mfs_assembler_zero", self._h, _dptr(Ax), _stream())

    def neohookean(self, X_rest, X, E=1e5, nu=0.3, E_per_elem=None, nu_per_elem=None, layout=MFS_HE_FULL12,
                   hessians=True, gradients=True):
        """Element Hessians and per-element gradients of the compressible neo-Hookean energy on the
        device (float64 CUDA tensors; positions num_nodes x 3).  Returns (He, g12)."""
        import torch
        dev = X.device
        per = 144 if layout == MFS_HE_FULL12 else 78
        He = torch.empty(per * self.ne, dtype=torch.float64, device=dev) if hessians else None
        g12 = torch.empty(12 * self.ne, dtype=torch.float64, device=dev) if gradients else None
        opt = lambda t: None if t is None else _dptr(t.contiguous())
        _lib.call("mfs_neohookean_elements", self._h, _dptr(X_rest.contiguous()), _dptr(X.contiguous()),
                  opt(E_per_elem), opt(nu_per_elem), float(E), float(nu), opt(He), layout, opt(g12), _stream())
        return He, g12

    def assembleGradient(self, g12, g):
        """g += sum_e scatter(g12_e) (SystemAssembler::assembleGradient accumulate semantics)."""
        _lib.call("mfs_assemble_gradient", self._h, _dptr(g12), _dptr(g), _stream())

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mfs_assembler_destroy(self._h)
            self._h = None
