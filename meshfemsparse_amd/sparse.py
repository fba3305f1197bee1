"""Host mirror of the reference's block-CSC Hessian and assembler interfaces.

BlockCSCHessian  <- BlockCSCHessian<OptimizationVarStructure<3>, ContiguousBlocks>
                    (src/lib/MeshFEMSparse/BlockCSCHessian.hh:52-97, 560-571, 798-875): upper
                    block pattern Ap/Ai (int64, host) and the value array Ax in device memory.
                    Contiguous blocks (the default): 9 doubles per block, column-major 3x3,
                    diagonal block last per column, upper triangle meaningful.  Non-contiguous:
                    Ax is the scalar upper-triangle CSC of the 3n x 3n matrix (diagonal blocks
                    store their upper triangle only), as StoreFullDiagonalBlocks = ContiguousBlocks
                    makes it in the reference.
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
    """Uniform 3x3-block upper-triangle CSC matrix, contiguous or non-contiguous blocks."""

    blockSize = 3

    def __init__(self, Ap, Ai, device="cuda", contiguous=True):
        self.Ap = np.ascontiguousarray(Ap, dtype=np.int64)
        self.Ai = np.ascontiguousarray(Ai, dtype=np.int64)
        self.n = len(self.Ap) - 1
        self.m = self.n
        self.nz = int(self.Ap[-1])
        self.Ax = None
        self.device = device
        self.contiguousBlocks = bool(contiguous)
        self._layout = None
        if not self.contiguousBlocks:
            self._check_diagonal_blocks()

    # -- BlockCSCHessianBase interface --------------------------------------------------------
    def numScalarCols(self):
        return 3 * self.n

    def scalarNNZ(self):
        """9 values per block; the non-contiguous layout drops the strict lower triangle of the
        diagonal blocks (BlockCSCHessian.hh:74-80)."""
        return 9 * self.nz if self.contiguousBlocks else 9 * self.nz - 3 * self.n

    def hasContiguousBlocks(self):
        return self.contiguousBlocks

    def isSparsityOnly(self):
        return self.Ax is None

    def maxBlockSize(self):
        return 3

    def uniformBlockSize(self):
        return True

    def _check_diagonal_blocks(self):
        last = self.Ap[1:] - 1
        if self.n and ((np.diff(self.Ap) <= 0).any() or (self.Ai[last] != np.arange(self.n)).any()):
            raise ValueError("the non-contiguous layout needs every diagonal block (BlockCSCHessian.hh:77)")

    def col_nnz(self, bj):
        return int(self.Ap[bj + 1] - self.Ap[bj])

    def scalarColStride(self, bj):
        """Entries of the first scalar column of block column bj (BlockCSCHessian.hh:64-71);
        each later scalar column of bj holds one more (the diagonal block's upper triangle)."""
        if self.contiguousBlocks:
            return 3
        return 3 * (self.col_nnz(bj) - 1) + 1

    def scalarOffsetForColumn(self, bj):
        """BlockCSCHessian.hh:73-80."""
        return 9 * int(self.Ap[bj]) - (0 if self.contiguousBlocks else 3 * bj)

    def findEntry(self, bi, bj):
        lo, hi = int(self.Ap[bj]), int(self.Ap[bj + 1])
        k = lo + int(np.searchsorted(self.Ai[lo:hi], bi))
        if k >= hi or self.Ai[k] != bi:
            raise KeyError(f"block ({bi}, {bj}) is not in the pattern")
        return k

    def scalarOffsetWithinColumn(self, bi, bj):
        """BlockCSCHessian.hh:88-92."""
        idx = self.findEntry(bi, bj) - int(self.Ap[bj])
        return (9 if self.contiguousBlocks else 3) * idx

    def locForBlock(self, bi, bj):
        """Position in Ax of entry (0, 0) of block (bi, bj) (BlockCSCHessian.hh:83-86)."""
        return self.scalarOffsetForColumn(bj) + self.scalarOffsetWithinColumn(bi, bj)

    def _layout_maps(self):
        """(pos, Sp, Si): pos[9 k + 3 c + r] = position of block k's entry (r, c) in the
        non-contiguous value array (-1 for the strict lower triangle of a diagonal block), and
        the scalar upper CSC pattern (Sp, Si) that array follows."""
        if self._layout is None:
            self._check_diagonal_blocks()
            n, Ap, Ai = self.n, self.Ap, self.Ai
            ne = np.diff(Ap)
            col = np.repeat(np.arange(n), ne)                   # block column of every block
            idx = np.arange(self.nz) - Ap[col]
            base = 9 * Ap[col] - 3 * col + 3 * idx
            stride = 3 * (ne[col] - 1)
            diag = Ai == col
            pos = np.empty((self.nz, 3, 3), np.int64)           # [k][c][r]
            for c in range(3):
                for r in range(3):
                    pos[:, c, r] = np.where(diag & (r > c), -1, base + c * stride + c * (c + 1) // 2 + r)
            pos = pos.reshape(-1)
            Sp = np.zeros(3 * n + 1, np.int64)
            lens = 3 * (ne[:, None] - 1) + np.arange(1, 4)[None, :]
            Sp[1:] = np.cumsum(lens.reshape(-1))
            Si = np.empty(int(Sp[-1]), np.int64)
            valid = pos >= 0
            rows = (3 * Ai[:, None, None] + np.arange(3)[None, None, :]).repeat(3, axis=1).reshape(-1)
            Si[pos[valid]] = rows[valid]
            self._layout = (pos, Sp, Si)
        return self._layout

    def cloneWithLayout(self, contiguous):
        """BlockCSCHessian::cloneWithLayout: the same matrix in the other value layout (the
        strict lower triangle of diagonal blocks becomes zero going to contiguous blocks)."""
        H = BlockCSCHessian(self.Ap, self.Ai, self.device, contiguous)
        if contiguous == self.contiguousBlocks or self.Ax is None:
            H.Ax = None if self.Ax is None else self.Ax.clone()
            return H
        pos = self._layout_maps()[0]
        H._layout = self._layout
        valid = np.nonzero(pos >= 0)[0]
        dev = self.Ax.device
        src_c = torch.from_numpy(valid).to(dev)
        dst_nc = torch.from_numpy(pos[valid]).to(dev)
        if contiguous:
            H.Ax = torch.zeros(9 * self.nz, dtype=self.Ax.dtype, device=dev)
            H.Ax[src_c] = self.Ax[dst_nc]
        else:
            H.Ax = torch.empty(9 * self.nz - 3 * self.n, dtype=self.Ax.dtype, device=dev)
            H.Ax[dst_nc] = self.Ax[src_c]
        return H

    def cloneWithContiguousBlocks(self):
        return self.cloneWithLayout(True)

    def cloneWithNoncontiguousBlocks(self):
        return self.cloneWithLayout(False)

    def toScalar(self):
        """The scalar upper-triangle CSC matrix (BlockCSCHessianBase::toScalar).  For
        non-contiguous blocks the value array is shared, not copied."""
        _, Sp, Si = self._layout_maps()
        if self.contiguousBlocks:
            Ax = None if self.Ax is None else self.cloneWithLayout(False).Ax
        else:
            Ax = self.Ax
        return SuiteSparseMatrix(3 * self.n, Sp, Si, Ax, self.device)

    def setZero(self):
        """BlockCSCHessian.hh:638-649: allocate Ax on first use, otherwise zero it."""
        if self.Ax is None or self.Ax.numel() != self.scalarNNZ():
            self.Ax = torch.zeros(self.scalarNNZ(), dtype=torch.float64, device=self.device)
        else:
            self.Ax.zero_()

    def data(self):
        return self.Ax

    def diagBlockLoc(self, j):
        return self.locForBlock(j, j)

    def diagEntryPositions(self):
        """Positions in Ax of the 3n diagonal scalar entries, in scalar order."""
        k = self.Ap[1:] - 1                                     # diagonal block of each column
        cont = 9 * k[:, None] + np.array([0, 4, 8])[None, :]
        if self.contiguousBlocks:
            return cont.reshape(-1)
        return self._layout_maps()[0][cont.reshape(-1)]

    def _host_contiguous(self):
        H = self if self.contiguousBlocks else self.cloneWithLayout(True)
        return H.Ax.detach().cpu().numpy() if isinstance(H.Ax, torch.Tensor) else np.asarray(H.Ax)

    def trace(self):
        """Sum of the diagonal scalar entries (BlockCSCHessianBase::trace)."""
        Ax = self.Ax.detach().cpu().numpy() if isinstance(self.Ax, torch.Tensor) else np.asarray(self.Ax)
        return float(Ax[self.diagEntryPositions()].sum())

    def dumpBinaryToFile(self, path):
        """BlockCSCHessianBase::dumpBinaryToStream (BlockCSCHessian.cc:11-40): the reference's
        binary format, non-contiguous layout on disk; a sparsity-only matrix writes no values."""
        Ax = None
        if self.Ax is not None:
            Ax = np.ascontiguousarray(self._host_contiguous(), np.float64)
        _lib.call("mfs_blockcsc_write", str(path).encode(), ptr(self.Ap, _lib.c_i64), ptr(self.Ai, _lib.c_i64),
                  self.n, None if Ax is None else ptr(Ax, _lib.c_double))

    @classmethod
    def constructFromBinaryFile(cls, path, device="cuda", contiguous=True):
        """BlockCSCHessianBase::constructFromBinaryStream (BlockCSCHessian.cc:88-103), returning
        the contiguous-block layout (ContiguousBlocksDefault) unless `contiguous` is False, with
        values on `device` (None for a sparsity-only file)."""
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
        return H if contiguous else H.cloneWithLayout(False)

    def toDense(self):
        """Dense symmetric scalar matrix (host, float64); for small tests."""
        Ax = self._host_contiguous()
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
        """Accumulate element Hessians He (device, ne x 144 or ne x 78 float64) into H.Ax, in
        H's value layout."""
        if H is None:
            raise RuntimeError("Attempted to assemble into a BorderedSparseHessian with no block sparsity pattern.")
        if H.isSparsityOnly():
            H.setZero()
        en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        self._assembler(H, en).assemble(He, H.Ax, layout, H.hasContiguousBlocks())


class DeviceAssembler:
    """Slot map of (pattern, connectivity), resident on the device."""

    def __init__(self, H, elem_nodes):
        en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        self.ne, npe = en.shape
        self._h = c_vp()
        _lib.call("mfs_assembler_create", ptr(H.Ap, _lib.c_i64), ptr(H.Ai, _lib.c_i64), H.n,
                  ptr(en, _lib.c_i32), self.ne, npe, self._h)

    def assemble(self, He, Ax, layout=MFS_HE_FULL12, contiguous=True):
        """Ax += the assembled element Hessians; `contiguous` selects the value layout of Ax
        (9 per block, or the non-contiguous scalar upper CSC)."""
        per = 144 if layout == MFS_HE_FULL12 else 78
        if He.numel() != per * self.ne:
            raise ValueError(f"He must hold {per} doubles per element")
        name = "mfs_assemble_hessian" if contiguous else "mfs_assemble_hessian_noncontiguous"
        _lib.call(name, self._h, _dptr(He), layout, _dptr(Ax), _stream())

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
