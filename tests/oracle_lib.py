"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).  Test infrastructure only."""
import ctypes
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_PATH = os.path.join(_ROOT, "oracle", "liboracle.so")
_lib = None

i32, i64, f64, u8 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_uint8


def _p(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError(f"{_PATH} missing; run __graft_entry__.build()")
        # OpenMP workers of the tree levels sleep instead of spinning while the BLAS threads run
        # the top levels (read by libgomp when the oracle library loads it)
        os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
        L = ctypes.CDLL(_PATH)
        P = ctypes.POINTER
        L.orc_cube_mesh.argtypes = [ctypes.c_int, P(i32), P(f64)]
        L.orc_perturb.argtypes = [ctypes.c_int, ctypes.c_uint64, f64, P(f64)]
        L.orc_neohookean_energy.argtypes = [P(f64), P(f64), P(i32), i64, P(f64), P(f64), f64, f64]
        L.orc_neohookean_energy.restype = f64
        L.orc_pattern.argtypes = [P(i32), i64, ctypes.c_int, i64, P(i64), P(i64), i64]
        L.orc_pattern.restype = i64
        L.orc_assemble.argtypes = [P(i64), P(i64), i64, P(i32), i64, P(f64), P(f64)]
        L.orc_row_col_removal.argtypes = [P(i64), P(i64), i64, P(u8), P(i64), P(i64), P(i64), P(i64)]
        L.orc_row_col_removal.restype = i64
        L.orc_spmv.argtypes = [P(i64), P(i64), i64, P(f64), P(f64), P(f64)]
        L.orc_permuted_lower.argtypes = [P(i64), P(i64), i64, P(f64), P(i64), P(i32), i64, f64, P(f64),
                                         P(i64), P(i64), P(f64), i64]
        L.orc_permuted_lower.restype = i64
        L.orc_chol.argtypes = [i64, P(i64), P(i64), P(f64), P(i64), P(i64), P(f64), i64]
        L.orc_chol.restype = i64
        L.orc_lsolve.argtypes = [i64, P(i64), P(i64), P(f64), P(f64)]
        L.orc_ltsolve.argtypes = [i64, P(i64), P(i64), P(f64), P(f64)]
        L.orc_multiply_lower_block.argtypes = [ctypes.c_int, P(i64), i64, i64, i64, P(f64), i64,
                                               ctypes.c_int, P(f64), i64]
        L.orc_blas_load.argtypes = [ctypes.c_char_p]
        L.orc_blas_load.restype = ctypes.c_int
        L.orc_sn_analyze.argtypes = [i64, P(i64), P(i64), ctypes.c_int]
        L.orc_sn_analyze.restype = ctypes.c_void_p
        L.orc_sn_info.argtypes = [ctypes.c_void_p, P(f64)]
        L.orc_sn_factor.argtypes = [ctypes.c_void_p, P(i64), P(i64), P(f64), ctypes.c_int]
        L.orc_sn_factor.restype = i64
        L.orc_sn_solve.argtypes = [ctypes.c_void_p, P(f64)]
        L.orc_sn_dense_L.argtypes = [ctypes.c_void_p, P(f64)]
        L.orc_sn_free.argtypes = [ctypes.c_void_p]
        L.orc_sn_phase_times.argtypes = [P(f64)]
        _lib = L
    return _lib


def cube_mesh(N):
    tets = np.empty((6 * N ** 3, 4), np.int32)
    X = np.empty(((N + 1) ** 3, 3))
    lib().orc_cube_mesh(N, _p(tets, i32), _p(X, f64))
    return tets, X


def perturb(N, X, seed, sigma_rel=1e-3):
    Y = np.array(X, dtype=np.float64, copy=True)
    lib().orc_perturb(N, seed, sigma_rel, _p(Y, f64))
    return Y


def energy(Xr, X, tets, E=1e5, nu=0.3, Ee=None, nue=None):
    Xr = np.ascontiguousarray(Xr, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    t = np.ascontiguousarray(tets, np.int32)
    Ee = None if Ee is None else np.ascontiguousarray(Ee, np.float64)
    nue = None if nue is None else np.ascontiguousarray(nue, np.float64)
    return lib().orc_neohookean_energy(_p(Xr, f64), _p(X, f64), _p(t, i32), len(t), _p(Ee, f64), _p(nue, f64),
                                       E, nu)


def pattern(en, nb):
    en = np.ascontiguousarray(en, np.int32)
    ne, npe = en.shape
    Ap = np.zeros(nb + 1, np.int64)
    cap = max(1, ne * npe * (npe + 1) // 2)
    Ai = np.zeros(cap, np.int64)
    nz = lib().orc_pattern(_p(en, i32), ne, npe, nb, _p(Ap, i64), _p(Ai, i64), cap)
    assert nz >= 0
    return Ap, Ai[:nz].copy()


def assemble(Ap, Ai, nb, en, He, Ax=None):
    Ap = np.ascontiguousarray(Ap, np.int64)
    Ai = np.ascontiguousarray(Ai, np.int64)
    en = np.ascontiguousarray(en, np.int32)
    He = np.ascontiguousarray(He, np.float64)
    out = np.zeros(9 * int(Ap[-1])) if Ax is None else np.array(Ax, dtype=np.float64, copy=True)
    lib().orc_assemble(_p(Ap, i64), _p(Ai, i64), nb, _p(en, i32), len(en), _p(He, f64), _p(out, f64))
    return out


def row_col_removal(Ap, Ai, removed):
    n = len(Ap) - 1
    Ap = np.ascontiguousarray(Ap, np.int64)
    Ai = np.ascontiguousarray(Ai, np.int64)
    rem = np.ascontiguousarray(removed, np.uint8)
    Rp = np.zeros(n + 1, np.int64)
    Ri = np.zeros(max(1, len(Ai)), np.int64)
    rr = np.zeros(n, np.int64)
    ee = np.zeros(max(1, len(Ai)), np.int64)
    nz = lib().orc_row_col_removal(_p(Ap, i64), _p(Ai, i64), n, _p(rem, u8), _p(Rp, i64), _p(Ri, i64), _p(rr, i64),
                                   _p(ee, i64))
    nr = int((rem == 0).sum())
    return Rp[:nr + 1].copy(), Ri[:nz].copy(), rr, ee[:nz].copy()


def spmv(Ap, Ai, nb, Ax, x):
    y = np.zeros(3 * nb)
    Ap = np.ascontiguousarray(Ap, np.int64)
    Ai = np.ascontiguousarray(Ai, np.int64)
    Ax = np.ascontiguousarray(Ax, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_spmv(_p(Ap, i64), _p(Ai, i64), nb, _p(Ax, f64), _p(x, f64), _p(y, f64))
    return y


def permuted_lower(Ap, Ai, nb, Ax, red_of_full, pos, sigma=0.0, Bx=None):
    Ap = np.ascontiguousarray(Ap, np.int64)
    Ai = np.ascontiguousarray(Ai, np.int64)
    Ax = np.ascontiguousarray(Ax, np.float64)
    rf = np.ascontiguousarray(red_of_full, np.int64)
    pos = np.ascontiguousarray(pos, np.int32)
    Bx = None if Bx is None else np.ascontiguousarray(Bx, np.float64)
    nred = len(pos)
    cap = 9 * len(Ai)
    Cp = np.zeros(3 * nred + 1, np.int64)
    Ci = np.zeros(cap, np.int64)
    Cx = np.zeros(cap)
    nz = lib().orc_permuted_lower(_p(Ap, i64), _p(Ai, i64), nb, _p(Ax, f64), _p(rf, i64), _p(pos, i32), nred,
                                  sigma, _p(Bx, f64), _p(Cp, i64), _p(Ci, i64), _p(Cx, f64), cap)
    assert nz >= 0
    return Cp, Ci[:nz].copy(), Cx[:nz].copy()


class NotSPD(Exception):
    pass


def chol(Cp, Ci, Cx):
    n = len(Cp) - 1
    Cp = np.ascontiguousarray(Cp, np.int64)
    Ci = np.ascontiguousarray(Ci, np.int64)
    Cx = np.ascontiguousarray(Cx, np.float64)
    Lp = np.zeros(n + 1, np.int64)
    cap = 1
    while True:
        Li = np.zeros(cap, np.int64)
        Lx = np.zeros(cap)
        r = lib().orc_chol(n, _p(Cp, i64), _p(Ci, i64), _p(Cx, f64), _p(Lp, i64), _p(Li, i64), _p(Lx, f64), cap)
        if r < -(1 << 62):
            raise NotSPD(r - np.iinfo(np.int64).min - 1)
        if r < 0:
            cap = -r
            continue
        return Lp, Li, Lx


def lsolve(Lp, Li, Lx, b):
    x = np.array(b, dtype=np.float64, copy=True)
    lib().orc_lsolve(len(Lp) - 1, _p(Lp, i64), _p(Li, i64), _p(Lx, f64), _p(x, f64))
    return x


def ltsolve(Lp, Li, Lx, b):
    x = np.array(b, dtype=np.float64, copy=True)
    lib().orc_ltsolve(len(Lp) - 1, _p(Lp, i64), _p(Li, i64), _p(Lx, f64), _p(x, f64))
    return x


def multiply_lower_block(BS, indices, start, size, degree, A, lda, nrhs, B, ldb):
    idx = np.ascontiguousarray(indices, np.int64)
    A = np.ascontiguousarray(A, np.float64)
    out = np.array(B, dtype=np.float64, copy=True)
    lib().orc_multiply_lower_block(BS, _p(idx, i64), start, size, degree, _p(A, f64), lda, nrhs, _p(out, f64), ldb)
    return out


def dense_from_lower(Cp, Ci, Cx):
    n = len(Cp) - 1
    A = np.zeros((n, n))
    for j in range(n):
        for p in range(Cp[j], Cp[j + 1]):
            A[Ci[p], j] = Cx[p]
            A[j, Ci[p]] = Cx[p]
    return A


def dense_L(Lp, Li, Lx):
    n = len(Lp) - 1
    L = np.zeros((n, n))
    for j in range(n):
        for p in range(Lp[j], Lp[j + 1]):
            L[Li[p], j] = Lx[p]
    return L


def blas_path():
    """The OpenBLAS that ships with scipy (LP64, symbols prefixed scipy_)."""
    import glob
    import scipy
    libs = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs", "libscipy_openblas*.so"))
    if not libs:
        raise ImportError("scipy's bundled OpenBLAS not found")
    return os.path.realpath(libs[0])


_blas_loaded = False


class Supernodal:
    """Multifrontal supernodal Cholesky on the host (oracle/supernodal.c) of a scalar lower CSC."""

    def __init__(self, Cp, Ci, relax=4):
        global _blas_loaded
        L = lib()
        if not _blas_loaded:
            # idle BLAS threads spin for 2^4 cycles only, so they do not compete with the OpenMP
            # threads of the tree levels (read when the library initialises)
            os.environ.setdefault("OPENBLAS_THREAD_TIMEOUT", "4")
            if L.orc_blas_load(blas_path().encode()) != 0:
                raise ImportError("cannot load the BLAS for the supernodal CPU factorization")
            _blas_loaded = True
        self.Cp = np.ascontiguousarray(Cp, np.int64)
        self.Ci = np.ascontiguousarray(Ci, np.int64)
        self.n = len(self.Cp) - 1
        self._h = L.orc_sn_analyze(self.n, _p(self.Cp, i64), _p(self.Ci, i64), relax)
        if not self._h:
            raise ValueError("scalar dimension must be a multiple of 3")
        info = np.zeros(6)
        L.orc_sn_info(self._h, _p(info, f64))
        self.info = dict(supernodes=int(info[0]), levels=int(info[1]), panel_scalars=int(info[2]),
                         flops=float(info[3]), max_width=int(info[4]), arena_scalars=int(info[5]))

    def factor(self, Cx, threads=None):
        Cx = np.ascontiguousarray(Cx, np.float64)
        t = threads or os.cpu_count() or 1
        r = lib().orc_sn_factor(self._h, _p(self.Cp, i64), _p(self.Ci, i64), _p(Cx, f64), int(t))
        if r == -2:
            raise ImportError("BLAS not loaded")
        if r >= 0:
            raise NotSPD(r)

    @staticmethod
    def phase_times():
        """Seconds of the last factorization: per phase summed over threads (assembly, potrf,
        trsm, syrk, rest), then wall time in tree-parallel levels and in BLAS-parallel levels."""
        t = np.zeros(7)
        lib().orc_sn_phase_times(_p(t, f64))
        return dict(zip(["assembly", "potrf", "trsm", "syrk", "rest", "wall_tree", "wall_blas"], t.tolist()))

    def solve(self, b):
        x = np.array(b, dtype=np.float64, copy=True)
        lib().orc_sn_solve(self._h, _p(x, f64))
        return x

    def dense_L(self):
        Ld = np.zeros((self.n, self.n))
        lib().orc_sn_dense_L(self._h, _p(Ld, f64))
        return Ld.T.copy()       # written column-major

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_sn_free(self._h)
            self._h = None
