"""Learned block incomplete-Cholesky preconditioner (device network, FP32 PCG, FP64 refinement).

A message-passing network (hidden 64, 4 rounds, bfloat16-rounded weights, FP32 accumulation) over
the free-node block graph of a Hessian emits a zero-fill lower block factor L on tril(A) in a
multicolor ordering; M = L L^T preconditions FP32 CG on the Jacobi-scaled system, inside FP64
refinement, with the supernodal Cholesky factorizer as fallback.  Training runs on the device
(forward/backward/Adam kernels of the library) with the probe loss
mean_p |(L L^T - A^) z_p|^2 / |A^ z_p|^2.  There is no reference counterpart: the reference
preconditions with exact factorizations only.
"""
import numpy as np
import torch

from . import _lib
from ._lib import c_vp, ptr
from .sparse import _dptr, _stream

PRECONDITIONERS = {"jacobi": 0, "learned": 1}
PARAM_NAMES = (["W0", "b0"] + [f"{k}{l}" for l in range(4) for k in ("Wm", "bm", "Wu", "bu")] +
               ["Wd1", "bd1", "Wd2", "bd2", "Wdd", "bdd"])


def graph_host(H, pins):
    """The preconditioner's graph from the pattern alone (no device): dict of numpy arrays."""
    pins = np.ascontiguousarray(np.unique(np.asarray(pins, np.int64)) if len(pins) else np.zeros(0, np.int64))
    sz = np.zeros(4, np.int64)
    args = (ptr(H.Ap, _lib.c_i64), ptr(H.Ai, _lib.c_i64), H.n, ptr(pins, _lib.c_i64), len(pins))
    _lib.call("mfs_learned_analyze_host", *args, ptr(sz, _lib.c_i64), None, None, None, None, None, None)
    n, E, nl, nc = (int(v) for v in sz)
    out = {k: np.zeros(m, np.int32) for k, m in (("edge_dst", E), ("edge_src", E), ("lower_edge", nl),
                                                 ("lower_dst", nl), ("lower_src", nl), ("color", n))}
    _lib.call("mfs_learned_analyze_host", *args, ptr(sz, _lib.c_i64),
              *(ptr(out[k], _lib.c_i32) for k in ("edge_dst", "edge_src", "lower_edge", "lower_dst", "lower_src",
                                                   "color")))
    out.update(n=n, edges=E, lower=nl, colors=nc)
    return out



def _contiguous_only(H):
    if not H.hasContiguousBlocks():
        raise ValueError("the learned preconditioner reads contiguous 3x3 blocks; use cloneWithLayout(True)")

class LearnedPreconditioner:
    def __init__(self, H, pinnedVars=(), seed=0):
        pins = np.unique(np.asarray(pinnedVars, np.int64)) if len(pinnedVars) else np.zeros(0, np.int64)
        self._pins = np.ascontiguousarray(pins)
        self._h = c_vp()
        _lib.call("mfs_learned_create", ptr(H.Ap, _lib.c_i64), ptr(H.Ai, _lib.c_i64), H.n,
                  ptr(self._pins, _lib.c_i64), len(self._pins), int(seed) & 0xFFFFFFFFFFFFFFFF, self._h)
        v = np.zeros(8, np.int64)
        _lib.call("mfs_learned_info", self._h, ptr(v, _lib.c_i64))
        self.n, self.edges, self.lower, self.colors, self.nparams = (int(x) for x in v[:5])
        off = np.zeros(len(PARAM_NAMES) + 1, np.int64)
        _lib.call("mfs_learned_param_offsets", self._h, ptr(off, _lib.c_i64))
        self.offsets = off

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mfs_learned_destroy(self._h)
            self._h = None

    def info(self):
        v = np.zeros(8, np.int64)
        _lib.call("mfs_learned_info", self._h, ptr(v, _lib.c_i64))
        return dict(zip(["nodes", "edges", "lower_blocks", "colors", "params", "steps", "hidden", "rounds"],
                        (int(x) for x in v)))

    def graph(self):
        out = {k: np.zeros(m, np.int32) for k, m in (("edge_dst", self.edges), ("edge_src", self.edges),
                                                     ("lower_edge", self.lower), ("lower_dst", self.lower),
                                                     ("lower_src", self.lower), ("color", self.n))}
        _lib.call("mfs_learned_graph", self._h, *(ptr(out[k], _lib.c_i32) for k in
                                                  ("edge_dst", "edge_src", "lower_edge", "lower_dst", "lower_src",
                                                   "color")))
        return out

    def setMatrix(self, H):
        _contiguous_only(H)
        _lib.call("mfs_learned_set_matrix", self._h, _dptr(H.Ax), _stream())

    def features(self):
        nf = np.zeros((self.n, 6), np.float32)
        ef = np.zeros((self.edges, 9), np.float32)
        dg = np.zeros((self.n, 9), np.float32)
        sc = np.zeros(3 * self.n)
        _lib.call("mfs_learned_features", self._h, nf.ctypes.data, ef.ctypes.data, dg.ctypes.data, sc.ctypes.data)
        return dict(node=nf, edge=ef, diag=dg, scale=sc)

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        _lib.call("mfs_learned_params", self._h, p.ctypes.data, None)
        return p

    def setParams(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _lib.call("mfs_learned_params", self._h, None, p.ctypes.data)

    def grads(self):
        g = np.zeros(self.nparams, np.float32)
        _lib.call("mfs_learned_grads", self._h, g.ctypes.data)
        return g

    def split(self, flat):
        return {k: flat[self.offsets[i]:self.offsets[i + 1]] for i, k in enumerate(PARAM_NAMES)}

    def loss(self, z, backward=True):
        z = np.ascontiguousarray(z, np.float32)
        nprobe = z.size // (3 * self.n)
        v = _lib.c_double()
        _lib.call("mfs_learned_loss", self._h, z.ctypes.data, nprobe, int(bool(backward)), v, _stream())
        return v.value

    def adamStep(self, lr=1e-3):
        _lib.call("mfs_learned_adam_step", self._h, float(lr), _stream())

    def trainStep(self, H, nprobe=4, lr=1e-3, rng=None):
        """One Adam step on the probe loss of H's current values; returns the loss."""
        rng = rng or np.random.default_rng()
        self.setMatrix(H)
        loss = self.loss(rng.standard_normal((nprobe, 3 * self.n)).astype(np.float32), True)
        self.adamStep(lr)
        return loss

    def factor(self):
        LE = np.zeros((self.lower, 9), np.float32)
        LD = np.zeros((self.n, 9), np.float32)
        _lib.call("mfs_learned_factor", self._h, LE.ctypes.data, LD.ctypes.data, _stream())
        return LE, LD

    def solve(self, H, b, precond="learned", tol=1e-10, maxPCG=200, fallback=None):
        """A x = b to tol (FP32 PCG + FP64 refinement; the fallback factorizer takes over after
        maxPCG CG iterations).  Returns (x, pcg iterations, refinement steps, relres, used fallback)."""
        _contiguous_only(H)
        x = torch.empty_like(b)
        it, ot, fb = _lib.c_i32(), _lib.c_i32(), _lib.c_i32()
        rr = _lib.c_double()
        _lib.call("mfs_learned_solve", self._h, fallback._h if fallback is not None else None, _dptr(H.Ax),
                  _dptr(b.contiguous()), _dptr(x), PRECONDITIONERS[precond], float(tol), int(maxPCG), it, ot, rr, fb,
                  _stream())
        return x, it.value, ot.value, rr.value, bool(fb.value)
