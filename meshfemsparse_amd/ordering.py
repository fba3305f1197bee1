"""Nested-dissection ordering with temporal reuse of the separator tree, as the factorizer uses it.

Mirrors the reference's stateful ordering entry point (ParallelNestedDissection/
cholmod_nesdis_parallel.hh:41-75: nested_dissection_from_graph with a caller-owned
TemporalReuseState) and its separator-tree metadata (CholeskyFactorizerBase.hh:51-59, NDOrdering).
The work happens in the native library (meshfemsparse_amd/csrc/ordering.cpp).
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import c_vp, ptr


@dataclass
class NDOrdering:
    """Separator tree of a dissection: CParent[c] is the parent component of c (components are
    postordered, so CParent[c] > c) or -1 for a root; CMember[v] is the component of ordering
    variable v; every ordering variable stands for blockSize scalar unknowns."""
    CParent: np.ndarray
    CMember: np.ndarray
    blockSize: int = 1


@dataclass
class TemporalReuseStatistics:
    """cholmod_nesdis_parallel.hh:33-39."""
    full_rebuild: bool = False
    graph_vertices: int = 0
    recomputed_subtrees: int = 0
    repartitioned_vertices: int = 0
    nd_seconds: float = 0.0


def _stats(fn, h):
    a = np.zeros(4, np.int64)
    t = _lib.c_double()
    _lib.call(fn, h, ptr(a, _lib.c_i64), t)
    return TemporalReuseStatistics(bool(a[0]), int(a[1]), int(a[2]), int(a[3]), t.value)


class NestedDissection:
    """One ordering stream with its own reuse history.  `order(p, i)` orders the symmetric graph
    given as CSC of both triangles (diagonal entries ignored) and returns perm[new] = old; with
    `temporalReusePeriod` P > 0, the next P calls on graphs over the same vertices repair only the
    subtrees whose separators the graph violates, then the tree is rebuilt from scratch."""

    def __init__(self, temporalReusePeriod=0):
        self._h = c_vp()
        _lib.call("mfs_nd_create", self._h)
        self._period = 0
        self._n = 0
        self._ncomp = 0
        self.temporalReusePeriod = temporalReusePeriod

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mfs_nd_destroy(self._h)
            self._h = None

    @property
    def temporalReusePeriod(self):
        return self._period

    @temporalReusePeriod.setter
    def temporalReusePeriod(self, period):
        _lib.call("mfs_nd_set_temporal_reuse_period", self._h, int(period))
        self._period = int(period)

    def resetTemporalReuse(self):
        _lib.call("mfs_nd_reset", self._h)

    def order(self, p, i):
        p = np.ascontiguousarray(p, np.int64)
        i = np.ascontiguousarray(i, np.int64)
        n = len(p) - 1
        perm = np.empty(max(n, 0), np.int64)
        nc = _lib.c_i64()
        _lib.call("mfs_nd_order", self._h, n, ptr(p, _lib.c_i64), ptr(i, _lib.c_i64), ptr(perm, _lib.c_i64), nc)
        self._n, self._ncomp = n, nc.value
        return perm

    def ndOrdering(self):
        parent = np.empty(self._ncomp, np.int64)
        member = np.empty(self._n, np.int64)
        _lib.call("mfs_nd_forest", self._h, ptr(parent, _lib.c_i64), ptr(member, _lib.c_i64))
        return NDOrdering(parent, member, 1)

    def temporalReuseStatistics(self):
        return _stats("mfs_nd_statistics", self._h)


__all__ = ["NDOrdering", "NestedDissection", "TemporalReuseStatistics"]
