"""SparsityLRU (SparsityLRU.hh:26-331): a cache of the block sparsity pattern for problems whose
pattern changes now and then (contact).  The pattern is a static part plus dynamic entries; a
dynamic entry that leaves the current pattern is kept (so a symbolic factorization of the cached
pattern stays valid if it comes back) and ages by one per update until it expires
(`expirationAge`), a rebuild is forced (`hardExpirationAge`), or the number of stale entries
exceeds `entryCacheBudgetRatio` x the current entries (the oldest go first).

`update(S)` returns 0 when the cached pattern can be reused as is (no new symbolic analysis),
the number of entries it had to add, or EXPIRED after a pruning / budget rebuild.  The pattern is
held as sorted int64 keys col * n + row (columns ascending, rows ascending inside a column, the
diagonal last), every merge is vectorized over the keys.  One choice the reference leaves open:
stale entries of equal age are evicted in pattern order (a stable sort), where the reference's
std::sort leaves the order of ties unspecified.
"""
import numpy as np

STATIC_ENTRY_AGE = -1
EXPIRED = np.iinfo(np.int32).max


class SparsityLRU:
    def __init__(self, S_static):
        self.n = int(S_static.n)
        self.Ap = np.asarray(S_static.Ap, np.int64).copy()
        self.Ai = np.asarray(S_static.Ai, np.int64).copy()
        self._ages = np.full(len(self.Ai), STATIC_ENTRY_AGE, np.int64)
        self._num_static = len(self.Ai)
        self._max_age = STATIC_ENTRY_AGE
        self.entryCacheBudgetRatio = 0.01
        self.expirationAge = 25
        self.hardExpirationAge = 250
        self.verbose = False

    # -- pattern access ------------------------------------------------------------------
    @property
    def nz(self):
        return len(self.Ai)

    def entryAges(self):
        return self._ages.copy()

    def _keys(self, Ap, Ai):
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(Ap))
        return cols * self.n + np.asarray(Ai, np.int64)

    def _set_from_keys(self, keys):
        cols = keys // self.n
        self.Ai = keys - cols * self.n
        self.Ap = np.zeros(self.n + 1, np.int64)
        np.add.at(self.Ap, cols + 1, 1)
        np.cumsum(self.Ap, out=self.Ap)

    # -- reference operations ------------------------------------------------------------
    def pruneEntries(self, remove):
        """Drop the entries where the boolean mask (or predicate on entry indices) is true."""
        mask = remove(np.arange(self.nz)) if callable(remove) else np.asarray(remove, bool)
        keep = ~mask
        keys = self._keys(self.Ap, self.Ai)[keep]
        self._ages = self._ages[keep]
        self._set_from_keys(keys)
        self._max_age = int(self._ages.max()) if len(self._ages) else int(np.iinfo(np.int32).min)

    def resetAge(self, ii):
        if self._ages[ii] != STATIC_ENTRY_AGE:
            self._ages[ii] = 0

    def increaseAgeOfOldEntries(self, threshold=0):
        """Age every entry older than `threshold`; prune expired entries (and return True) once
        an entry reaches hardExpirationAge."""
        if self._max_age > threshold:
            old = self._ages > threshold
            self._ages[old] += 1
            aged = self._ages[old]
            max_age = int(aged.max()) if len(aged) else STATIC_ENTRY_AGE
            below = aged[aged < self.expirationAge]
            max_below = max(STATIC_ENTRY_AGE, int(below.max())) if len(below) else STATIC_ENTRY_AGE
            max_age = max(STATIC_ENTRY_AGE, max_age)
            if max_age >= self.hardExpirationAge:
                self.pruneEntries(self._ages >= self.expirationAge)
                self._max_age = max_below
                return True
            self._max_age = max_age
        return False

    def update(self, S_dynamic):
        if S_dynamic.n != self.n:
            raise RuntimeError("SparsityLRU: S_dynamic size mismatch")
        if len(self._ages) != self.nz:
            raise RuntimeError("SparsityLRU: age size mismatch")
        ratio = self.entryCacheBudgetRatio
        Dnz = int(np.asarray(S_dynamic.Ap)[-1])
        if Dnz == 0 and len(self._ages) <= ratio * self._num_static:
            return EXPIRED if self.increaseAgeOfOldEntries(STATIC_ENTRY_AGE) else 0
        kS = self._keys(self.Ap, self.Ai)
        kD = self._keys(np.asarray(S_dynamic.Ap, np.int64), np.asarray(S_dynamic.Ai, np.int64))
        both_S = np.isin(kS, kD, assume_unique=True)
        in_S_D = np.isin(kD, kS, assume_unique=True)
        ages = self._ages
        static = ages == STATIC_ENTRY_AGE
        # entries in both: reset (static ones stay static)
        ages[both_S & ~static] = 0
        # entries only in the cache: dynamic ones age and become stale
        s_only_dyn = ~both_S & ~static
        ages[s_only_dyn] += 1
        aged = ages[s_only_dyn]
        max_age = max(STATIC_ENTRY_AGE, int(aged.max())) if len(aged) else STATIC_ENTRY_AGE
        below = aged[aged < self.expirationAge]
        max_below = max(STATIC_ENTRY_AGE, int(below.max())) if len(below) else STATIC_ENTRY_AGE
        old_loc = np.nonzero(s_only_dyn)[0]                     # in pattern order
        new_keys = kD[~in_S_D]
        total_new = len(new_keys)
        total_current = int(both_S.sum() + (~both_S & static).sum() + total_new)
        if self.verbose:
            print(f"totalNewEntries: {total_new}")
        if total_new == 0:
            if max_age >= self.hardExpirationAge:
                self.pruneEntries(ages >= self.expirationAge)
                self._max_age = max_below
                return EXPIRED
            self._max_age = max_age
            if len(old_loc) <= total_current * ratio:
                return 0
        # rebuild: keep at most `budget` stale entries (the youngest), drop expired ones
        budget = int(total_current * ratio)
        if len(old_loc) > budget:
            order = np.argsort(ages[old_loc], kind="stable")
            old_loc = old_loc[order]
            ages[old_loc[budget:]] = EXPIRED
            total_new = EXPIRED
        keep = both_S | (~both_S & (ages < self.expirationAge))
        kept_keys = kS[keep]
        kept_ages = np.where(both_S[keep], np.minimum(0, ages[keep]), ages[keep])
        keys = np.concatenate([kept_keys, new_keys])
        all_ages = np.concatenate([kept_ages, np.zeros(len(new_keys), np.int64)])
        order = np.argsort(keys, kind="stable")
        keys, all_ages = keys[order], all_ages[order]
        self._set_from_keys(keys)
        self._ages = all_ages
        last = self.Ap[1:] - 1
        if (np.diff(self.Ap) == 0).any() or not np.array_equal(self.Ai[last], np.arange(self.n)):
            raise RuntimeError("SparsityLRU: missing diagonal entry in new sparsity pattern")
        self._max_age = int(all_ages.max()) if len(all_ages) else int(np.iinfo(np.int32).min)
        return int(total_new)


def expand_values(H, Ap, Ai):
    """The block values of H scattered into the (larger) cached pattern Ap/Ai; entries the cache
    holds beyond H's pattern are zero ("enlarge numeric sparsity" of the reference's replay,
    tests/benchmark_sparsity_lru.cc)."""
    import torch
    from .sparse import BlockCSCHessian
    if not H.hasContiguousBlocks():
        return expand_values(H.cloneWithLayout(True), Ap, Ai).cloneWithLayout(False)
    n = H.n
    Ap = np.asarray(Ap, np.int64)
    Ai = np.asarray(Ai, np.int64)
    kd = np.repeat(np.arange(n, dtype=np.int64), np.diff(Ap)) * n + Ai
    ks = np.repeat(np.arange(n, dtype=np.int64), np.diff(H.Ap)) * n + np.asarray(H.Ai, np.int64)
    idx = np.searchsorted(kd, ks)
    if (idx >= len(kd)).any() or not np.array_equal(kd[np.minimum(idx, len(kd) - 1)], ks):
        raise RuntimeError("pattern of H is not contained in the cached pattern")
    out = BlockCSCHessian(Ap, Ai, H.Ax.device)
    Ax = torch.zeros(9 * len(Ai), dtype=H.Ax.dtype, device=H.Ax.device)
    Ax.view(-1, 9)[torch.from_numpy(idx).to(H.Ax.device)] = H.Ax.view(-1, 9)
    out.Ax = Ax
    return out
