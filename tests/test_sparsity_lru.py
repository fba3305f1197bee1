"""SparsityLRU (meshfemsparse_amd/sparsity_lru.py, vectorized over pattern keys) against the
per-entry loop restatement of SparsityLRU.hh (tests/lru_oracle.py) on random sequences of
changing patterns, plus hand-checked cases of each rule: reuse (0), additions (count), budget
eviction and hard expiration (EXPIRED), expiry of stale entries, static entries never aging."""
import numpy as np
import pytest

from meshfemsparse_amd.sparsity_lru import EXPIRED, STATIC_ENTRY_AGE, SparsityLRU, expand_values
from tests.lru_oracle import LRUOracle


class Pat:
    def __init__(self, cols):
        self.n = len(cols)
        self.Ap = np.zeros(self.n + 1, np.int64)
        self.Ap[1:] = np.cumsum([len(c) for c in cols])
        self.Ai = np.array([r for c in cols for r in c], np.int64)


def static_cols(n, rng):
    cols = []
    for j in range(n):
        rows = sorted(set(rng.choice(j, size=min(j, 2), replace=False).tolist()) if j else set())
        cols.append(rows + [j])
    return cols


def with_dynamic(static, pool, on):
    cols = [set(c) for c in static]
    for (i, j), o in zip(pool, on):
        if o:
            cols[j].add(i)
    return [sorted(c) for c in cols]


def cols_of(lru):
    return [lru.Ai[lru.Ap[j]:lru.Ap[j + 1]].tolist() for j in range(lru.n)]


@pytest.mark.parametrize("seed,ratio,exp,hard", [(0, 0.01, 25, 250), (1, 0.05, 4, 9), (2, 0.2, 3, 6),
                                                 (3, 0.0, 2, 5)])
def test_matches_loop_oracle_on_random_sequences(seed, ratio, exp, hard):
    rng = np.random.default_rng(seed)
    n = 60
    static = static_cols(n, rng)
    pool = []
    while len(pool) < 40:
        j = int(rng.integers(1, n))
        i = int(rng.integers(0, j))
        if i not in static[j] and (i, j) not in pool:
            pool.append((i, j))
    lru = SparsityLRU(Pat(static))
    lru.entryCacheBudgetRatio, lru.expirationAge, lru.hardExpirationAge = ratio, exp, hard
    orc = LRUOracle(static, ratio, exp, hard)
    on = rng.random(len(pool)) < 0.3
    codes = set()
    for step in range(300):
        flip = rng.random(len(pool)) < (0.02 if step % 50 < 40 else 0.3)
        on = on ^ flip
        dyn = with_dynamic(static, pool, on)
        r = lru.update(Pat(dyn))
        ro = orc.update(dyn)
        assert r == ro, step
        codes.add("reuse" if r == 0 else "expired" if r == EXPIRED else "added")
        assert cols_of(lru) == orc.cols, step
        ages = lru.entryAges()
        assert ages.tolist() == [a for col in orc.ages for a in col], step
        assert (ages == STATIC_ENTRY_AGE).sum() == sum(len(c) for c in static)
        if step % 17 == 0:     # an update request without changes
            assert lru.increaseAgeOfOldEntries() == orc.increase_age()
            assert cols_of(lru) == orc.cols
    assert len(codes) >= 2


def test_rules_by_hand():
    static = [[0], [0, 1], [1, 2], [3]]
    lru = SparsityLRU(Pat(static))
    lru.entryCacheBudgetRatio, lru.expirationAge, lru.hardExpirationAge = 0.5, 3, 5
    assert lru.update(Pat(static)) == 0                       # nothing changed: reuse
    dyn = [[0], [0, 1], [0, 1, 2], [2, 3]]                     # two new entries
    assert lru.update(Pat(dyn)) == 2
    assert cols_of(lru) == dyn
    assert lru.entryAges().tolist() == [-1, -1, -1, 0, -1, -1, 0, -1]
    # the dynamic entries leave: they stay cached (the budget 0.5 x 6 = 3 holds them) and age
    assert lru.update(Pat(static)) == 0
    assert cols_of(lru) == dyn
    assert lru.entryAges().tolist() == [-1, -1, -1, 1, -1, -1, 1, -1]
    assert lru.update(Pat(static)) == 0 and lru.update(Pat(static)) == 0
    assert lru.entryAges().tolist() == [-1, -1, -1, 3, -1, -1, 3, -1]
    # age 5 = hard expiration: the stale entries (>= 3) are pruned, a rebuild is requested
    assert lru.update(Pat(static)) == 0
    assert lru.update(Pat(static)) == EXPIRED
    assert cols_of(lru) == static
    # a budget too small for the stale entries: the oldest go, the result is EXPIRED
    lru2 = SparsityLRU(Pat(static))
    lru2.entryCacheBudgetRatio = 0.0
    assert lru2.update(Pat(dyn)) == 2                         # no stale entries yet: additions
    assert lru2.update(Pat(static)) == EXPIRED                # 2 stale entries > budget 0
    assert cols_of(lru2) == static


def test_missing_diagonal_and_size_checks():
    lru = SparsityLRU(Pat([[0], [0, 1]]))
    with pytest.raises(RuntimeError):
        lru.update(Pat([[0]]))
    assert lru.update(Pat([[0], [0]])) == 0       # the cached static diagonal stays
    bad = SparsityLRU(Pat([[0], [0], [1, 2]]))   # a cache whose column 1 has no diagonal
    with pytest.raises(RuntimeError):
        bad.update(Pat([[0], [0], [0, 1, 2]]))   # a rebuild (new entry (0, 2)) finds it


def test_expand_values_into_cached_pattern():
    import torch
    from meshfemsparse_amd import BlockCSCHessian
    H = BlockCSCHessian(np.array([0, 1, 3]), np.array([0, 0, 1]), "cpu")
    H.Ax = torch.arange(27, dtype=torch.float64)
    with pytest.raises(RuntimeError):                          # (0, 1) is missing from the target
        expand_values(H, np.array([0, 1, 2]), np.array([0, 1]))
    Ap, Ai = np.array([0, 1, 3]), np.array([0, 0, 1])
    E = expand_values(H, Ap, Ai)
    assert torch.equal(E.Ax, H.Ax)
    Ap2, Ai2 = np.array([0, 1, 3, 6]), np.array([0, 0, 1, 0, 1, 2])
    H3 = BlockCSCHessian(np.array([0, 1, 3, 4]), np.array([0, 0, 1, 2]), "cpu")
    H3.Ax = torch.arange(36, dtype=torch.float64) + 1
    E3 = expand_values(H3, Ap2, Ai2)
    v = E3.Ax.view(-1, 9)
    assert torch.equal(v[[0, 1, 2, 5]], H3.Ax.view(-1, 9))
    assert (v[[3, 4]] == 0).all()
