"""Oracle for SparsityLRU (test infrastructure only): a per-entry, per-column restatement of
SparsityLRU.hh:40-295 in plain Python loops (pruneEntries :40-60, resetAge :62-65,
increaseAgeOfOldEntries :79-104, update :113-295), for small patterns.  Patterns are lists of
sorted row lists per column.  Ties among stale entries of equal age are evicted in pattern order
(the reference's std::sort leaves that order open)."""
STATIC = -1
EXPIRED = 2 ** 31 - 1


class LRUOracle:
    def __init__(self, cols, ratio=0.01, expiration=25, hard=250):
        self.cols = [list(c) for c in cols]
        self.ages = [[STATIC] * len(c) for c in cols]
        self.num_static = sum(len(c) for c in cols)
        self.max_age = STATIC
        self.ratio, self.expiration, self.hard = ratio, expiration, hard

    def _prune(self, remove):
        self.max_age = -(2 ** 31)
        for j in range(len(self.cols)):
            keep = [k for k in range(len(self.cols[j])) if not remove(self.ages[j][k])]
            self.cols[j] = [self.cols[j][k] for k in keep]
            self.ages[j] = [self.ages[j][k] for k in keep]
            for a in self.ages[j]:
                self.max_age = max(self.max_age, a)

    def increase_age(self, threshold=0):
        if self.max_age > threshold:
            mx, mxb = STATIC, STATIC
            for j in range(len(self.cols)):
                for k in range(len(self.ages[j])):
                    if self.ages[j][k] > threshold:
                        self.ages[j][k] += 1
                        mx = max(mx, self.ages[j][k])
                        if self.ages[j][k] < self.expiration:
                            mxb = max(mxb, self.ages[j][k])
            if mx >= self.hard:
                self._prune(lambda a: a >= self.expiration)
                self.max_age = mxb
                return True
            self.max_age = mx
        return False

    def update(self, dyn):
        total_entries = sum(len(a) for a in self.ages)
        dnz = sum(len(c) for c in dyn)
        if dnz == 0 and total_entries <= self.ratio * self.num_static:
            return EXPIRED if self.increase_age(STATIC) else 0
        mx, mxb = STATIC, STATIC
        current, new = 0, 0
        old = []                          # (column, position) of stale dynamic entries
        new_cols = [0] * len(self.cols)
        for j in range(len(self.cols)):
            S, D, A = self.cols[j], dyn[j], self.ages[j]
            s = d = 0
            while s < len(S) or d < len(D):
                if s < len(S) and d < len(D) and S[s] == D[d]:
                    if A[s] != STATIC:
                        A[s] = 0
                    current += 1
                    s += 1
                    d += 1
                elif d >= len(D) or (s < len(S) and S[s] < D[d]):
                    if A[s] != STATIC:
                        A[s] += 1
                        mx = max(mx, A[s])
                        if A[s] < self.expiration:
                            mxb = max(mxb, A[s])
                        old.append((j, s))
                    else:
                        current += 1
                    s += 1
                else:
                    new_cols[j] += 1
                    new += 1
                    current += 1
                    d += 1
        if new == 0:
            if mx >= self.hard:
                self._prune(lambda a: a >= self.expiration)
                self.max_age = mxb
                return EXPIRED
            self.max_age = mx
            if len(old) <= current * self.ratio:
                return 0
        budget = int(current * self.ratio)
        if len(old) > budget:
            old = sorted(old, key=lambda loc: self.ages[loc[0]][loc[1]])     # stable
            new = EXPIRED
        for i, (j, s) in enumerate(old):
            if i >= budget:
                self.ages[j][s] = EXPIRED
        cols, ages = [], []
        self.max_age = -(2 ** 31)
        for j in range(len(self.cols)):
            S, D, A = self.cols[j], dyn[j], self.ages[j]
            c, a = [], []
            s = d = 0
            while s < len(S) or d < len(D):
                if s < len(S) and d < len(D) and S[s] == D[d]:
                    c.append(S[s])
                    a.append(min(0, A[s]))
                    s += 1
                    d += 1
                elif d >= len(D) or (s < len(S) and S[s] < D[d]):
                    if A[s] < self.expiration:
                        c.append(S[s])
                        a.append(A[s])
                    s += 1
                else:
                    c.append(D[d])
                    a.append(0)
                    d += 1
            if not c or c[-1] != j:
                raise RuntimeError("missing diagonal entry")
            cols.append(c)
            ages.append(a)
            for x in a:
                self.max_age = max(self.max_age, x)
        self.cols, self.ages = cols, ages
        return new
