"""The host supernodal factorization (oracle/supernodal.c), which bench.py times as the CPU
baseline, against the oracle's up-looking Cholesky of the same ordering (the factor of a fixed
ordering is unique), against dense solves, and on a matrix that is not positive definite."""
import numpy as np
import pytest

from tests import oracle_lib as orc
from meshfemsparse_amd import workload as wl


def problem(N, order, seed=0):
    """Reduced (z = 0 pinned) Hessian of the N^3 cube as the scalar lower CSC of P A P^T."""
    tets, Xr = orc.cube_mesh(N)
    X = orc.perturb(N, Xr, seed)
    He, _, _ = wl.neohookean(Xr, X, tets)
    nb = len(Xr)
    Ap, Ai = orc.pattern(tets, nb)
    Ax = orc.assemble(Ap, Ai, nb, tets, He)
    npin = (N + 1) ** 2
    nred = nb - npin
    red = -np.ones(nb, np.int64)
    red[npin:] = np.arange(nred)
    rng = np.random.default_rng(seed)
    if order == "identity":
        pos = np.arange(nred, dtype=np.int32)
    elif order == "random":
        pos = rng.permutation(nred).astype(np.int32)
    else:   # reverse: long chains, wide fronts
        pos = np.arange(nred, dtype=np.int32)[::-1].copy()
    return orc.permuted_lower(Ap, Ai, nb, Ax, red, pos)


@pytest.mark.parametrize("N,order,relax", [(3, "identity", 1), (4, "random", 4), (4, "reverse", 8),
                                           (5, "identity", 4)])
def test_factor_matches_up_looking_cholesky(N, order, relax):
    Cp, Ci, Cx = problem(N, order)
    S = orc.Supernodal(Cp, Ci, relax=relax)
    S.factor(Cx, threads=4)
    Ld = S.dense_L()
    Lo = orc.dense_L(*orc.chol(Cp, Ci, Cx))
    assert np.abs(Ld - Lo).max() <= 1e-12 * np.abs(Lo).max()
    # flops = sum W^3/3 + W^2 D + W D^2: the leading terms of sum_j c_j^2 over the factor's column
    # counts (equal up to O(W^2 + W D) per supernode with fundamental supernodes; relaxed ones add
    # explicit zeros)
    assert S.info["supernodes"] >= 1 and S.info["max_width"] % 3 == 0
    cc = np.diff(orc.chol(Cp, Ci, Cx)[0]).astype(np.float64)
    ratio = S.info["flops"] / (cc ** 2).sum()
    assert 0.9 <= ratio and (relax > 1 or ratio <= 1.0)


@pytest.mark.parametrize("threads", [1, 3])
def test_solve_matches_dense(threads):
    Cp, Ci, Cx = problem(5, "random", seed=1)
    A = orc.dense_from_lower(Cp, Ci, Cx)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0])
    S = orc.Supernodal(Cp, Ci, relax=4)
    S.factor(Cx, threads=threads)
    x = S.solve(b)
    xe = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xe) <= 1e-11 * np.linalg.norm(xe)
    assert np.linalg.norm(A @ x - b) <= 1e-13 * np.linalg.norm(A, 1) * np.linalg.norm(x)


def test_not_positive_definite_reports_first_column():
    Cp, Ci, Cx = problem(3, "identity")
    S = orc.Supernodal(Cp, Ci)
    with pytest.raises(orc.NotSPD) as e:
        S.factor(-Cx, threads=2)
    assert e.value.args[0] == 0
    S.factor(Cx, threads=2)       # reusable after a failure
