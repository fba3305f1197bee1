"""GPU: the reduced-precision factorizations against the FP64 oracle.

Forms (no reference counterpart beyond the float state of CatamariFactorizer.cc:52-77,648-656):
  * FP32: symmetric diagonal scaling S A S, FP32 factor on MFMA, FP64 iterative refinement;
  * BF16: as FP32, the Schur-complement updates (in-panel and update-matrix GEMMs) on bfloat16
    MFMA with FP32 accumulation;
  * BF16X3: as BF16 with every operand split hi + lo = bf16(v) + bf16(v - bf16(v)) and the three
    products hi*hi + hi*lo + lo*hi accumulated in FP32 (about 16 significant bits per operand).

Checks:
  * the exported (unscaled) factor against the oracle's up-looking Cholesky factor of the same
    ordering (unique), column-relative; the bound of each form follows its operand rounding
    (bf16: unit roundoff 2^-9, split: about 2^-17, FP32: 2^-24), and each form must show an error
    of its own size (it really ran in that precision);
  * FP64 refinement of every form reaches 1e-10 on the 16^3 stress configuration (per-element E
    log-uniform in [1e3, 1e7], nu uniform in [0.2, 0.45]), and the refined solution agrees with
    the oracle's FP64 factor solve;
  * loss of definiteness in the bfloat16 forms is reported like in FP32
    (test_catamari_factorizer.cc:197-212);
  * the diagonal blocks of every single-precision form are factored in FP64 (a dense one-tile
    matrix gives the FP64 factor of its FP32 entries, rounded once).
The 64^3 stress configuration runs in tests/test_gpu_large.py.
"""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from tests import oracle_lib as orc
from tests.test_gpu_factor import mesh_problem, oracle_factor, one_block

pytestmark = pytest.mark.gpu

# column-relative factor error bounds: FP32 as the reference's single-precision tolerance
# (test_catamari_factorizer.cc:31); bf16 operands carry 8 significant bits (u = 2^-9 ~ 2e-3),
# the split operands about 16 (~8e-6), before accumulation over the updates of a column
FACTOR_BOUND = {None: 2e-5, "bf16": 4e-2, "bf16x3": 2e-4}
FACTOR_FLOOR = {None: 0.0, "bf16": 1e-4, "bf16x3": 0.0}


def factorizer(form):
    if form == "fp64":
        return mfs.make_cholesky_factorizer()
    return mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, True, form)


def colrel_error(f, H):
    (Lp, Li, Lx), _, _ = oracle_factor(H, f)
    Lo = orc.dense_L(Lp, Li, Lx)
    Lg = orc.dense_L(*f.exportFactorCSC())
    cn = np.abs(Lo).max(axis=0)
    return (np.abs(Lg - Lo).max(axis=0) / cn).max(), Lg


@pytest.mark.parametrize("N", [8, 12])
def test_bf16_factors_match_unique_cholesky(dev, N):
    """Exported factors of the three single-precision forms against the oracle factor."""
    H, pins, _, _, _ = mesh_problem(dev, N, seed=4)
    errs = {}
    for form in (None, "bf16", "bf16x3"):
        f = factorizer(form)
        f.factorizeSymbolic(H, pins)
        f.factorizeNumeric(H)
        errs[form], Lg = colrel_error(f, H)
        print(f"N={N} schur={form}: max column-relative factor error {errs[form]:.3e}")
        assert errs[form] <= FACTOR_BOUND[form], (form, errs[form])
        assert errs[form] >= FACTOR_FLOOR[form], (form, errs[form])
        # bitwise reproducible run to run
        f.factorizeNumeric(H)
        assert np.array_equal(orc.dense_L(*f.exportFactorCSC()), Lg)
    # the bfloat16 forms really round their Schur operands: plain bf16 is far less accurate than
    # FP32, and the split recovers most of the lost bits
    assert errs["bf16"] > 20 * errs[None]
    assert errs["bf16x3"] < errs["bf16"] / 20


@pytest.mark.parametrize("form", ["fp32", "bf16", "bf16x3"])
def test_stress_16cube_refines_to_oracle_solution(dev, form):
    """16^3 stress configuration: the refined solution of every reduced-precision form meets
    1e-10 and matches the oracle's FP64 factor solve; the refinement improves on the plain
    reduced-precision solve."""
    N = 16
    rng = np.random.default_rng(7)
    ne = 6 * N ** 3
    Ee, nue = 10 ** rng.uniform(3, 7, ne), rng.uniform(0.2, 0.45, ne)
    H, pins, g, _, _ = mesh_problem(dev, N, seed=1, Ee=Ee, nue=nue)
    f = factorizer(None if form == "fp32" else form)
    f.factorizeSymbolic(H, pins)
    f.factorizeNumeric(H)
    b = torch.from_numpy(-g).to(dev)
    x0 = f.solve(b).cpu().numpy()
    x, it, rr = f.solveRefined(H, b, tol=1e-10, maxIters=40)
    assert rr <= 1e-10
    # the oracle's solve in the device ordering (the factor of a fixed ordering is unique)
    (Lp, Li, Lx), _, _ = oracle_factor(H, f)
    ip = f.getInversePermutation()
    free = np.setdiff1d(np.arange(3 * H.n), pins)
    xp = orc.ltsolve(Lp, Li, Lx, orc.lsolve(Lp, Li, Lx, -g[free][ip]))
    xo = np.zeros(3 * H.n)
    xo[free[ip]] = xp
    xg = x.cpu().numpy()
    # independent residual of the refined solution with the oracle's SpMV
    r = orc.spmv(H.Ap, H.Ai, H.n, H.Ax.cpu().numpy(), xg) + g
    assert np.linalg.norm(r[free]) <= 1e-10 * np.linalg.norm(g[free])
    e1 = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    e0 = np.linalg.norm(x0 - xo) / np.linalg.norm(xo)
    print(f"{form}: {it} refinement steps, relres {rr:.2e}, forward error {e1:.2e} (unrefined {e0:.2e})")
    assert e1 <= 1e-6 and e1 < e0
    assert 1 <= it <= {"fp32": 5, "bf16x3": 10, "bf16": 40}[form]


@pytest.mark.parametrize("form", ["bf16", "bf16x3"])
def test_bf16_rounding_loss_of_definiteness(dev, form):
    """The reference's float SPD-loss matrix (test_catamari_factorizer.cc:197-212) fails in the
    bfloat16 forms too, and a valid matrix refactors afterwards."""
    H = one_block(dev, [[1, 1, 0], [1, 1 + 1e-8, 0], [0, 0, 1]])
    f = factorizer(form)
    with pytest.raises(mfs.NotPositiveDefiniteError):
        f.factorize(H)
    assert not f.hasNumericFactorization()
    H2 = one_block(dev, [[4, 1, 0], [1, 3, 0], [0, 0, 2]])
    f.factorize(H2)
    x = f.solve(torch.tensor([5.0, 4.0, 2.0], dtype=torch.float64, device=dev))
    assert torch.allclose(x, torch.ones(3, dtype=torch.float64, device=dev), rtol=0, atol=1e-6)


@pytest.mark.parametrize("form", ["bf16", "bf16x3"])
def test_bf16_negated_matrix_fails_and_recovers(dev, form):
    """-A fails in the bfloat16 forms (non-positive diagonal under the scaling), the factor drops
    back to the symbolic state, and A factors again."""
    H, pins, g, _, _ = mesh_problem(dev, 6, seed=2)
    f = factorizer(form)
    f.factorizeSymbolic(H, pins)
    H.Ax.neg_()
    with pytest.raises(mfs.NotPositiveDefiniteError):
        f.factorizeNumeric(H)
    assert not f.hasNumericFactorization()
    H.Ax.neg_()
    f.factorizeNumeric(H)
    b = torch.from_numpy(-g).to(dev)
    _, _, rr = f.solveRefined(H, b, tol=1e-10, maxIters=40)
    assert rr <= 1e-10


def chol_fp32(A):
    """Column Cholesky with every operation in float32 (numpy.linalg computes float32 input in
    double), to show what FP32 pivots would give."""
    A = A.astype(np.float32)
    n = len(A)
    L = np.zeros_like(A)
    for j in range(n):
        d = np.float32(A[j, j] - np.dot(L[j, :j], L[j, :j]))
        L[j, j] = np.sqrt(d)
        L[j + 1:, j] = (A[j + 1:, j] - L[j + 1:, :j] @ L[j, :j]) / L[j, j]
    return L.astype(np.float64)


@pytest.mark.parametrize("form", [None, "bf16", "bf16x3"])
def test_diagonal_block_pivots_in_fp64(dev, form):
    """In the single-precision forms the diagonal blocks are factored in FP64 from their FP32
    entries: a dense 60 x 60 SPD matrix with unit diagonal (one supernode, one diagonal tile; the
    diagonal scaling is exactly 1) and condition number ~1e5 gives the FP64 Cholesky factor of its
    FP32-rounded entries, rounded once to FP32 (column-relative error <= 1e-7).  Pivots in FP32
    would miss that bound by three orders of magnitude (shown on the host)."""
    rng = np.random.default_rng(7)
    n = 60
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    M = (Q * np.logspace(0, -5, n)) @ Q.T
    d = np.sqrt(np.diag(M))
    A = M / d[:, None] / d[None, :]
    np.fill_diagonal(A, 1.0)
    iu, ju = np.nonzero(np.triu(np.ones((n, n))))
    S = mfs.SuiteSparseMatrix.fromTriplets(n, iu, ju, A[iu, ju], device=dev)
    f = factorizer(form)
    f.factorize(S)
    assert f.supernodeStats()["num_supernodes"] == 1 and f.supernodeStats()["max_width"] == n
    order = f.getInversePermutation()
    A32 = A.astype(np.float32).astype(np.float64)[np.ix_(order, order)]
    Lref = np.linalg.cholesky(A32)
    cm = np.abs(Lref).max(axis=0)
    Lg = orc.dense_L(*f.exportFactorCSC())
    err = (np.abs(Lg - Lref).max(axis=0) / cm).max()
    err32 = (np.abs(chol_fp32(A32) - Lref).max(axis=0) / cm).max()
    print(f"form {form}: FP64-pivot tile error {err:.2e}; FP32 pivots would give {err32:.2e}")
    assert err <= 1e-7
    assert err32 > 1e3 * 1e-7
    assert np.array_equal(Lg, Lg.astype(np.float32).astype(np.float64))     # stored in FP32
