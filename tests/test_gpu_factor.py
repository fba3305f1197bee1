"""GPU: supernodal factorization and solves against the oracle and the reference's own tests.

The Cholesky factor of a fixed ordering is unique, so the device factor is compared entrywise
with the oracle's up-looking factor under the ordering the device analysis chose; solves follow
the reference's checkSolves (tests/test_catamari_factorizer.cc:13-98): known solutions
x_ij = sin(0.2 (i+1)(j+1)), multi-RHS and single solves, raw/permuted/in-place reduced solves,
shifts, shift-by-matrix, stash/swap, rescaling, failure on -A and recovery; tolerance 1e-11
relative to ||x|| in double precision and 2e-5 in single precision, the reference's own bounds.

Single precision factors the symmetrically scaled matrix S A S (S = diag(A)^-1/2) in FP32 and
solves in FP32; the reference's single-precision factorizers cast A to float and factor that
(CatamariFactorizer.cc:640-660), so both lose definiteness on the same rounding case
(test_catamari_factorizer.cc:197-212).
"""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import workload as wl
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu
TOL = 1e-11
TOL_SINGLE = 2e-5


def mesh_problem(dev, N, seed=0, E=1e5, nu=0.3, Ee=None, nue=None):
    tets, Xr = wl.cube_mesh(N)
    X = wl.perturb(N, Xr, seed)
    He, g, _ = wl.neohookean(Xr, X, tets, E=E, nu=nu, E_per_elem=Ee, nu_per_elem=nue)
    nb = len(Xr)
    asm = mfs.SystemAssembler(nb)
    H = asm.blockSparsityPattern(tets)
    H.setZero()
    asm.assembleHessian(H, tets, torch.from_numpy(He.ravel()).to(dev))
    return H, wl.z0_pins(N), g, tets, He


def chain_problem(dev, nb=20):
    """Reference checkBlocks<3>: chain stencils {i, i+1}, scalar values 10.1 on the diagonal,
    -0.1 (1 + (i + j) % 5) elsewhere (test_catamari_factorizer.cc:100-122)."""
    en = np.array([[i, i + 1] for i in range(nb - 1)], np.int32)
    H = mfs.SystemAssembler(nb).blockSparsityPattern(en)
    Ax = np.zeros(9 * H.nz)
    for j in range(nb):
        for k in range(H.Ap[j], H.Ap[j + 1]):
            bi = H.Ai[k]
            for c in range(3):
                for r in range(3):
                    if bi == j and r > c:
                        continue
                    i_, j_ = 3 * bi + r, 3 * j + c
                    Ax[9 * k + 3 * c + r] = 10.1 if i_ == j_ else -0.1 * (1 + (i_ + j_) % 5)
    H.Ax = torch.from_numpy(Ax).to(dev)
    return H


def oracle_factor(H, f, sigma=0.0, Bx=None):
    ip = f.getInversePermutation()
    nred = len(ip) // 3
    iperm_b = ip[0::3] // 3
    pos = np.empty(nred, np.int32)
    pos[iperm_b] = np.arange(nred)
    pinned = np.zeros(H.n, bool)
    pinned[np.array(f.getFixedVars(), dtype=np.int64) // 3] = True
    red = -np.ones(H.n, np.int64)
    red[~pinned] = np.arange(nred)
    Ax = H.Ax.cpu().numpy()
    Cp, Ci, Cx = orc.permuted_lower(H.Ap, H.Ai, H.n, Ax, red, pos, sigma, Bx)
    return orc.chol(Cp, Ci, Cx), red, pos


@pytest.mark.parametrize("N,outer,fuse,big", [(2, None, 0, 128), (5, None, 0, 128), (8, None, 0, 128),
                                              (8, 128, 0, 128), (9, 64, 0, 128), (5, None, 100000, 128),
                                              (8, None, 240, 128), (9, 64, 100000, 128), (12, None, 0, 256),
                                              (12, None, 600, 512)])
def test_factor_matches_unique_cholesky(dev, N, outer, fuse, big, monkeypatch):
    """fuse > 0 sends every supernode one diagonal tile wide with at most `fuse` structure rows
    through the fused level kernel (one workgroup per supernode); big is the region size from
    which the GEMM tasks use 128x128 tiles."""
    if outer:
        monkeypatch.setenv("MFS_OUTER", str(outer))      # several outer blocks per supernode
    monkeypatch.setenv("MFS_FUSE_DEGREE", str(fuse))
    monkeypatch.setenv("MFS_BIG_TILE", str(big))
    H, pins, _, _, _ = mesh_problem(dev, N)
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(H, pins)
    f.factorizeNumeric(H)
    (Lp, Li, Lx), _, _ = oracle_factor(H, f)
    Lo = orc.dense_L(Lp, Li, Lx)
    Lg = orc.dense_L(*f.exportFactorCSC())
    assert np.abs(Lg - Lo).max() <= 1e-12 * np.abs(Lo).max()
    # run-to-run bitwise determinism of the factor
    f.factorizeNumeric(H)
    assert np.array_equal(orc.dense_L(*f.exportFactorCSC()), Lg)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("N,panel", [(8, 4), (9, 512), (12, 40)])
def test_fused_panel_steps_match_separate_launches(dev, N, panel, single, monkeypatch):
    """Steps whose TRSM needs at most `panel` row tiles run as one launch (every workgroup
    factors the diagonal tile in LDS, the last one to load it stores it).  In FP64 the fused TRSM
    forms the same MFMA products in the same order as the separate launch, so the factor is
    bitwise the one of the unfused schedule; in FP32 it agrees to FP32 accuracy.  Refactoring
    reproduces it bit for bit (the shared-tile counters are left at zero)."""
    H, pins, _, _, _ = mesh_problem(dev, N)

    def factor(pm):
        monkeypatch.setenv("MFS_PANEL_MAX", str(pm))
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
        f.factorizeSymbolic(H, pins)
        f.factorizeNumeric(H)
        return f, orc.dense_L(*f.exportFactorCSC())

    f0, L0 = factor(0)
    f1, L1 = factor(panel)
    st0, st1 = f0.scheduleStats(), f1.scheduleStats()
    assert st0["panel_launches"] == 0 and st1["panel_launches"] > 0
    assert st1["panel_launches"] + st1["potrf_launches"] == st0["potrf_launches"]
    if single:
        (Lp, Li, Lx), _, _ = oracle_factor(H, f1)
        Lo = orc.dense_L(Lp, Li, Lx)
        cn = np.abs(Lo).max(axis=0)
        assert (np.abs(L1 - Lo).max(axis=0) <= 2e-5 * cn).all()
    else:
        assert np.array_equal(L1, L0)
    for _ in range(2):
        f1.factorizeNumeric(H)
        assert np.array_equal(orc.dense_L(*f1.exportFactorCSC()), L1)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("outer", [128, 192])
def test_lookahead_matches_in_order_schedule(dev, outer, single, monkeypatch):
    """With lookahead the rest of every outer trailing update runs on a second stream beside the
    next outer block's steps; each entry still receives its updates in block order, so the factor
    is bitwise the in-order schedule's, eagerly and replayed as a captured graph."""
    H, pins, _, _, _ = mesh_problem(dev, 10)
    monkeypatch.setenv("MFS_OUTER", str(outer))

    def factor(la, graph=False):
        monkeypatch.setenv("MFS_LOOKAHEAD", str(la))
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
        f.factorizeSymbolic(H, pins)
        if graph:
            f.setGraphCapture(True)
        f.factorizeNumeric(H)
        L1 = orc.dense_L(*f.exportFactorCSC())
        if graph:
            f.factorizeNumeric(H)          # a replay of the captured graph
            assert np.array_equal(orc.dense_L(*f.exportFactorCSC()), L1)
        return L1

    L0 = factor(0)
    assert np.array_equal(factor(1), L0)
    assert np.array_equal(factor(1, graph=True), L0)


@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("N,fuse", [(10, "0"), (14, "0"), (14, "300")])
def test_tile_children_lists_match_extend_add(dev, N, fuse, form, monkeypatch):
    """The Schur-update tiles add the children's update-matrix blocks listed on them in their
    epilogue (FP64 and FP32 forms), and the panel tiles those that land in the parents' panels
    (every form), in child order, instead of the extend-add passes: the same additions in the
    same order, so the factor is bitwise the one of the passes, eagerly and replayed as a captured
    graph."""
    H, pins, _, _, _ = mesh_problem(dev, N)
    monkeypatch.setenv("MFS_FUSE_DEGREE", fuse)

    def factor(pull, graph=False):
        monkeypatch.setenv("MFS_PULL", str(pull))
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, form != "fp64",
                                         "bf16x3" if form == "bf16x3" else None)
        f.factorizeSymbolic(H, pins)
        if graph:
            f.setGraphCapture(True)
        f.factorizeNumeric(H)
        Lp, Li, Lx = (np.asarray(a).copy() for a in f.exportFactorCSC())
        if graph:
            f.factorizeNumeric(H)          # a replay of the captured graph
            assert np.array_equal(np.asarray(f.exportFactorCSC()[2]), Lx)
        return (Lp, Li, Lx), f.scheduleStats()["extend_launches"]

    (Lp0, Li0, Lx0), e0 = factor(0)
    (Lp1, Li1, Lx1), e1 = factor(1)
    assert e1 < e0                         # the passes into U are gone
    assert np.array_equal(Lp0, Lp1) and np.array_equal(Li0, Li1)
    assert np.array_equal(Lx0.view(np.uint8), Lx1.view(np.uint8))
    (_, _, Lx2), _ = factor(1, graph=True)
    assert np.array_equal(Lx2.view(np.uint8), Lx0.view(np.uint8))


@pytest.mark.parametrize("single", [False, True])
def test_small_critical_updates_on_64_tiles_bitwise(dev, single, monkeypatch):
    """An update launch on the per-step critical path with few 128x128 tiles runs on 64x64 tiles
    (MFS_SPLIT_BIG: the tile-count bound).  Either tile adds the same products in the same order,
    so the factor is bitwise the one of the 128x128 tiles."""
    H, pins, _, _, _ = mesh_problem(dev, 16)

    def factor(split):
        monkeypatch.setenv("MFS_SPLIT_BIG", str(split))
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single, None)
        f.factorizeSymbolic(H, pins)
        f.factorizeNumeric(H)
        return np.asarray(f.exportFactorCSC()[2]).copy()

    # at 16^3 the top separator (867 columns) gets a next-outer-block trailing update of a few
    # 128x128 tiles (512-column outer blocks), which the large bound splits
    L0 = factor(0)
    L1 = factor(100000)
    assert np.array_equal(L0.view(np.uint8), L1.view(np.uint8))


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("N", [10, 16])
def test_direct_load_tile_gemm_matches_register_staged(dev, N, single, monkeypatch):
    """The tile GEMM that loads its operand chunks straight into LDS (MFS_GEMM_V=3: every tile;
    the default uses it for the 128x128 tiles, and for the 64x64 ones in FP64) and the one that
    stages them through registers (MFS_GEMM_V=1: every tile) issue the same MFMAs in the same
    order, so they give the same factor, bit for bit, eagerly and replayed as a captured graph."""
    H, pins, _, _, _ = mesh_problem(dev, N)

    def factor(v, graph=False):
        monkeypatch.setenv("MFS_GEMM_V", str(v))
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
        f.factorizeSymbolic(H, pins)
        if graph:
            f.setGraphCapture(True)
        f.factorizeNumeric(H)
        Lx = np.asarray(f.exportFactorCSC()[2]).copy()
        if graph:
            f.factorizeNumeric(H)
            assert np.array_equal(np.asarray(f.exportFactorCSC()[2]).view(np.uint8), Lx.view(np.uint8))
        return Lx

    L1 = factor(1)
    assert np.array_equal(factor(3).view(np.uint8), L1.view(np.uint8))
    assert np.array_equal(factor(3, graph=True).view(np.uint8), L1.view(np.uint8))


def test_fused_panel_step_reports_failing_pivot(dev, monkeypatch):
    """-A through the fused steps: the pivot failure is recorded once by the storing workgroup
    and raised like in the separate launches."""
    monkeypatch.setenv("MFS_PANEL_MAX", "512")
    H, pins, _, _, _ = mesh_problem(dev, 6)
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(H, pins)
    assert f.scheduleStats()["panel_launches"] > 0
    H.Ax.neg_()
    with pytest.raises(mfs.NotPositiveDefiniteError):
        f.factorizeNumeric(H)
    assert not f.checkPosDef()
    H.Ax.neg_()
    f.factorizeNumeric(H)
    assert f.checkPosDef()


def dense_of(H):
    return H.toDense()


def check_solves(dev, H, pins, single=False):
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
    TOL = TOL_SINGLE if single else globals()["TOL"]
    A = dense_of(H)
    n = A.shape[0]
    expected = np.array([[np.sin(0.2 * (i + 1) * (j + 1)) for j in range(3)] for i in range(n)])
    expected[np.array(pins, dtype=np.int64)] = 0
    rhs = A @ expected
    free = np.setdiff1d(np.arange(n), pins)
    assert np.allclose(rhs[free], A[np.ix_(free, free)] @ expected[free])
    E = torch.from_numpy(expected).to(dev)

    def verify(b):
        bt = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
        X = f.solveMultiRHS(bt)
        assert torch.linalg.norm(X - E) < TOL * torch.linalg.norm(E)
        b0 = bt[:, 0].contiguous()
        x0 = f.solve(b0)
        e0 = E[:, 0]
        assert torch.linalg.norm(x0 - e0) < TOL * torch.linalg.norm(e0)
        for permuted in (False, True):
            red = f.removeFixedEntries(b0, permuted)
            rx = torch.empty(f.n_reduced(), dtype=torch.float64, device=dev)
            f.solveRawReduced(red, rx, mfs.CholeskySys.A, permuted)
            full = f.extractFullSolution(rx, permuted)
            assert torch.linalg.norm(full - e0) < TOL * torch.linalg.norm(e0)
            f.solveRawReducedInPlace(red, mfs.CholeskySys.A, permuted)
            full = f.extractFullSolution(red, permuted)
            assert torch.linalg.norm(full - e0) < TOL * torch.linalg.norm(e0)

    for repeat in range(2):
        f.factorizeSymbolic(H, pins)
        assert f.getFactorNNZ() > 0 and f.getFlopEstimate() > 0
        f.factorizeNumeric(H)
        verify(rhs)
        f.stashFactorization()
        assert f.hasStashedFactorization()
        for sigma in (0.75, 0.25):
            f.factorizeNumericWithShift(H, sigma)
            verify(rhs + sigma * expected)
        f.factorizeNumericWithShift(H, 0.5, H)     # A itself as a same-pattern shift matrix
        verify(1.5 * rhs)
        f.swapStashedFactorization()
        verify(rhs)
        f.swapStashedFactorization()
        verify(1.5 * rhs)
        f.clearStashedFactorization()
        assert not f.hasStashedFactorization()
        H.Ax *= 2.0
        f.factorizeNumeric(H)
        verify(2.0 * rhs)
        H.Ax *= 0.5
    f.clearFactors()
    assert not f.hasNumericFactorization()
    f.factorizeSymbolic(H, pins)
    f.factorizeNumeric(H)
    verify(rhs)
    # failure must invalidate the previous numeric factor and allow recovery
    H.Ax *= -1.0
    with pytest.raises(RuntimeError):
        f.factorizeNumeric(H)
    assert not f.hasNumericFactorization()
    with pytest.raises(RuntimeError):
        f.solve(torch.from_numpy(rhs[:, 0].copy()).to(dev))
    H.Ax *= -1.0
    f.factorizeNumeric(H)
    verify(rhs)
    f.stashFactorization()
    f.clearFactors()
    f.swapStashedFactorization()
    verify(rhs)
    f.factorizeNumeric(H)
    verify(rhs)
    f.clearStashedFactorization()


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("pin_mode", [0, 1])
def test_check_solves_chain(dev, pin_mode, single):
    H = chain_problem(dev)
    pins = [0, 1, 2] if pin_mode == 1 else []
    check_solves(dev, H, pins, single)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("fuse", [0, 100000])
def test_check_solves_mesh(dev, single, fuse, monkeypatch):
    monkeypatch.setenv("MFS_FUSE_DEGREE", str(fuse))
    H, pins, _, _, _ = mesh_problem(dev, 4, seed=3)
    check_solves(dev, H, pins, single)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("block,chunk", [(64, 64), (128, 64), (320, 128)])
def test_solves_with_small_solve_blocks(dev, block, chunk, single, monkeypatch):
    """Several column blocks per supernode and several backward partial slots per block."""
    monkeypatch.setenv("MFS_SOLVE_BLOCK", str(block))
    monkeypatch.setenv("MFS_SOLVE_CHUNK", str(chunk))
    H, pins, _, _, _ = mesh_problem(dev, 7, seed=2)
    check_solves(dev, H, pins, single)


@pytest.mark.parametrize("N", [5, 8])
def test_single_factor_matches_unique_cholesky(dev, N):
    """The FP32 factor, unscaled, against the FP64 oracle factor: agreement to FP32 accuracy,
    and bitwise run-to-run determinism."""
    H, pins, _, _, _ = mesh_problem(dev, N)
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, True)
    f.factorizeSymbolic(H, pins)
    f.factorizeNumeric(H)
    (Lp, Li, Lx), _, _ = oracle_factor(H, f)
    Lo = orc.dense_L(Lp, Li, Lx)
    Lg = orc.dense_L(*f.exportFactorCSC())
    # column-relative error: each column of L scales with sqrt of its pivot
    cn = np.abs(Lo).max(axis=0)
    assert (np.abs(Lg - Lo).max(axis=0) <= 2e-5 * cn).all()
    assert not np.array_equal(Lg, Lo)          # really computed in single precision
    f.factorizeNumeric(H)
    assert np.array_equal(orc.dense_L(*f.exportFactorCSC()), Lg)


def one_block(dev, M):
    """A single 3x3 diagonal block holding the upper triangle of M (column-major)."""
    H = mfs.BlockCSCHessian(np.array([0, 1]), np.array([0]), dev)
    H.Ax = torch.tensor(np.asarray(M, np.float64).T.ravel(), device=dev)
    return H


def test_single_rounding_loss_of_definiteness(dev):
    """[[1, 1], [1, 1 + 1e-8]] is SPD in double but singular once rounded to float
    (test_catamari_factorizer.cc:197-212), embedded as the leading 2x2 of one 3x3 block."""
    H = one_block(dev, [[1, 1, 0], [1, 1 + 1e-8, 0], [0, 0, 1]])
    dbl = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, False)
    single = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, True)
    dbl.factorize(H)
    assert dbl.hasNumericFactorization()
    with pytest.raises(mfs.NotPositiveDefiniteError):
        single.factorize(H)
    assert not single.hasNumericFactorization()
    b = torch.ones(3, dtype=torch.float64, device=dev)
    x = torch.empty(3, dtype=torch.float64, device=dev)
    with pytest.raises(RuntimeError):
        single.solveRawReduced(b, x)
    with pytest.raises(RuntimeError):
        single.solveMultiRHS(torch.ones(3, 2, dtype=torch.float64, device=dev))


def test_single_non_positive_diagonal_fails(dev):
    """A non-positive diagonal entry cannot be scaled: reported as a pivot failure."""
    H = one_block(dev, [[4, 0, 0], [0, -1, 0], [0, 0, 1]])
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, True)
    with pytest.raises(mfs.NotPositiveDefiniteError):
        f.factorize(H)
    assert f.numSuccessfulPivots() == 1


def test_single_refined_newton_step_16cube(dev):
    """FP32 factor + FP64 iterative refinement on a stiffness field spread over four decades
    (per-element E log-uniform in [1e3, 1e7], nu uniform in [0.2, 0.45]): relative residual
    <= 1e-10, and the solution agrees with the FP64 factor's."""
    N = 16
    rng = np.random.default_rng(7)
    ne = 6 * N ** 3
    Ee, nue = 10 ** rng.uniform(3, 7, ne), rng.uniform(0.2, 0.45, ne)
    H, pins, g, _, _ = mesh_problem(dev, N, seed=1, Ee=Ee, nue=nue)
    b = torch.from_numpy(-g).to(dev)
    f64 = mfs.make_cholesky_factorizer()
    f64.factorize(H, pins)
    x64, it64, rr64 = f64.solveRefined(H, b, tol=1e-10, maxIters=5)
    f32 = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, True)
    f32.factorize(H, pins)
    x0 = f32.solve(b)                     # unrefined single-precision solve
    x32, it32, rr32 = f32.solveRefined(H, b, tol=1e-10, maxIters=30)
    assert rr64 <= 1e-10 and rr32 <= 1e-10
    assert 1 <= it32 <= 30
    e0 = torch.linalg.norm(x0 - x64) / torch.linalg.norm(x64)
    e1 = torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64)
    assert e1 < 1e-6 and e1 < e0


@pytest.mark.parametrize("single", [False, True])
def test_check_solves_chain_partial_pin(dev, single):
    """The reference's pin mode 2 (test_catamari_factorizer.cc:121): one component of block 0
    pinned.  The block stays with that component masked (the reference's scalar fallback,
    CatamariFactorizer.cc:207-216, solves the same reduced system)."""
    H = chain_problem(dev)
    check_solves(dev, H, [1], single)


@pytest.mark.parametrize("single", [False, True])
def test_check_solves_mesh_partial_pins(dev, single):
    """A mesh with whole-block (z = 0) pins and single components pinned on other nodes."""
    H, pins, _, _, _ = mesh_problem(dev, 4, seed=5)
    extra = [3 * 40 + 0, 3 * 41 + 2, 3 * 57 + 1, 3 * 57 + 2, 3 * 90 + 1]
    check_solves(dev, H, sorted(set(list(pins) + extra)), single)


def scalar_chain(dev, BS, nb=20):
    """checkBlocks<BS> of the reference (test_catamari_factorizer.cc:100-122) as the scalar
    matrix it expands to: chain stencils {i, i+1} of BS x BS blocks, 10.1 on the diagonal and
    -0.1 (1 + (i + j) % 5) elsewhere."""
    rows, cols, vals = [], [], []
    for b in range(nb):
        for bj in (b, b + 1):
            if bj >= nb:
                continue
            for r in range(BS):
                for c in range(BS):
                    i, j = BS * b + r, BS * bj + c
                    if i <= j:
                        rows.append(i)
                        cols.append(j)
                        vals.append(10.1 if i == j else -0.1 * (1 + (i + j) % 5))
    return mfs.SuiteSparseMatrix.fromTriplets(BS * nb, rows, cols, vals, device=dev)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("BS", [1, 2])
@pytest.mark.parametrize("pin_mode", [0, 1, 2])
def test_check_solves_scalar_blocks(dev, BS, pin_mode, single):
    """The reference's "scalar blocks" / "blocks of size two" sections through the scalar
    SuiteSparseMatrix overloads (CholeskyFactorizerBase.hh:95-113)."""
    A = scalar_chain(dev, BS)
    pins = list(range(BS)) if pin_mode == 1 else [1] if pin_mode == 2 else []
    check_solves(dev, A, pins, single)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("pinned", [False, True])
def test_check_solves_plain_scalar_matrix(dev, pinned, single):
    """The reference's "plain scalar matrix" section (test_catamari_factorizer.cc:152-161): a
    20 x 20 tridiagonal matrix, 4.1 on the diagonal and -0.7 beside it, pins {} or {1}."""
    n = 20
    rows = list(range(n)) + list(range(n - 1))
    cols = list(range(n)) + list(range(1, n))
    vals = [4.1] * n + [-0.7] * (n - 1)
    A = mfs.SuiteSparseMatrix.fromTriplets(n, rows, cols, vals, device=dev)
    check_solves(dev, A, [1] if pinned else [], single)


def test_scalar_matrix_factor_matches_lapack(dev):
    """A random sparse SPD scalar matrix (n = 61, not a multiple of 3) with pins: the exported
    factor is the unique Cholesky factor of the reduced matrix in the analysis order, which
    getInversePermutation reports (LAPACK via numpy), and the solve matches a dense solve."""
    rng = np.random.default_rng(3)
    n = 61
    M = (rng.random((n, n)) < 0.08) * rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    iu, ju = np.nonzero(np.triu(A))
    S = mfs.SuiteSparseMatrix.fromTriplets(n, iu, ju, A[iu, ju], device=dev)
    pins = [0, 5, 6, 60]
    f = mfs.make_cholesky_factorizer()
    f.factorize(S, pins)
    free = np.setdiff1d(np.arange(n), pins)
    assert f.n_reduced() == len(free)
    order = f.getInversePermutation()            # factor position -> reduced index
    assert np.array_equal(np.sort(order), np.arange(len(free)))
    Ared = A[np.ix_(free, free)][np.ix_(order, order)]
    Lref = np.linalg.cholesky(Ared)
    Lg = orc.dense_L(*f.exportFactorCSC())
    assert np.abs(Lg - Lref).max() <= 1e-12 * np.abs(Lref).max()
    b = rng.standard_normal(n)
    x = f.solve(torch.from_numpy(b).to(dev)).cpu().numpy()
    xo = np.zeros(n)
    xo[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
    y = f.apply(S, torch.from_numpy(x).to(dev)).cpu().numpy()
    assert np.linalg.norm(y - A @ x) <= 1e-13 * np.linalg.norm(A @ x)
    xr, it, rr = f.solveRefined(S, torch.from_numpy(b).to(dev), tol=1e-12)
    assert rr <= 1e-12


def test_solve_without_factor_raises(dev):
    f = mfs.make_cholesky_factorizer()
    with pytest.raises(RuntimeError):
        f.solve(torch.zeros(6, dtype=torch.float64, device=dev))


def test_spmv_matches_oracle(dev):
    H, pins, _, _, _ = mesh_problem(dev, 5)
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(H, pins)
    x = np.random.default_rng(0).standard_normal(3 * H.n)
    y = f.apply(H, torch.from_numpy(x).to(dev)).cpu().numpy()
    yo = orc.spmv(H.Ap, H.Ai, H.n, H.Ax.cpu().numpy(), x)
    assert np.linalg.norm(y - yo) <= 1e-14 * np.linalg.norm(yo) * 10


def test_newton_step_16cube(dev):
    """The 16^3 CPU configuration (13,872 DOFs): assemble, factor, solve RHS = -gradient to
    <= 1e-10 relative residual; solution matches the oracle's factor solve."""
    N = 16
    H, pins, g, _, _ = mesh_problem(dev, N, seed=0)
    assert 3 * H.n - len(pins) == 13872
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(H, pins)
    f.factorizeNumeric(H)
    b = torch.from_numpy(-g).to(dev)
    x, iters, rr = f.solveRefined(H, b, tol=1e-10, maxIters=5)
    assert rr <= 1e-10
    (Lp, Li, Lx), red, pos = oracle_factor(H, f)
    # oracle solve in factor order
    ip = f.getInversePermutation()
    free = np.setdiff1d(np.arange(3 * H.n), pins)
    bred = -g[free]
    bp = bred[ip]
    xp = orc.ltsolve(Lp, Li, Lx, orc.lsolve(Lp, Li, Lx, bp))
    xo = np.zeros(3 * H.n)
    xo[free[ip]] = xp
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-11 * np.linalg.norm(xo)
