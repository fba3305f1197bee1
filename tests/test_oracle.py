"""Pins the CPU oracle: reference golden vectors, closed forms and brute force (CPU only)."""
import itertools
import json
import os

import numpy as np
import pytest

from tests import oracle_lib as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_cube_mesh_geometry():
    N = 3
    tets, X = orc.cube_mesh(N)
    assert tets.shape == (6 * N ** 3, 4) and X.shape == ((N + 1) ** 3, 3)
    # node id formula x + (N+1)(y + (N+1) z)
    M = N + 1
    for nid in [0, 5, 17, M ** 3 - 1]:
        x, y, z = nid % M, (nid // M) % M, nid // (M * M)
        assert np.allclose(X[nid], np.array([x, y, z]) / N)
    D = X[tets[:, 1:]] - X[tets[:, :1]]
    vol = np.linalg.det(np.transpose(D, (0, 2, 1))) / 6
    assert np.all(vol > 0)
    assert abs(vol.sum() - 1.0) < 1e-12
    # every tet contains its hex's main diagonal
    assert np.all(tets[:, 0] + M * M + M + 1 == tets[:, 3]) or np.all(
        np.sort(tets, 1)[:, 0] + M * M + M + 1 == np.sort(tets, 1)[:, 3])


def test_perturb_moments():
    N = 20
    _, X = orc.cube_mesh(N)
    Y = orc.perturb(N, X, 7, 1e-3)
    d = (Y - X).ravel() / (1e-3 / N)
    assert abs(d.mean()) < 0.02 and abs(d.std() - 1.0) < 0.02
    assert np.array_equal(orc.perturb(N, X, 7, 1e-3), Y)
    assert not np.array_equal(orc.perturb(N, X, 8, 1e-3), Y)


def test_energy_closed_forms():
    N = 2
    tets, X = orc.cube_mesh(N)
    assert abs(orc.energy(X, X, tets)) < 1e-9          # zero at rest
    # rigid rotation invariance
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Y = orc.perturb(N, X, 3, 1e-2)
    assert abs(orc.energy(X, Y @ R.T, tets) - orc.energy(X, Y, tets)) < 1e-8 * abs(orc.energy(X, Y, tets))
    # uniform stretch F = s I: psi = mu/2 (3 s^2 - 3) - mu ln s^3 + lam/2 (ln s^3)^2, volume 1
    s, E, nu = 1.05, 1e5, 0.3
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    psi = mu / 2 * (3 * s * s - 3) - mu * np.log(s ** 3) + lam / 2 * np.log(s ** 3) ** 2
    assert abs(orc.energy(X, s * X, tets) - psi) < 1e-9 * psi


def brute_pattern(tets, nb):
    cols = [set() for _ in range(nb)]
    for t in tets:
        for a, b in itertools.product(t, t):
            if a <= b:
                cols[b].add(a)
    Ap = [0]
    Ai = []
    for j in range(nb):
        Ai += sorted(cols[j])
        Ap.append(len(Ai))
    return np.array(Ap), np.array(Ai)


def test_pattern_bruteforce():
    N = 3
    tets, X = orc.cube_mesh(N)
    Ap, Ai = orc.pattern(tets, len(X))
    bAp, bAi = brute_pattern(tets, len(X))
    assert np.array_equal(Ap, bAp) and np.array_equal(Ai, bAi)
    # diagonal last in each column
    assert all(Ai[Ap[j + 1] - 1] == j for j in range(len(X)))


def dense_assembly_upper(tets, He, nb):
    A = np.zeros((3 * nb, 3 * nb))
    for e, t in enumerate(tets):
        h = He[e].reshape(12, 12).T   # column-major
        idx = np.concatenate([3 * n + np.arange(3) for n in t])
        A[np.ix_(idx, idx)] += h
    return A


def ax_to_dense_upper(Ap, Ai, Ax, nb):
    A = np.zeros((3 * nb, 3 * nb))
    for j in range(nb):
        for k in range(Ap[j], Ap[j + 1]):
            i = Ai[k]
            B = Ax[9 * k:9 * k + 9].reshape(3, 3).T
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = B
    return A


def test_assembly_matches_dense_sum():
    rng = np.random.default_rng(0)
    N = 2
    tets, X = orc.cube_mesh(N)
    nb = len(X)
    # scramble node order within elements so both transposition branches are exercised
    tets = np.array([rng.permutation(t) for t in tets], dtype=np.int32)
    He = rng.standard_normal((len(tets), 144))
    Ap, Ai = orc.pattern(tets, nb)
    Ax = orc.assemble(Ap, Ai, nb, tets, He)
    # the reference reads local-upper blocks (transposed as needed) and only the upper
    # triangle of diagonal blocks; with a symmetrised He this equals sum_e P^T He P
    Hs = He.reshape(-1, 12, 12)
    Hs = (Hs + np.transpose(Hs, (0, 2, 1))) / 2
    Ax_s = orc.assemble(Ap, Ai, nb, tets, Hs.reshape(-1, 144))
    Ad = dense_assembly_upper(tets, Hs.reshape(-1, 144), nb)
    G = ax_to_dense_upper(Ap, Ai, Ax_s, nb)
    U = np.triu(np.ones((3 * nb, 3 * nb)))
    assert np.allclose(G * U, Ad * U, rtol=1e-13, atol=1e-12)
    # strict lower parts of diagonal blocks are untouched
    for j in range(nb):
        k = Ap[j + 1] - 1
        B = Ax[9 * k:9 * k + 9]
        assert B[1] == 0 and B[2] == 0 and B[5] == 0
    # with a non-symmetric He only local-upper entries are read: perturbing local-lower
    # entries changes nothing
    He2 = He.reshape(-1, 12, 12).copy()
    low = np.tril(np.ones((12, 12)), -1).astype(bool)
    for e in range(len(He2)):
        He2[e].T[low] += 1.0        # He2[e].T is the column-major view: entries (r > c)
    assert np.array_equal(orc.assemble(Ap, Ai, nb, tets, He2.reshape(-1, 144)), Ax)


def test_row_col_removal():
    N = 2
    tets, X = orc.cube_mesh(N)
    nb = len(X)
    Ap, Ai = orc.pattern(tets, nb)
    removed = np.zeros(nb, np.uint8)
    removed[[0, 4, 7]] = 1
    Rp, Ri, rr, ee = orc.row_col_removal(Ap, Ai, removed)
    assert np.array_equal(rr[removed == 0], np.arange(nb - 3)) and np.all(rr[removed == 1] == -1)
    keep = np.flatnonzero(removed == 0)
    for jr, j in enumerate(keep):
        for q in range(Rp[jr], Rp[jr + 1]):
            k = ee[q]
            assert Ap[j] <= k < Ap[j + 1] and rr[Ai[k]] == Ri[q]
    assert len(ee) == sum(1 for j in keep for k in range(Ap[j], Ap[j + 1]) if not removed[Ai[k]])


def test_spmv_dense():
    rng = np.random.default_rng(1)
    N = 2
    tets, X = orc.cube_mesh(N)
    nb = len(X)
    Ap, Ai = orc.pattern(tets, nb)
    Ax = rng.standard_normal(9 * len(Ai))
    A = ax_to_dense_upper(Ap, Ai, Ax, nb)
    for j in range(nb):
        A[3 * j:3 * j + 3, 3 * j:3 * j + 3] = np.triu(A[3 * j:3 * j + 3, 3 * j:3 * j + 3])
    A = np.triu(A) + np.triu(A, 1).T
    x = rng.standard_normal(3 * nb)
    assert np.allclose(orc.spmv(Ap, Ai, nb, Ax, x), A @ x, rtol=1e-13, atol=1e-12)


def random_spd_lower(n, density, rng):
    A = np.zeros((n, n))
    mask = rng.random((n, n)) < density
    A[mask] = rng.standard_normal(mask.sum())
    A = A + A.T + 2 * n * np.eye(n)
    Cp, Ci, Cx = [0], [], []
    for j in range(n):
        for i in range(j, n):
            if A[i, j] != 0:
                Ci.append(i)
                Cx.append(A[i, j])
        Cp.append(len(Ci))
    return A, np.array(Cp), np.array(Ci), np.array(Cx)


@pytest.mark.parametrize("n,density", [(1, 1.0), (7, 0.5), (40, 0.1), (120, 0.03)])
def test_chol_unique_factor_matches_lapack(n, density):
    rng = np.random.default_rng(n)
    A, Cp, Ci, Cx = random_spd_lower(n, density, rng)
    Lp, Li, Lx = orc.chol(Cp, Ci, Cx)
    L = orc.dense_L(Lp, Li, Lx)
    assert np.allclose(L, np.linalg.cholesky(A), rtol=1e-12, atol=1e-12)
    b = rng.standard_normal(n)
    x = orc.ltsolve(Lp, Li, Lx, orc.lsolve(Lp, Li, Lx, b))
    assert np.allclose(A @ x, b, rtol=1e-11, atol=1e-11)


def test_chol_detects_non_spd():
    # reference test_catamari_factorizer.cc:197-211 matrix, negated: not SPD
    A = -np.array([[1.0, 1.0], [1.0, 1.0 + 1e-8]])
    with pytest.raises(orc.NotSPD):
        orc.chol(np.array([0, 2, 3]), np.array([0, 1, 1]), np.array([A[0, 0], A[1, 0], A[1, 1]]))
    # and the original factors in double precision
    A = -A
    Lp, Li, Lx = orc.chol(np.array([0, 2, 3]), np.array([0, 1, 1]), np.array([A[0, 0], A[1, 0], A[1, 1]]))
    assert np.allclose(orc.dense_L(Lp, Li, Lx), np.linalg.cholesky(A))


def test_multiply_lower_block_reference_kat():
    with open(os.path.join(GOLDEN, "multiply_lower_block_kat.json")) as f:
        kat = json.load(f)
    out = orc.multiply_lower_block(kat["BS"], kat["indices"], kat["supernode_start"], kat["supernode_size"],
                                   kat["degree"], np.array(kat["A"], float), kat["lda"], kat["nrhs"],
                                   np.array(kat["rhs"], float), kat["ldb"])
    assert out.tolist() == kat["expected"]


@pytest.mark.parametrize("BS", [1, 2, 3])
def test_multiply_lower_block_dense(BS):
    # reference test_catamari_solve_kernels.cc:53-90 construction
    rng = np.random.default_rng(BS)
    for n in (BS, 2 * BS, 5 * BS):
        for degree in (0, BS, 3 * BS, 7 * BS):
            for nrhs in (1, 3):
                start = BS
                rows = start + n + 2 * degree
                lda, ldb = degree + 2 * BS, rows + 2 * BS
                A = rng.standard_normal((n, lda)).T.copy()   # column-major lda x n
                B = rng.standard_normal((nrhs, ldb)).T.copy()
                idx = np.zeros(max(1, degree // BS), np.int64)
                upd = np.zeros((rows, rows))
                for i in range(degree):
                    row = start + n + 2 * BS * (i // BS) + i % BS
                    if i % BS == 0:
                        idx[i // BS] = row
                    upd[row, start:start + n] = A[i, :n]
                Bm = B[:rows, :]
                exp = Bm - upd @ Bm
                out = orc.multiply_lower_block(BS, idx, start, n, degree, A.T.ravel(), lda, nrhs, B.T.ravel(), ldb)
                out = out.reshape(nrhs, ldb).T
                assert np.allclose(out[:rows], exp, rtol=1e-12, atol=1e-12)
                assert np.array_equal(out[rows:], B[rows:])
