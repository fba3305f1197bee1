"""GPU: building-block kernels through the C ABI test hooks."""
import json
import os

import numpy as np
import pytest
import torch

from meshfemsparse_amd import _lib
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _s():
    return _lib.c_vp(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (64, 64, 64), (70, 33, 17), (130, 200, 300),
                                   (64, 64, 1000)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0)])
def test_gemm_nt_tile_kernel(dev, M, N, K, alpha, beta):
    rng = np.random.default_rng(M * 1000 + N + K)
    lda, ldb, ldc = M + 3, N + 5, M + 1
    A = rng.standard_normal((K, lda))        # column-major lda x K stored as K rows
    B = rng.standard_normal((K, ldb))
    C = rng.standard_normal((N, ldc))
    Ad, Bd, Cd = (torch.from_numpy(a.ravel().copy()).to(dev) for a in (A, B, C))
    _lib.call("mfs_test_gemm_nt", M, N, K, alpha, _lib.c_vp(Ad.data_ptr()), lda, _lib.c_vp(Bd.data_ptr()), ldb,
              beta, _lib.c_vp(Cd.data_ptr()), ldc, _s())
    Am, Bm, Cm = A[:, :M].T, B[:, :N].T, C[:, :M].T
    exp = alpha * Am @ Bm.T + beta * Cm
    got = Cd.cpu().numpy().reshape(N, ldc)
    assert np.allclose(got[:, :M].T, exp, rtol=1e-13, atol=1e-12 * np.sqrt(K))
    assert np.array_equal(got[:, M:], C[:, M:])       # padding rows untouched


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (64, 64, 64), (70, 33, 17), (130, 200, 300),
                                   (256, 256, 1000)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0)])
def test_gemm_nt_tile_kernel_fp32(dev, M, N, K, alpha, beta):
    """The FP32 tile kernel against an FP64 product of the same float inputs: FP32 rounding of a
    K-term sum bounds the error by about K * 2^-24 * sum |a||b| per entry."""
    rng = np.random.default_rng(M * 1000 + N + K + 7)
    lda, ldb, ldc = M + 3, N + 5, M + 1
    A = rng.standard_normal((K, lda)).astype(np.float32)
    B = rng.standard_normal((K, ldb)).astype(np.float32)
    C = rng.standard_normal((N, ldc)).astype(np.float32)
    Ad, Bd, Cd = (torch.from_numpy(a.ravel().copy()).to(dev) for a in (A, B, C))
    _lib.call("mfs_test_gemm_nt_f32", M, N, K, alpha, _lib.c_vp(Ad.data_ptr()), lda, _lib.c_vp(Bd.data_ptr()),
              ldb, beta, _lib.c_vp(Cd.data_ptr()), ldc, _s())
    Am, Bm, Cm = (x.astype(np.float64) for x in (A[:, :M].T, B[:, :N].T, C[:, :M].T))
    exp = alpha * Am @ Bm.T + beta * Cm
    bound = (K + 2) * 2.0 ** -24 * (np.abs(Am) @ np.abs(Bm).T + np.abs(Cm))
    got = Cd.cpu().numpy().reshape(N, ldc)
    assert (np.abs(got[:, :M].T - exp) <= bound).all()
    assert np.array_equal(got[:, M:], C[:, M:])       # padding rows untouched


def _bf16(x):
    """float32 -> nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


# shapes of at least 256 in both dimensions run on the 256x256 tile kernel (512 threads)
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (64, 64, 64), (70, 33, 17), (130, 200, 300),
                                   (256, 256, 1000), (300, 260, 100), (513, 300, 77)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.0, 1.0)])
@pytest.mark.parametrize("split", [0, 1])
def test_gemm_nt_tile_kernel_bf16(dev, M, N, K, alpha, beta, split):
    """The bfloat16 tile kernel of the BF16 Schur-update variants.  Each operand is rounded to
    bfloat16 (ties to even); with split the rounding residual is rounded again and the product is
    hi*hi + hi*lo + lo*hi.  bfloat16 products are exact in FP32, so the kernel equals the FP64
    product of the rounded operands up to FP32 accumulation, (terms + 2) 2^-24 sum |a||b|; the
    split form is also within 2^-15 sum |a||b| of the exact product of the float inputs."""
    rng = np.random.default_rng(M * 1000 + N + K + 11 + split)
    lda, ldb, ldc = M + 3, N + 5, M + 1
    A = rng.standard_normal((K, lda)).astype(np.float32)
    B = rng.standard_normal((K, ldb)).astype(np.float32)
    C = rng.standard_normal((N, ldc)).astype(np.float32)
    Ad, Bd, Cd = (torch.from_numpy(a.ravel().copy()).to(dev) for a in (A, B, C))
    _lib.call("mfs_test_gemm_nt_bf16", M, N, K, alpha, _lib.c_vp(Ad.data_ptr()), lda, _lib.c_vp(Bd.data_ptr()),
              ldb, beta, _lib.c_vp(Cd.data_ptr()), ldc, split, _s())
    Am, Bm = A[:, :M].T, B[:, :N].T
    Cm = C[:, :M].T.astype(np.float64)
    Ah, Bh = _bf16(Am), _bf16(Bm)
    terms = [(Ah, Bh)]
    if split:
        Al, Bl = _bf16(Am - Ah), _bf16(Bm - Bh)
        terms += [(Ah, Bl), (Al, Bh)]
    prod = sum(a.astype(np.float64) @ b.astype(np.float64).T for a, b in terms)
    mag = sum(np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64)).T for a, b in terms)
    exp = alpha * prod + beta * Cm
    bound = (len(terms) * K + 2) * 2.0 ** -24 * (mag + np.abs(Cm))
    got = Cd.cpu().numpy().reshape(N, ldc)
    assert (np.abs(got[:, :M].T - exp) <= bound).all()
    assert np.array_equal(got[:, M:], C[:, M:])       # padding rows untouched
    if split:
        exact = Am.astype(np.float64) @ Bm.astype(np.float64).T
        absab = np.abs(Am.astype(np.float64)) @ np.abs(Bm.astype(np.float64)).T
        assert (np.abs(prod - exact) <= 2.0 ** -15 * absab).all()


def test_multiply_lower_block_reference_kat_on_device(dev):
    with open(os.path.join(GOLDEN, "multiply_lower_block_kat.json")) as f:
        kat = json.load(f)
    idx = torch.tensor(kat["indices"], dtype=torch.int64, device=dev)
    A = torch.tensor(kat["A"], dtype=torch.float64, device=dev)
    B = torch.tensor(kat["rhs"], dtype=torch.float64, device=dev)
    _lib.call("mfs_test_multiply_lower_block", kat["BS"], _lib.c_vp(idx.data_ptr()), kat["supernode_start"],
              kat["supernode_size"], kat["degree"], _lib.c_vp(A.data_ptr()), kat["lda"], kat["nrhs"],
              _lib.c_vp(B.data_ptr()), kat["ldb"], _s())
    assert B.cpu().tolist() == kat["expected"]


@pytest.mark.parametrize("BS", [1, 2, 3])
def test_multiply_lower_block_matches_oracle(dev, BS):
    rng = np.random.default_rng(10 + BS)
    n, degree, nrhs, start = 5 * BS, 7 * BS, 3, BS
    rows = start + n + 2 * degree
    lda, ldb = degree + 2 * BS, rows + 2 * BS
    A = rng.standard_normal(lda * n)
    B = rng.standard_normal(ldb * nrhs)
    idx = np.array([start + n + 2 * BS * q for q in range(degree // BS)], np.int64)
    exp = orc.multiply_lower_block(BS, idx, start, n, degree, A, lda, nrhs, B, ldb)
    Bd = torch.from_numpy(B.copy()).to(dev)
    Ad = torch.from_numpy(A).to(dev)
    idd = torch.from_numpy(idx).to(dev)
    _lib.call("mfs_test_multiply_lower_block", BS, _lib.c_vp(idd.data_ptr()), start, n, degree,
              _lib.c_vp(Ad.data_ptr()), lda, nrhs, _lib.c_vp(Bd.data_ptr()), ldb, _s())
    assert np.allclose(Bd.cpu().numpy(), exp, rtol=1e-13, atol=1e-13)
