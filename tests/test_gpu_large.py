"""GPU: the full-size configurations, checked through size-independent properties.

* 64^3 cube (811,200 free DOFs), E = 1e5, nu = 0.3: assembly -> FP64 factorization -> solve of
  RHS = -gradient; the residual is recomputed on the host with the oracle's SpMV
  (BlockCSCHessian.hh:1085-1120 semantics) and must be <= 1e-10 (benchmark_linear_systems.cc:
  316-319 is the reference's own check); factor nnz and flops equal the host-only analysis.
* 64^3 stress configuration (per-element E log-uniform in [1e3, 1e7], nu uniform in
  [0.2, 0.45]): FP64, FP32 + FP64 refinement and the bfloat16 Schur-update forms all reach
  <= 1e-10 (oracle SpMV residual), with bounded refinement steps and a forward error against
  the FP64 factor's solution within the conditioning the mixed forms are built for.
* 512 independent cubes, per-side n ~ U{6..20}, E log-uniform in [1e3, 1e6] (the batching
  configuration): every mesh's segment of the refined batch solution meets 1e-10 by the oracle's
  SpMV, in FP64 and FP32.
"""
import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import _lib, workload as wl
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu


def cube_on_device(dev, N, seed=0, Ee=None, nue=None):
    tets, Xr = wl.cube_mesh(N)
    nb = len(Xr)
    H = mfs.SystemAssembler(nb).blockSparsityPattern(tets)
    H.setZero()
    da = mfs.DeviceAssembler(H, tets)
    T = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    He, g12 = da.neohookean(T(Xr), T(wl.perturb(N, Xr, seed)), E_per_elem=T(Ee), nu_per_elem=T(nue))
    da.assemble(He, H.Ax)
    g = torch.zeros(3 * nb, dtype=torch.float64, device=dev)
    da.assembleGradient(g12, g)
    del He, g12
    return H, wl.z0_pins(N), -g


def oracle_relres(H, x, b, pins):
    """||b - A x|| / ||b|| over the free rows, with the oracle's SpMV on the host."""
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    r = bh - orc.spmv(H.Ap, H.Ai, H.n, H.Ax.cpu().numpy(), xh)
    free = np.ones(3 * H.n, bool)
    free[pins] = False
    return np.linalg.norm(r[free]) / np.linalg.norm(bh[free])


def test_newton_step_64cube_fp64(dev):
    N = 64
    H, pins, b = cube_on_device(dev, N, seed=0)
    f = mfs.make_cholesky_factorizer()
    f.factorizeSymbolic(H, pins)
    assert f.n_reduced() == 811200
    # the device analysis is the host-only analysis (same ordering, supernodes and counts)
    st = np.zeros(8, np.int64)
    fl = np.zeros(2)
    _lib.call("mfs_analyze_host", _lib.ptr(H.Ap, _lib.c_i64), _lib.ptr(H.Ai, _lib.c_i64), H.n,
              _lib.ptr(pins, _lib.c_i64), len(pins), None, _lib.ptr(st, _lib.c_i64), _lib.ptr(fl, _lib.c_double))
    assert f.getFactorNNZ() == st[3]
    assert f.getFlopEstimate() == fl[0]
    assert f.flopBreakdown()["unrelaxed"] == fl[1]
    f.factorizeNumeric(H)
    x, it, rr = f.solveRefined(H, b, tol=1e-10, maxIters=5)
    ro = oracle_relres(H, x, b, pins)
    print(f"64^3 FP64: nnz(L) {st[3]:.4e}, flops {fl[0]:.4e}, refinement steps {it}, relres {rr:.2e} "
          f"(oracle SpMV {ro:.2e})")
    assert rr <= 1e-10 and ro <= 1e-10
    assert it == 0          # the FP64 factor solves to 1e-10 without refinement
    assert x[torch.from_numpy(pins).to(dev)].abs().max() == 0


def test_stress_64cube_precision_forms(dev):
    N = 64
    rng = np.random.default_rng(12345)
    ne = 6 * N ** 3
    Ee, nue = 10 ** rng.uniform(3, 7, ne), rng.uniform(0.2, 0.45, ne)
    H, pins, b = cube_on_device(dev, N, seed=0, Ee=Ee, nue=nue)
    f64 = mfs.make_cholesky_factorizer()
    f64.factorizeSymbolic(H, pins)
    f64.factorizeNumeric(H)
    x64, it64, rr64 = f64.solveRefined(H, b, tol=1e-12, maxIters=5)
    assert oracle_relres(H, x64, b, pins) <= 1e-10
    del f64
    steps = {"fp32": 5, "bf16x3": 10, "bf16": 80}
    for form in ("fp32", "bf16x3", "bf16"):
        f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, True,
                                         None if form == "fp32" else form)
        f.factorizeSymbolic(H, pins)
        f.factorizeNumeric(H)
        x, it, rr = f.solveRefined(H, b, tol=1e-10, maxIters=steps[form])
        ro = oracle_relres(H, x, b, pins)
        fe = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
        print(f"64^3 stress {form}: {it} refinement steps, relres {rr:.2e} (oracle SpMV {ro:.2e}), "
              f"forward error vs FP64 {fe:.2e}")
        assert rr <= 1e-10 and ro <= 1e-10
        assert 1 <= it <= steps[form]
        assert fe <= 1e-5
        del f


@pytest.mark.parametrize("single", [False, True])
def test_batch_512_meshes(dev, single):
    B = wl.batch_meshes(512, 6, 20, seed=0)
    nb = int(B["node_off"][-1])
    tets = B["tets"]
    H = mfs.SystemAssembler(nb).blockSparsityPattern(tets)
    H.setZero()
    da = mfs.DeviceAssembler(H, tets)
    rng = np.random.default_rng(12345)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    He, g12 = da.neohookean(T(B["X_rest"]), T(wl.batch_perturb(B, 0)), E_per_elem=T(10 ** rng.uniform(3, 6, len(tets))),
                            nu=0.3)
    da.assemble(He, H.Ax)
    g = torch.zeros(3 * nb, dtype=torch.float64, device=dev)
    da.assembleGradient(g12, g)
    del He, g12
    b = -g
    f = mfs.make_cholesky_factorizer(mfs.CholeskyProvider.GpuSupernodal, single)
    f.factorizeSymbolic(H, B["pins"])
    assert f.n_reduced() == 3 * nb - len(B["pins"])
    f.factorizeNumeric(H)
    x, it, rr, per = f.solveRefinedSegments(H, b, B["seg"], tol=1e-10, maxIters=30 if single else 10)
    assert rr <= 1e-10 and (per <= 1e-10).all()
    # per-mesh residuals recomputed with the oracle's SpMV
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    r = bh - orc.spmv(H.Ap, H.Ai, H.n, H.Ax.cpu().numpy(), xh)
    free = np.ones(3 * nb, bool)
    free[B["pins"]] = False
    seg = B["seg"]
    worst = 0.0
    for m in range(512):
        s = slice(seg[m], seg[m + 1])
        fr = free[s]
        worst = max(worst, np.linalg.norm(r[s][fr]) / np.linalg.norm(bh[s][fr]))
    print(f"512-mesh batch ({'FP32' if single else 'FP64'}): {3 * nb - len(B['pins'])} free DOFs, {it} refinement "
          f"steps, worst per-mesh relres {worst:.2e}")
    assert worst <= 1e-10
