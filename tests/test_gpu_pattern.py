"""GPU: the block sparsity pattern built on the device (mfs_pattern_build_device) equals the host
construction (mfs_pattern_build, itself pinned by brute-force pair sets against the oracle) for
cube meshes, a batch of meshes, shuffled and relabelled connectivity, other element sizes, repeated
and unused nodes; bad node indices raise."""
import time

import numpy as np
import pytest
import torch

import meshfemsparse_amd as mfs
from meshfemsparse_amd import _lib
from meshfemsparse_amd import workload as wl
from tests import oracle_lib as orc

pytestmark = pytest.mark.gpu


def both(en, nb, dev):
    asm = mfs.SystemAssembler(nb)
    Hh = asm.blockSparsityPattern(np.asarray(en))
    Hd = asm.blockSparsityPattern(torch.as_tensor(np.asarray(en, np.int32)).to(dev))
    return Hh, Hd


def check_equal(Hh, Hd):
    assert np.array_equal(Hh.Ap, Hd.Ap)
    assert np.array_equal(Hh.Ai, Hd.Ai)


@pytest.mark.parametrize("N", [1, 2, 5, 9])
def test_cube_pattern_matches_host(dev, N):
    tets, Xr = wl.cube_mesh(N)
    Hh, Hd = both(tets, len(Xr), dev)
    check_equal(Hh, Hd)
    Ap, Ai = orc.pattern(tets, len(Xr))          # the oracle's pair-set construction
    assert np.array_equal(Hd.Ap, Ap) and np.array_equal(Hd.Ai, Ai)


def test_batch_pattern_matches_host(dev):
    B = wl.batch_meshes(12, 3, 7, seed=4)
    nb = int(B["node_off"][-1])
    Hh, Hd = both(B["tets"], nb, dev)
    check_equal(Hh, Hd)


@pytest.mark.parametrize("seed", [0, 1])
def test_shuffled_relabelled_connectivity(dev, seed):
    rng = np.random.default_rng(seed)
    tets, Xr = wl.cube_mesh(6)
    nb = len(Xr) + 17                                # trailing unused nodes: empty columns
    relabel = rng.permutation(nb)[: len(Xr)]
    en = relabel[tets][rng.permutation(len(tets))]
    en = np.array([rng.permutation(row) for row in en])     # local node order shuffled too
    Hh, Hd = both(en, nb, dev)
    check_equal(Hh, Hd)


@pytest.mark.parametrize("npe", [1, 2, 3, 8, 10, 16])
def test_other_element_sizes_and_repeated_nodes(dev, npe):
    rng = np.random.default_rng(npe)
    nb = 300
    en = rng.integers(0, nb, size=(2000, npe))     # random nodes, repeats inside elements allowed
    Hh, Hd = both(en, nb, dev)
    check_equal(Hh, Hd)


def test_no_elements(dev):
    Hh, Hd = both(np.zeros((0, 4), np.int32), 10, dev)
    check_equal(Hh, Hd)
    assert Hd.Ap[-1] == 0


def test_device_arrays_and_bad_index(dev):
    tets, Xr = wl.cube_mesh(3)
    en = torch.as_tensor(tets).to(dev)
    lib = _lib.load()
    h = _lib.c_vp()
    _lib.check(lib.mfs_pattern_build_device(_lib.c_vp(en.data_ptr()), len(tets), 4, len(Xr),
                                            _lib.c_vp(torch.cuda.current_stream().cuda_stream), h))
    try:
        Apd, Aid = _lib.P_i64(), _lib.P_i64()
        _lib.check(lib.mfs_pattern_get_device(h, Apd, Aid))
        nb_, nz = _lib.c_i64(), _lib.c_i64()
        Ap, Ai = _lib.P_i64(), _lib.P_i64()
        _lib.check(lib.mfs_pattern_get(h, nb_, nz, Ap, Ai))
        ptr_ap = _lib.ctypes.cast(Apd, _lib.ctypes.c_void_p).value
        ptr_ai = _lib.ctypes.cast(Aid, _lib.ctypes.c_void_p).value
        # copy the device arrays back through torch and compare with the host copies
        ap_d = torch.empty(len(Xr) + 1, dtype=torch.int64, device=dev)
        ai_d = torch.empty(nz.value, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(ctypes.c_void_p(ap_d.data_ptr()), ctypes.c_void_p(ptr_ap), ap_d.numel() * 8, 3) == 0
        assert hip.hipMemcpy(ctypes.c_void_p(ai_d.data_ptr()), ctypes.c_void_p(ptr_ai), ai_d.numel() * 8, 3) == 0
        host_ap = np.ctypeslib.as_array(Ap, shape=(len(Xr) + 1,))
        host_ai = np.ctypeslib.as_array(Ai, shape=(nz.value,))
        assert np.array_equal(ap_d.cpu().numpy(), host_ap)
        assert np.array_equal(ai_d.cpu().numpy(), host_ai)
    finally:
        lib.mfs_pattern_destroy(h)
    bad = tets.copy()
    bad[5, 2] = len(Xr)
    with pytest.raises(ValueError):
        mfs.SystemAssembler(len(Xr)).blockSparsityPattern(torch.as_tensor(bad).to(dev))


def test_64cube_pattern_matches_host(dev):
    tets, Xr = wl.cube_mesh(64)
    asm = mfs.SystemAssembler(len(Xr))
    t0 = time.perf_counter()
    Hh = asm.blockSparsityPattern(tets)
    t1 = time.perf_counter()
    en = torch.as_tensor(tets).to(dev)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    Hd = asm.blockSparsityPattern(en)
    t3 = time.perf_counter()
    check_equal(Hh, Hd)
    assert Hd.nz == 2146689
    print(f"64^3 pattern: host {t1 - t0:.3f}s, device {t3 - t2:.3f}s (incl. copy-back)")
