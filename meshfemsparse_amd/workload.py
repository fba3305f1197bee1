"""Synthetic Newton-step workloads: the unit-cube Kuhn-tet mesh with a seeded perturbation and
compressible neo-Hookean element Hessians / gradient (computed by the library's host code)."""
import numpy as np

from . import _lib
from ._lib import ptr


def cube_mesh(N):
    """(tets[6N^3, 4] int32, X_rest[(N+1)^3, 3] float64); node id = x + (N+1)(y + (N+1)z)."""
    tets = np.empty((6 * N ** 3, 4), np.int32)
    X = np.empty(((N + 1) ** 3, 3), np.float64)
    _lib.call("mfs_cube_mesh", N, ptr(tets, _lib.c_i32), ptr(X, _lib.c_double))
    return tets, X


def perturb(N, X_rest, seed, sigma_rel=1e-3):
    X = np.array(X_rest, dtype=np.float64, copy=True)
    _lib.call("mfs_perturb_nodes", N, int(seed) & 0xFFFFFFFFFFFFFFFF, float(sigma_rel), ptr(X, _lib.c_double))
    return X


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def neohookean(X_rest, X, tets, E=1e5, nu=0.3, E_per_elem=None, nu_per_elem=None, hessians=True,
               gradient=True, energy=False, threads=0):
    """Per-element 12x12 Hessians (ne x 144, column-major), assembled gradient, total energy."""
    ne = len(tets)
    Xr = np.ascontiguousarray(X_rest, np.float64)
    Xc = np.ascontiguousarray(X, np.float64)
    tt = np.ascontiguousarray(tets, np.int32)
    He = np.empty((ne, 144), np.float64) if hessians else None
    g = np.zeros(Xc.size, np.float64) if gradient else None
    en = np.zeros(1) if energy else None
    Ee = None if E_per_elem is None else np.ascontiguousarray(E_per_elem, np.float64)
    nue = None if nu_per_elem is None else np.ascontiguousarray(nu_per_elem, np.float64)
    P = lambda a: None if a is None else ptr(a, _lib.c_double)
    _lib.call("mfs_neohookean_host", ptr(Xr, _lib.c_double), ptr(Xc, _lib.c_double), ptr(tt, _lib.c_i32), ne,
              P(Ee), P(nue), float(E), float(nu), P(He), P(g), P(en), int(threads))
    return He, g, (en[0] if energy else None)


def z0_pins(N):
    """Scalar variables of all z = 0 nodes (whole 3-blocks)."""
    M = N + 1
    nodes = np.arange(M * M, dtype=np.int64)
    return (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()


def batch_meshes(nmeshes=512, nmin=6, nmax=20, seed=0):
    """Independent unit cubes with per-side hex counts n ~ U{nmin..nmax}, numbered one after
    another as one disconnected mesh (its Hessian is block diagonal, one block per mesh).
    Returns dict(n, tets, X_rest, node_off[m+1], pins, seg[m+1] (scalar offsets per mesh))."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, nmeshes)
    tets, Xs, pins, node_off = [], [], [], [0]
    for n in ns:
        t, X = cube_mesh(int(n))
        tets.append(t + node_off[-1])
        Xs.append(X)
        pins.append(z0_pins(int(n)) + 3 * node_off[-1])
        node_off.append(node_off[-1] + len(X))
    node_off = np.array(node_off, np.int64)
    return dict(n=ns, tets=np.ascontiguousarray(np.concatenate(tets), np.int32), X_rest=np.concatenate(Xs),
                node_off=node_off, pins=np.concatenate(pins), seg=3 * node_off)


def batch_perturb(batch, seed, sigma_rel=1e-3):
    """Each mesh perturbed with its own h = 1/n and a seed derived from (seed, mesh index)."""
    Xr = batch["X_rest"]
    X = np.empty_like(Xr)
    no = batch["node_off"]
    for m, n in enumerate(batch["n"]):
        X[no[m]:no[m + 1]] = perturb(int(n), Xr[no[m]:no[m + 1]], seed * 1000003 + m, sigma_rel)
    return X


def grf_log_E(X_rest, tets, seed, mean=np.log(1e5), std=1.5, corr_len=0.1, features=256):
    """Per-element Young's modulus E = exp(g) with g a Gaussian random field at the element
    centroids: mean `mean`, standard deviation `std`, squared-exponential correlation of length
    `corr_len`, sampled by random Fourier features (omega ~ N(0, I / corr_len^2))."""
    rng = np.random.default_rng(seed)
    om = rng.standard_normal((features, 3)) / corr_len
    ph = rng.uniform(0.0, 2.0 * np.pi, features)
    if not isinstance(X_rest, np.ndarray):
        # torch tensors (e.g. on the GPU, for bulk training-data synthesis): same field
        import torch
        Xt = X_rest.to(torch.float64)
        c = Xt[tets.long()].mean(dim=1)
        omt = torch.as_tensor(om, dtype=torch.float64, device=Xt.device)
        pht = torch.as_tensor(ph, dtype=torch.float64, device=Xt.device)
        g = np.sqrt(2.0 / features) * torch.cos(c @ omt.T + pht).sum(dim=1)
        return torch.exp(mean + std * g)
    c = X_rest[tets].mean(axis=1)                         # centroids (ne, 3)
    g = np.sqrt(2.0 / features) * np.cos(c @ om.T + ph).sum(axis=1)
    return np.exp(mean + std * g)
