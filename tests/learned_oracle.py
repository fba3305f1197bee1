"""CPU oracle of the learned preconditioner (test infrastructure): the network, factor and probe
loss restated in float64 torch on the CPU, so autograd gives the reference gradients for the
device's hand-written backward kernels.  Conventions as the device: weights [in, out] row-major,
rounded to bfloat16 before use; 3x3 blocks column-major (B[3c + r] = B(r, c))."""
import numpy as np
import torch

NAMES = (["W0", "b0"] + [f"{k}{l}" for l in range(4) for k in ("Wm", "bm", "Wu", "bu")] +
         ["Wd1", "bd1", "Wd2", "bd2", "Wdd", "bdd"])
SHAPES = {"W0": (6, 64), "b0": (64,), "Wd1": (128, 64), "bd1": (64,), "Wd2": (64, 9), "bd2": (9,),
          "Wdd": (64, 6), "bdd": (6,)}
for _l in range(4):
    SHAPES.update({f"Wm{_l}": (73, 64), f"bm{_l}": (64,), f"Wu{_l}": (128, 64), f"bu{_l}": (64,)})


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def blocks(x9):
    """(m, 9) column-major -> (m, 3, 3) [row, col] float64 tensor."""
    return torch.as_tensor(np.asarray(x9, np.float64)).reshape(-1, 3, 3).transpose(1, 2)


def loss_and_grads(flat, offsets, g, feats, z, round_weights=True):
    """flat: master parameters; g: graph arrays; feats: features(); z: (np, 3n).  The device uses
    the bfloat16-rounded weights; round_weights=False evaluates the given values exactly (for
    finite-difference checks of the oracle itself)."""
    P = {}
    for i, k in enumerate(NAMES):
        w = flat[offsets[i]:offsets[i + 1]]
        w = (bf16(w) if round_weights else np.asarray(w)).astype(np.float64).reshape(SHAPES[k])
        P[k] = torch.tensor(w, requires_grad=True)
    n = feats["node"].shape[0]
    X0 = torch.as_tensor(feats["node"].astype(np.float64))
    EF = torch.as_tensor(feats["edge"].astype(np.float64))
    EFm = blocks(feats["edge"])
    AD = blocks(feats["diag"])
    dst = torch.as_tensor(g["edge_dst"].astype(np.int64))
    src = torch.as_tensor(g["edge_src"].astype(np.int64))
    le = torch.as_tensor(g["lower_edge"].astype(np.int64))
    ld = torch.as_tensor(g["lower_dst"].astype(np.int64))
    ls = torch.as_tensor(g["lower_src"].astype(np.int64))
    relu = torch.relu
    h = relu(X0 @ P["W0"] + P["b0"])
    for l in range(4):
        m = relu(torch.cat([h[src], EF], 1) @ P[f"Wm{l}"] + P[f"bm{l}"])
        M = torch.zeros(n, 64, dtype=torch.float64).index_add(0, dst, m)
        h = h + relu(torch.cat([h, M], 1) @ P[f"Wu{l}"] + P[f"bu{l}"])
    Dl = relu(torch.cat([h[ld], h[ls]], 1) @ P["Wd1"] + P["bd1"]) @ P["Wd2"] + P["bd2"]
    Gm = h @ P["Wdd"] + P["bdd"]
    LE = EFm[le] + Dl.reshape(-1, 3, 3).transpose(1, 2)
    zero = torch.zeros(n, dtype=torch.float64)
    LD = torch.stack([torch.stack([torch.exp(Gm[:, 0]), zero, zero], 1),
                      torch.stack([AD[:, 1, 0] + Gm[:, 3], torch.exp(Gm[:, 1]), zero], 1),
                      torch.stack([AD[:, 2, 0] + Gm[:, 4], AD[:, 2, 1] + Gm[:, 5], torch.exp(Gm[:, 2])], 1)], 1)
    Z = torch.as_tensor(np.asarray(z, np.float64)).reshape(z.shape[0], n, 3)
    loss = 0
    for p in range(Z.shape[0]):
        zp = Z[p]
        y = torch.einsum("nrc,nr->nc", LD, zp).index_add(0, ls, torch.einsum("qrc,qr->qc", LE, zp[ld]))
        w = torch.einsum("nrc,nc->nr", LD, y).index_add(0, ld, torch.einsum("qrc,qc->qr", LE, y[ls]))
        a = torch.einsum("nrc,nc->nr", AD, zp).index_add(0, dst, torch.einsum("erc,ec->er", EFm, zp[src]))
        loss = loss + ((w - a) ** 2).sum() / (a ** 2).sum()
    loss = loss / Z.shape[0]
    loss.backward()
    grads = np.concatenate([P[k].grad.numpy().ravel() for k in NAMES])
    return float(loss), grads
