"""FP64 CPU reference of the learned block-Jacobi preconditioner network (test infrastructure only).

Restates the definition the HIP kernels implement (finite_elements_amd/csrc/fe_gnn.hip): scale-free
node features of the assembled stiffness, 3 mean-aggregation message-passing layers (width 32, ReLU),
a linear head, s = exp(clamp(o, -4, 4)), learned blocks diag(s) K_ii^-1 diag(s), and the loss
L = 1 - (r.z)^2 / ((z.Kz)(r.e)) with z = M r.  torch autograd in float64 on the CPU provides the
reference gradients; the parameter vector layout matches the kernels' (per layer W1 [32][fin],
W2 [32][fin], b [32]; then the head Wo [d][32], bo [d]).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

WIDTH = 32
NF = 8


def unpack(params: torch.Tensor, d: int):
    out, off = [], 0
    for layer in range(3):
        fin = NF if layer == 0 else WIDTH
        W1 = params[off:off + WIDTH * fin].reshape(WIDTH, fin)
        off += WIDTH * fin
        W2 = params[off:off + WIDTH * fin].reshape(WIDTH, fin)
        off += WIDTH * fin
        b = params[off:off + WIDTH]
        off += WIDTH
        out.append((W1, W2, b))
    Wo = params[off:off + d * WIDTH].reshape(d, WIDTH)
    off += d * WIDTH
    bo = params[off:off + d]
    off += d
    assert off == params.numel()
    return out, (Wo, bo)


def features(bptr, bcol, diag_blocks, fixed, d):
    """(N, 8) features from the diagonal blocks (N, d, d) and the fixed-DOF mask (N*d,)."""
    N = len(bptr) - 1
    tr = np.abs(np.einsum('nii->n', diag_blocks))
    safe = np.where(tr > 0, tr, 1.0)
    logtr = np.log(safe)
    f = np.zeros((N, NF))
    for k in range(min(d, 3)):
        f[:, 1 + k] = np.abs(diag_blocks[:, k, k]) / safe
    off = diag_blocks.copy()
    idx = np.arange(d)
    off[:, idx, idx] = 0.0
    f[:, 4] = np.sqrt((off ** 2).sum(axis=(1, 2))) / safe
    f[:, 5] = fixed.reshape(N, d).sum(axis=1) / d
    deg = np.diff(bptr)
    f[:, 6] = deg / 16.0
    for i in range(N):
        nb = [j for j in bcol[bptr[i]:bptr[i + 1]] if j != i]
        if nb:
            v = logtr[nb]
            f[i, 0] = logtr[i] - v.mean()
            f[i, 7] = v.std()
    return f


def neighbour_mean(bptr, bcol):
    N = len(bptr) - 1
    rows, cols, vals = [], [], []
    for i in range(N):
        nb = [j for j in bcol[bptr[i]:bptr[i + 1]] if j != i]
        for j in nb:
            rows.append(i)
            cols.append(j)
            vals.append(1.0 / len(nb))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N)).tocoo()
    return torch.sparse_coo_tensor(np.vstack([A.row, A.col]), A.data, (N, N), dtype=torch.float64)


def forward(params, feat, avg, d):
    layers, (Wo, bo) = unpack(params, d)
    h = torch.as_tensor(feat, dtype=torch.float64)
    for W1, W2, b in layers:
        agg = torch.sparse.mm(avg, h)
        h = torch.relu(h @ W1.T + agg @ W2.T + b)
    return h @ Wo.T + bo


def loss(params, feat, avg, binv, A, pairs, d, zc=None, kind=0):
    """Mean loss over the (r, e) pairs; binv (N, d, d); A scipy (n, n).  zc: optional per-pair vectors
    added to z = M r (a network-independent part of the preconditioner, e.g. a coarse correction).
    kind 0: 1 - (r.z)^2/((z.Az)(r.e)), the energy left after an optimally damped step; kind 1:
    |e - z|_A^2 / |e|_A^2, the energy left after the undamped step, evaluated as written."""
    o = forward(params, feat, avg, d)
    s = torch.exp(torch.clamp(o, -4.0, 4.0))  # (N, d)
    B = torch.as_tensor(binv, dtype=torch.float64)
    Ac = A.tocoo()
    At = torch.sparse_coo_tensor(np.vstack([Ac.row, Ac.col]), Ac.data, A.shape, dtype=torch.float64)
    total = 0.0
    for k, (r, e) in enumerate(pairs):
        r = torch.as_tensor(r, dtype=torch.float64)
        e = torch.as_tensor(e, dtype=torch.float64)
        rn = r.reshape(-1, d)
        z = (s * torch.einsum('nij,nj->ni', B, s * rn)).reshape(-1)
        if zc is not None:
            z = z + torch.as_tensor(zc[k], dtype=torch.float64)
        az = torch.sparse.mm(At, z[:, None])[:, 0]
        if kind == 0:
            rz = torch.dot(r, z)
            total = total + 1.0 - rz * rz / (torch.dot(z, az) * torch.dot(r, e))
        else:
            ez = e - z
            aez = torch.sparse.mm(At, ez[:, None])[:, 0]
            total = total + torch.dot(ez, aez) / torch.dot(e, torch.sparse.mm(At, e[:, None])[:, 0])
    return total / len(pairs), o
