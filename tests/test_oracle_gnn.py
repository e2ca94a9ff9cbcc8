"""Pins of the learned-preconditioner oracle (oracle/gnn.py) by what the mathematics fixes:

* the loss 1 - (r.z)^2 / ((z.Az)(r.e)) is the energy fraction left by an optimally damped step, so it
  lies in [0, 1], vanishes when M is the exact inverse (block-diagonal A, zero network output), and is
  unchanged when the scaling s is multiplied by a constant;
* the node features are scale-free (K -> cK leaves them unchanged) and see only the node graph;
* autograd gradients agree with central differences.
"""

import numpy as np
import scipy.sparse as sp
import torch

from oracle import gnn


def _block_diag_system(N=12, d=3, seed=0):
    rng = np.random.default_rng(seed)
    blocks = []
    for _ in range(N):
        a = rng.standard_normal((d, d))
        blocks.append(a @ a.T + d * np.eye(d))
    A = sp.block_diag(blocks, format='csr')
    # ring graph over the nodes (the network's neighbours); A itself stays block diagonal
    bptr = [0]
    bcol = []
    for i in range(N):
        nb = sorted({(i - 1) % N, i, (i + 1) % N})
        bcol.extend(nb)
        bptr.append(len(bcol))
    return np.array(blocks), A, np.array(bptr), np.array(bcol)


def _params(d, head_scale=0.0, seed=1):
    rng = np.random.default_rng(seed)
    n = 2 * (32 * 8) + 32 + 2 * (2 * 32 * 32 + 32) + d * 32 + d
    p = rng.normal(0, 0.2, n)
    p[-(d * 32 + d):] *= head_scale
    return torch.tensor(p, dtype=torch.float64, requires_grad=True)


def _pairs(A, k, seed=2):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        e = rng.standard_normal(A.shape[0])
        out.append((A @ e, e))
    return out


def test_exact_inverse_gives_zero_loss():
    blocks, A, bptr, bcol = _block_diag_system()
    binv = np.linalg.inv(blocks)
    feat = gnn.features(bptr, bcol, blocks, np.zeros(A.shape[0], bool), 3)
    avg = gnn.neighbour_mean(bptr, bcol)
    L, _ = gnn.loss(_params(3, 0.0), feat, avg, binv, A, _pairs(A, 3), 3)
    assert abs(L.item()) < 1e-12


def test_loss_bounds_and_damping_invariance():
    blocks, A, bptr, bcol = _block_diag_system(seed=3)
    rng = np.random.default_rng(4)
    binv = np.linalg.inv(blocks) + 0.05 * rng.standard_normal(blocks.shape)
    binv = 0.5 * (binv + binv.transpose(0, 2, 1))
    feat = gnn.features(bptr, bcol, blocks, np.zeros(A.shape[0], bool), 3)
    avg = gnn.neighbour_mean(bptr, bcol)
    p = _params(3, 1.0, seed=5)
    L, o = gnn.loss(p, feat, avg, binv, A, _pairs(A, 4), 3)
    assert 0.0 < L.item() < 1.0
    # adding a constant to every output scales s by a constant: the optimal damping absorbs it
    d = 3
    q = p.detach().clone()
    q[-d:] += 0.7
    L2, _ = gnn.loss(q, feat, avg, binv, A, _pairs(A, 4), 3)
    assert abs(L2.item() - L.item()) < 1e-12


def test_undamped_loss_pins():
    """kind 1 is |e - z|_A^2/|e|_A^2: zero for the exact inverse, one for z = 0 (all outputs at the
    lower clamp give z -> e^-8 M_bj r, nearly zero), equal to 1 - 2(r.z)/(r.e) + (z.Az)/(r.e) (the
    form the kernel uses), and not invariant to a uniform rescaling of M (unlike kind 0)."""
    blocks, A, bptr, bcol = _block_diag_system(seed=11)
    feat = gnn.features(bptr, bcol, blocks, np.zeros(A.shape[0], bool), 3)
    avg = gnn.neighbour_mean(bptr, bcol)
    L0, _ = gnn.loss(_params(3, 0.0), feat, avg, np.linalg.inv(blocks), A, _pairs(A, 3), 3, kind=1)
    assert abs(L0.item()) < 1e-12
    rng = np.random.default_rng(12)
    binv = np.linalg.inv(blocks) + 0.05 * rng.standard_normal(blocks.shape)
    binv = 0.5 * (binv + binv.transpose(0, 2, 1))
    p = _params(3, 1.0, seed=13)
    pairs = _pairs(A, 3)
    L1, o = gnn.loss(p, feat, avg, binv, A, pairs, 3, kind=1)
    s = np.exp(np.clip(o.detach().numpy(), -4, 4))
    total = 0.0
    for r, e in pairs:
        z = (s * np.einsum('nij,nj->ni', binv, s * r.reshape(-1, 3))).reshape(-1)
        total += 1 - 2 * (r @ z) / (r @ e) + (z @ (A @ z)) / (r @ e)
    assert abs(L1.item() - total / len(pairs)) < 1e-12
    q = p.detach().clone()
    q[-3:] += 0.7
    L2, _ = gnn.loss(q, feat, avg, binv, A, pairs, 3, kind=1)
    assert abs(L2.item() - L1.item()) > 1e-3
    qlow = p.detach().clone()
    qlow[-3:] -= 1e3
    Llow, _ = gnn.loss(qlow, feat, avg, binv, A, pairs, 3, kind=1)
    assert abs(Llow.item() - 1.0) < 1e-3


def test_features_scale_free():
    blocks, A, bptr, bcol = _block_diag_system(seed=6)
    fixed = np.zeros(A.shape[0], bool)
    fixed[:3] = True
    f1 = gnn.features(bptr, bcol, blocks, fixed, 3)
    f2 = gnn.features(bptr, bcol, 1e7 * blocks, fixed, 3)
    np.testing.assert_allclose(f1, f2, rtol=1e-12, atol=1e-12)
    assert np.all(f1[:1, 5] == 1.0)  # fully fixed node


def test_autograd_matches_central_differences():
    blocks, A, bptr, bcol = _block_diag_system(N=6, seed=7)
    rng = np.random.default_rng(8)
    binv = np.linalg.inv(blocks) * (1 + 0.1 * rng.standard_normal(blocks.shape))
    binv = 0.5 * (binv + binv.transpose(0, 2, 1))
    feat = gnn.features(bptr, bcol, blocks, np.zeros(A.shape[0], bool), 3)
    avg = gnn.neighbour_mean(bptr, bcol)
    pairs = _pairs(A, 2)
    p = _params(3, 1.0, seed=9)
    L, _ = gnn.loss(p, feat, avg, binv, A, pairs, 3)
    L.backward()
    g = p.grad.numpy()
    idx = np.random.default_rng(10).choice(p.numel(), 12, replace=False)
    for i in idx:
        h = 1e-6
        q1 = p.detach().clone()
        q1[i] += h
        q2 = p.detach().clone()
        q2[i] -= h
        fd = (gnn.loss(q1, feat, avg, binv, A, pairs, 3)[0].item() -
              gnn.loss(q2, feat, avg, binv, A, pairs, 3)[0].item()) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9
