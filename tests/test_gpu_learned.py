"""Learned block-Jacobi preconditioner: the HIP forward / loss / backward kernels against the FP64
torch-CPU oracle (oracle/gnn.py), and the effect of training on PCG iteration counts against plain
block Jacobi (the ablation)."""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from finite_elements_amd import _hip, device, learned, meshgen
from finite_elements_amd.system import FESystem
from oracle import gnn as ognn

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _system(n=4, seed=1000):
    pr = meshgen.cube_elasticity(n=n, seed=seed)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    K = s.stiffness_bsr()
    fixed, g = s.dirichlet(pr.fixed_dofs)
    KD, b = s.prepare(K, fixed, g, s.ctx.tensor(pr.f))
    return pr, s, KD, b, fixed


def _host_blocks(KD):
    """(nnzb, d, d) FP64 blocks and the scipy matrix of a plane-layout device matrix."""
    d = KD.dof
    vals = KD.convert(_hip.FE_LAYOUT_BSR).vals.cpu().numpy().reshape(-1, d, d)
    g = KD.graph.pattern
    rows = np.repeat(np.arange(g.n_nodes), np.diff(g.bptr))
    A = sp.bsr_matrix((vals, g.bcol, g.bptr), shape=(g.n_nodes * d, g.n_nodes * d)).tocsr()
    diag = np.zeros((g.n_nodes, d, d))
    sel = g.bcol == rows
    diag[rows[sel]] = vals[sel]
    return vals, A, diag


def test_features_and_forward_match_oracle(ctx):
    _say('gnn features/forward')
    pr, s, KD, b, fixed = _system(4)
    rng = np.random.default_rng(5)
    model = learned.LearnedPreconditioner(3, seed=3)
    p = model.params_host.copy()
    p[-(3 * 32 + 3):] = rng.normal(0, 0.3, 3 * 32 + 3)  # non-trivial head
    model = learned.LearnedPreconditioner(3, params=p)
    binv = device.precond_setup(KD, 'block_jacobi')
    M, o, feat, acts, aggs = model.forward(KD, fixed, binv)
    _, A, diag = _host_blocks(KD)
    g = KD.graph.pattern
    f_ref = ognn.features(g.bptr, g.bcol, diag, fixed.cpu().numpy().astype(bool), 3)
    np.testing.assert_allclose(feat.cpu().numpy().reshape(-1, 8), f_ref, rtol=2e-5, atol=2e-6)
    avg = ognn.neighbour_mean(g.bptr, g.bcol)
    o_ref = ognn.forward(torch.as_tensor(p, dtype=torch.float64), f_ref, avg, 3).numpy()
    np.testing.assert_allclose(o.cpu().numpy().reshape(-1, 3), o_ref, rtol=1e-4, atol=1e-4)
    # learned blocks: diag(s) Binv diag(s), SPD
    Mh = M.cpu().numpy().reshape(-1, 3, 3)
    sv = np.exp(np.clip(o_ref, -4, 4))
    Bh = binv.cpu().numpy().reshape(-1, 3, 3)
    np.testing.assert_allclose(Mh, sv[:, :, None] * Bh * sv[:, None, :], rtol=1e-3)
    assert np.all(np.linalg.eigvalsh(0.5 * (Mh + Mh.transpose(0, 2, 1))) > 0)


@pytest.mark.parametrize('two_level,kind', [(False, 0), (True, 0), (True, 1)])
def test_gradients_match_oracle(ctx, two_level, kind):
    _say(f'gnn gradients (two_level={two_level}, loss kind {kind})')
    pr, s, KD, b, fixed = _system(4)
    rng = np.random.default_rng(6)
    p = learned.init_params(3, seed=4)
    p[-(3 * 32 + 3):] = rng.normal(0, 0.3, 3 * 32 + 3)
    model = learned.LearnedPreconditioner(3, params=p)
    model.loss_kind = kind
    coarse = s.coarse_space(pr.fixed_dofs, target=8) if two_level else None
    solver = s.solver(KD, 'two_level' if two_level else 'block_jacobi', 'fp64', coarse=coarse)
    xstar, _ = solver.solve(b, tol=1e-12)
    pairs, zcs = [], []
    for k in (2, 5):
        xk, _ = solver.solve(b, tol=1e-12, max_iter=k)
        rk = b - device.spmv(solver.A64, xk)
        pairs.append((rk, xstar - xk))
        if two_level:
            zcs.append(coarse.apply(solver.Ainv64, rk, torch.zeros_like(rk)))
    binv = device.precond_setup(KD, 'block_jacobi')
    ts = learned.TrainingSystem(solver.A64, fixed, pairs, binv, zcs=zcs if two_level else None)
    grad = torch.zeros(p.size, dtype=torch.float32, device=ctx.torch_device)
    loss_gpu = model.loss_and_grad(ts, grad)
    _, A, diag = _host_blocks(KD)
    g = KD.graph.pattern
    f_ref = ognn.features(g.bptr, g.bcol, diag, fixed.cpu().numpy().astype(bool), 3)
    avg = ognn.neighbour_mean(g.bptr, g.bcol)
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    host_pairs = [(r.cpu().numpy(), e.cpu().numpy()) for r, e in pairs]
    host_zc = [z.cpu().numpy() for z in zcs] if two_level else None
    L, _ = ognn.loss(P, f_ref, avg, binv.cpu().numpy().reshape(-1, 3, 3), A, host_pairs, 3, zc=host_zc,
                     kind=kind)
    L.backward()
    g_ref = P.grad.numpy()
    g_gpu = grad.cpu().numpy()
    rel = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    _say(f'  loss gpu {loss_gpu:.6f} ref {L.item():.6f}; gradient rel err {rel:.2e}')
    assert abs(loss_gpu - L.item()) < 1e-5 * max(1.0, abs(L.item()))
    assert rel < 2e-3
    # every parameter block receives gradient
    assert np.abs(g_gpu[:512]).max() > 0 and np.abs(g_gpu[-99:]).max() > 0


def test_zero_head_is_block_jacobi(ctx):
    _say('gnn zero head')
    pr, s, KD, b, fixed = _system(5)
    model = learned.LearnedPreconditioner(3, seed=0)
    M = model.preconditioner(s, KD, fixed)
    binv = device.precond_setup(KD, 'block_jacobi')
    assert torch.equal(M, binv)


def test_training_reduces_loss_and_reports_gain(ctx):
    """Train on residual/error pairs of plain block-Jacobi PCG on 8^3 cubes; report iterations of
    learned vs plain block-Jacobi PCG on held-out cubes (the ablation)."""
    _say('gnn training ablation')
    systems = learned.make_training_systems(n=8, seeds=range(1000, 1012), snapshots=(5, 20, 60))
    model = learned.LearnedPreconditioner(3, seed=0)
    hist = model.train(systems, epochs=15, lr=3e-3, log=_say)
    assert hist[-1] < hist[0]
    gains = []
    for seed in (2000, 2001, 2002):
        pr, s, KD, b, fixed = _system(10, seed)
        base, ib = s.solver(KD, 'block_jacobi', 'fp64').solve(b, tol=1e-8)
        M = model.preconditioner(s, KD, fixed)
        xl, il = s.solver(KD, 'block_jacobi', 'fp64', M64=M).solve(b, tol=1e-8)
        assert il['status'] == 0 and il['relres'] <= 1e-8
        gains.append(ib['iterations'] / il['iterations'])
        _say(f'  seed {seed}: block-Jacobi {ib["iterations"]} it, learned {il["iterations"]} it')
    _say(f'  iteration ratio plain/learned: {np.round(gains, 3)}')
    # the learned preconditioner must stay a valid SPD preconditioner that converges.  It does not
    # gain on these fields (measured ratios 0.88-0.90: a per-node rescaling of block Jacobi has no
    # leverage here), which is why the bench's headline is the method without it; the bound keeps
    # the loss from growing unnoticed
    assert min(gains) >= 0.8


def test_two_level_training_ablation(ctx):
    """The network as the smoother of the two-level preconditioner: train on residual/error pairs of
    plain two-level PCG on 8^3 cubes; report iterations with and without the network on held-out
    cubes (the ablation).  The learned preconditioner must stay valid (converge to the tolerance)."""
    _say('gnn two-level ablation')
    systems = learned.make_training_systems(n=8, seeds=range(1000, 1012), snapshots=(2, 6, 15),
                                            two_level=True, coarse_target=27)
    model = learned.LearnedPreconditioner(3, seed=0)
    hist = model.train(systems, epochs=15, lr=3e-3, log=_say)
    assert hist[-1] < hist[0]
    gains = []
    for seed in (2000, 2001, 2002):
        pr, s, KD, b, fixed = _system(10, seed)
        coarse = s.coarse_space(pr.fixed_dofs, target=27)
        _, ib = s.solver(KD, 'two_level', 'fp64', coarse=coarse).solve(b, tol=1e-8)
        M = model.preconditioner(s, KD, fixed)
        _, il = s.solver(KD, 'two_level', 'fp64', M64=M, coarse=coarse).solve(b, tol=1e-8)
        assert il['status'] == 0 and il['relres'] <= 1e-8
        gains.append(ib['iterations'] / il['iterations'])
        _say(f'  seed {seed}: two-level {ib["iterations"]} it, learned two-level {il["iterations"]} it')
    _say(f'  iteration ratio without/with network: {np.round(gains, 3)}')
    # measured 0.89-0.98: the network costs iterations as the two-level smoother too (see above)
    assert min(gains) >= 0.8
