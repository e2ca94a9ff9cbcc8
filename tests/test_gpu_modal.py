"""GPU modal path: the batched Jacobi eigensolver against numpy, and batched LOBPCG against the
eigenvalue oracle (oracle/eigen.py: the reference's eigsh calls) on the oracle's own assembled K, M."""

import os

import numpy as np
import pytest
import torch

from finite_elements_amd import _hip, device, meshgen, modal
from finite_elements_amd.system import FESystem
from oracle import eigen, vectorized as vz

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

EIG_RTOL = 1e-8


def _say(m):
    print(m, flush=True)


@pytest.mark.parametrize('r', [1, 5, 24, 47, 72, 96])
def test_sym_eig_matches_numpy(ctx, r):
    _say(f'sym_eig r={r}')
    rng = np.random.default_rng(r)
    A = rng.standard_normal((3, r, r))
    A = A + A.transpose(0, 2, 1)
    A[1] = np.diag(np.arange(r, dtype=float))  # already diagonal
    if r > 2:
        A[2, :2, :2] = [[1.0, 1e-9], [1e-9, 1.0]]  # nearly equal eigenvalues
    w, V = modal.sym_eig(torch.as_tensor(A, device=ctx.torch_device))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    for k in range(3):
        ref = np.linalg.eigvalsh(A[k])
        scale = max(1.0, np.abs(ref).max())
        assert np.max(np.abs(w[k] - ref)) <= 1e-13 * scale * r
        assert np.max(np.abs(A[k] @ V[k] - V[k] * w[k][None, :])) <= 1e-12 * scale * r
        assert np.max(np.abs(V[k].T @ V[k] - np.eye(r))) <= 1e-12 * r


def _modal_system(n, seed):
    pr = meshgen.cube_modal(n=n, seed=seed)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    return pr, s


def _oracle_pencil(pr):
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, what='stiffness')
    M = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, what='mass')
    return K, M


def _check_pairs(vals, vecs, K, M, fixed_dofs, ref_vals, ref_vecs=None):
    np.testing.assert_allclose(vals, ref_vals, rtol=EIG_RTOL)
    X = vecs
    if fixed_dofs is not None and len(fixed_dofs):
        assert np.all(X[fixed_dofs] == 0.0)
    G = X.T @ (M @ X)
    np.testing.assert_allclose(G, np.eye(X.shape[1]), atol=1e-8)
    R = K @ X - (M @ X) * vals[None, :]
    if fixed_dofs is not None and len(fixed_dofs):
        R[fixed_dofs] = 0.0  # reactions of the clamped DOFs are not part of the reduced problem
    rel = np.linalg.norm(R, axis=0) / (np.abs(vals) * np.linalg.norm(M @ X, axis=0))
    assert rel.max() < 1e-5, rel


# blocks 4 .. 32: the recombination kernel is built per block size, and the Gram kernel's 16-column
# tiles are partly empty except at 3 * block = 96
@pytest.mark.parametrize('n,k,block', [(4, 3, 4), (5, 6, 8), (7, 20, 24), (6, 24, 32)])
def test_lobpcg_smallest_clamped_vs_eigsh(ctx, n, k, block):
    _say(f'lobpcg smallest clamped n={n} k={k}')
    pr, s = _modal_system(n, 5000)
    K = s.stiffness_bsr()
    Mm = s.mass_bsr()
    fixed, _ = s.dirichlet(pr.fixed_dofs)
    KD = device.apply_dirichlet(K, fixed)
    MD = device.apply_dirichlet(Mm, fixed)
    vals, vecs, info = modal.lobpcg(KD, MD, k, block=block, fixed=fixed, tol=1e-7, max_iter=3000)
    _say(f'  iterations {info["iterations"]}, status {info["status"]}, max resid {info["resid"].max():.2e}')
    assert info['status'] == 0
    Ko, Mo = _oracle_pencil(pr)
    ref_vals, ref_vecs = eigen.smallest_clamped(Ko, Mo, pr.fixed_dofs, k)
    _check_pairs(vals[0], vecs[0].cpu().numpy(), Ko, Mo, pr.fixed_dofs, ref_vals, ref_vecs)


def test_lobpcg_two_level_vs_eigsh(ctx):
    """The two-level preconditioner (block Jacobi + rigid-body coarse correction, one coarse inverse
    per variant) gives the same pairs as eigsh in fewer iterations than block Jacobi alone."""
    _say('lobpcg two-level')
    seeds = (5000, 5001)
    pr0, s = _modal_system(8, seeds[0])
    fixed, _ = s.dirichlet(pr0.fixed_dofs)
    MD = device.apply_dirichlet(s.mass_bsr(), fixed)
    KDs, prs = [], []
    for sd in seeds:
        pr = meshgen.cube_modal(n=8, seed=sd)
        s.set_materials(pr.materials)
        KDs.append(device.apply_dirichlet(s.stiffness_bsr(), fixed))
        prs.append(pr)
    coarse = s.coarse_space(pr0.fixed_dofs, target=27)
    v2, x2, i2 = modal.lobpcg(KDs, MD, 12, block=16, fixed=fixed, tol=1e-7, max_iter=3000, coarse=coarse)
    v1, _, i1 = modal.lobpcg(KDs, MD, 12, block=16, fixed=fixed, tol=1e-7, max_iter=3000)
    _say(f'  iterations two-level {i2["iterations"]}, block Jacobi {i1["iterations"]}')
    assert i2['status'] == 0 and i1['status'] == 0
    # the variants split over two concurrent calls on two streams: the same bits
    v2s, x2s, i2s = modal.lobpcg(KDs, MD, 12, block=16, fixed=fixed, tol=1e-7, max_iter=3000,
                                 coarse=coarse, streams=2)
    assert np.array_equal(v2s, v2) and torch.equal(x2s, x2)
    assert np.array_equal(i2s['iterations'], i2['iterations']) and i2s['status'] == 0
    assert np.all(i2['iterations'] < i1['iterations'])
    for v, pr in enumerate(prs):
        Ko, Mo = _oracle_pencil(pr)
        ref_vals, _ = eigen.smallest_clamped(Ko, Mo, pr.fixed_dofs, 12)
        _check_pairs(v2[v], x2[v].cpu().numpy(), Ko, Mo, pr.fixed_dofs, ref_vals)


@pytest.mark.timeout(400)
def test_modal48_against_golden_eigsh(ctx):
    """The full 64-variant batch of the 48^3 modal workload (seeds 5000-5063, 352,947 DOF each) in
    batched LOBPCG over two streams, as the bench runs it: every variant must reach the residual tolerance, and the
    variants whose FP64 eigsh values the oracle wrote (tests/golden/make_modal_golden.py) must match
    them within 1e-8 relative in all 20 lowest eigenvalues."""
    path = os.path.join(GOLDEN, 'modal48_eigs.npz')
    if not os.path.exists(path):
        pytest.skip('tests/golden/modal48_eigs.npz not generated')
    g = np.load(path, allow_pickle=False)
    golden = {int(sd): v for sd, v in zip(g['seeds'], g['eigenvalues'])}
    seeds = list(range(5000, 5064))
    assert set(golden) <= set(seeds)
    _say(f'lobpcg 48^3, 64 variants in one call; golden eigsh seeds {sorted(golden)}')
    pr0, s = _modal_system(48, seeds[0])
    fixed, _ = s.dirichlet(pr0.fixed_dofs)
    MD = device.apply_dirichlet(s.mass_bsr(), fixed)
    KDs = []
    for sd in seeds:
        s.set_materials(meshgen.cube_modal(n=48, seed=sd).materials)
        KDs.append(device.apply_dirichlet(s.stiffness_bsr(), fixed))
    coarse = s.coarse_space(pr0.fixed_dofs)
    tol = 1e-6
    vals, _, info = modal.lobpcg(KDs, MD, 20, fixed=fixed, tol=tol, max_iter=3000, coarse=coarse,
                                 streams=2)
    its = info['iterations']
    _say(f'  iterations {its.min()}..{its.max()}, status {info["status"]}, max residual '
         f'{info["resid"].max():.2e}')
    assert info['status'] == 0
    assert info['resid'].shape == (64, 20) and info['resid'].max() <= tol
    assert np.all(np.diff(vals, axis=1) >= 0)
    for sd, ref in sorted(golden.items()):
        rel = np.abs(vals[sd - 5000] - ref) / ref
        _say(f'  seed {sd}: max relative difference to eigsh {rel.max():.2e}')
        assert rel.max() < 1e-8


def test_lobpcg_largest_unconstrained_vs_eigsh(ctx):
    _say('lobpcg largest unconstrained')
    pr, s = _modal_system(4, 5001)
    vals, vecs, info = modal.lobpcg(s.stiffness_bsr(), s.mass_bsr(), 6, block=8, precond=None,
                                    largest=True, tol=1e-8, max_iter=5000)
    _say(f'  iterations {info["iterations"]}, status {info["status"]}')
    assert info['status'] == 0
    Ko, Mo = _oracle_pencil(pr)
    ref_vals, _ = eigen.largest(Ko, Mo, 6)
    assert np.all(np.diff(vals[0]) >= 0)
    _check_pairs(vals[0], vecs[0].cpu().numpy(), Ko, Mo, None, ref_vals)


def test_lobpcg_batched_variants_share_mass(ctx):
    """Three stiffness variants of one mesh (cfg5's lognormal per-element perturbation) in one call;
    each against its own oracle."""
    _say('lobpcg batched variants')
    seeds = (5000, 5001, 5002)
    pr0, s = _modal_system(5, seeds[0])
    fixed, _ = s.dirichlet(pr0.fixed_dofs)
    MD = device.apply_dirichlet(s.mass_bsr(), fixed)
    KDs, prs = [], []
    for sd in seeds:
        pr = meshgen.cube_modal(n=5, seed=sd)
        s.set_materials(pr.materials)
        KDs.append(device.apply_dirichlet(s.stiffness_bsr(), fixed))
        prs.append(pr)
    vals, vecs, info = modal.lobpcg(KDs, MD, 8, block=12, fixed=fixed, tol=1e-7, max_iter=3000)
    _say(f'  iterations {info["iterations"]}')
    assert info['status'] == 0
    for v, pr in enumerate(prs):
        Ko, Mo = _oracle_pencil(pr)
        ref_vals, _ = eigen.smallest_clamped(Ko, Mo, pr.fixed_dofs, 8)
        _check_pairs(vals[v], vecs[v].cpu().numpy(), Ko, Mo, pr.fixed_dofs, ref_vals)
    # a single-variant call reproduces the batched result of that variant
    v1, _, _ = modal.lobpcg(KDs[1], MD, 8, block=12, fixed=fixed, tol=1e-7, max_iter=3000,
                            X0=None, seed=0)
    np.testing.assert_allclose(v1[0], vals[1], rtol=1e-9)
    # the plane layout of K and M gives the same iteration bit for bit (same per-row sum order)
    vp_, xp_, ip_ = modal.lobpcg(KDs, MD, 8, block=12, fixed=fixed, tol=1e-7, max_iter=3000,
                                 iteration_layout=_hip.FE_LAYOUT_PLANES)
    assert np.array_equal(vp_, vals) and torch.equal(xp_, vecs)
    assert np.array_equal(ip_['iterations'], info['iterations'])
    # three variants over two streams (2 + 1): the same bits
    vs_, xs_, is_ = modal.lobpcg(KDs, MD, 8, block=12, fixed=fixed, tol=1e-7, max_iter=3000, streams=2)
    assert np.array_equal(vs_, vals) and torch.equal(xs_, vecs)
    assert np.array_equal(is_['iterations'], info['iterations'])
