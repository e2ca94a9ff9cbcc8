"""Two-level preconditioner on the GPU: coarse operator, its inverse and the coarse correction against
the oracle (oracle/coarse.py), and two-level PCG against direct solves."""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from finite_elements_amd import _hip, device, meshgen
from finite_elements_amd.system import FESystem
from oracle import coarse as oc, vectorized as vz

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _host_matrix(KD):
    d = KD.dof
    vals = KD.convert(_hip.FE_LAYOUT_BSR).vals.cpu().numpy().reshape(-1, d, d)
    g = KD.graph.pattern
    n = g.n_nodes * d
    return sp.bsr_matrix((vals, g.bcol, g.bptr), shape=(n, n)).tocsr()


def _setup(pr, target):
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=pr.plane_mode)
    fixed, g = s.dirichlet(pr.fixed_dofs, pr.fixed_values)
    KD, b = s.prepare(s.stiffness_bsr(), fixed, g, s.ctx.tensor(pr.f))
    dc = s.coarse_space(pr.fixed_dofs, target=target)
    return s, KD, b, dc


@pytest.mark.parametrize('case', ['cube', 'cantilever', 'magnetic', 'cube16'])
def test_coarse_operator_and_inverse(ctx, case):
    """cube16: 216 aggregates, 1,296 coarse DOF -- 21 panels of the blocked Gauss-Jordan inverse."""
    _say(f'coarse operator {case}')
    pr = {'cube': lambda: meshgen.cube_elasticity(n=6),
          'cantilever': lambda: meshgen.cantilever_2d(),
          'magnetic': lambda: meshgen.magnetostatic_2d(n=30),
          'cube16': lambda: meshgen.cube_elasticity(n=16)}[case]()
    s, KD, b, dc = _setup(pr, 27 if case in ('cube', 'cube16') else 16)
    A = _host_matrix(KD)
    ref, P = oc.coarse_operator(dc.host, A)
    n = ref.shape[0]
    Ac, _ = dc.operator(KD, invert=False)
    Ac = Ac.cpu().numpy().reshape(dc.ld, dc.ld)
    assert np.abs(Ac[:n, :n] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(Ac[n:, n:], np.eye(dc.ld - n)) and not Ac[:n, n:].any()
    inv64, inv32 = dc.operator(KD, invert=True)
    inv64 = inv64.cpu().numpy().reshape(dc.ld, dc.ld)[:n, :n]
    refinv = np.linalg.inv(ref)
    err = np.abs(inv64 - refinv).max() / np.abs(refinv).max()
    _say(f'  n_coarse {n}, inverse rel err {err:.2e}, cond {np.linalg.cond(ref):.2e}')
    assert err <= 1e-12 * np.linalg.cond(ref)
    np.testing.assert_allclose(inv32.cpu().numpy().reshape(dc.ld, dc.ld)[:n, :n], inv64.astype(np.float32),
                               rtol=0, atol=0)
    # the correction u += P A_c^-1 P^T r in both precisions
    rng = np.random.default_rng(1)
    r = rng.standard_normal(A.shape[0])
    want = P @ np.linalg.solve(ref, P.T @ r)
    u = ctx.tensor(np.zeros(A.shape[0]))
    dc.apply(dc.operator(KD)[0], ctx.tensor(r), u)
    assert np.abs(u.cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()
    u32 = torch.zeros(A.shape[0], dtype=torch.float32, device=ctx.torch_device)
    dc.apply(dc.operator(KD)[1], ctx.tensor(r).float(), u32)
    assert np.abs(u32.cpu().numpy() - want).max() <= 2e-3 * np.abs(want).max()
    # the stored-P form of the correction (the moment form's arrays withheld) gives the same result
    saved = (dc.xhat, dc.xhat32, dc.T, dc.T32)
    dc.xhat = dc.xhat32 = dc.T = dc.T32 = None
    try:
        Ac_p, _ = dc.operator(KD, invert=False)
        Ac_p = Ac_p.cpu().numpy().reshape(dc.ld, dc.ld)
        assert np.abs(Ac_p[:n, :n] - Ac[:n, :n]).max() <= 1e-12 * np.abs(ref).max()
        up = ctx.tensor(np.zeros(A.shape[0]))
        dc.apply(dc.operator(KD)[0], ctx.tensor(r), up)
        assert np.abs(up.cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()
        assert np.abs(up.cpu().numpy() - u.cpu().numpy()).max() <= 1e-12 * np.abs(want).max()
        up32 = torch.zeros(A.shape[0], dtype=torch.float32, device=ctx.torch_device)
        dc.apply(dc.operator(KD)[1], ctx.tensor(r).float(), up32)
        assert np.abs(up32.cpu().numpy() - u32.cpu().numpy()).max() <= 1e-5 * np.abs(want).max()
    finally:
        dc.xhat, dc.xhat32, dc.T, dc.T32 = saved


@pytest.mark.parametrize('case', ['cube20', 'magnetic80', 'cantilever160'])
def test_coarse_correction_large_aggregates(ctx, case):
    """Aggregates of 600+ nodes in one system (the restriction's 1,024-thread form, two nodes per
    thread for some): the correction u += P A_c^-1 P^T r against the oracle and against the stored-P
    form in both precisions, on 3, 1 and 2 DOF per node."""
    _say(f'coarse correction, large aggregates: {case}')
    # 8 aggregates of 1,000-1,331 nodes (two per thread for some); 9 of 729 (coarse dimension 9: a
    # column past the last 4-wide piece); 5 of 1,312-1,353
    pr, target = {'cube20': lambda: (meshgen.cube_elasticity(n=20), 729),
                  'magnetic80': lambda: (meshgen.magnetostatic_2d(n=80), 900),
                  'cantilever160': lambda: (meshgen.cantilever_2d(nx=160, ny=40), 900)}[case]()
    s, KD, b, dc = _setup(pr, target)
    n_nodes, n_agg = dc.host.node_agg.size, dc.host.n_agg
    assert dc.n_seg == 0 and n_nodes // n_agg >= 600, (n_nodes, n_agg)
    A = _host_matrix(KD)
    ref, P = oc.coarse_operator(dc.host, A)
    rng = np.random.default_rng(7)
    r = rng.standard_normal(A.shape[0])
    want = P @ np.linalg.solve(ref, P.T @ r)
    inv64, inv32 = dc.operator(KD)
    u = ctx.tensor(np.zeros(A.shape[0]))
    dc.apply(inv64, ctx.tensor(r), u)
    u32 = torch.zeros(A.shape[0], dtype=torch.float32, device=ctx.torch_device)
    dc.apply(inv32, ctx.tensor(r).float(), u32)
    scale = np.abs(want).max()
    e64 = np.abs(u.cpu().numpy() - want).max() / scale
    e32 = np.abs(u32.cpu().numpy() - want).max() / scale
    _say(f'  {n_agg} aggregates of ~{n_nodes // n_agg} nodes: rel err fp64 {e64:.2e}, fp32 {e32:.2e}')
    assert e64 <= 1e-10 and e32 <= 2e-3
    saved = (dc.xhat, dc.xhat32, dc.T, dc.T32)
    dc.xhat = dc.xhat32 = dc.T = dc.T32 = None
    try:
        p64, p32 = dc.operator(KD)
        up = ctx.tensor(np.zeros(A.shape[0]))
        dc.apply(p64, ctx.tensor(r), up)
        up32 = torch.zeros(A.shape[0], dtype=torch.float32, device=ctx.torch_device)
        dc.apply(p32, ctx.tensor(r).float(), up32)
    finally:
        dc.xhat, dc.xhat32, dc.T, dc.T32 = saved
    assert np.abs(up.cpu().numpy() - u.cpu().numpy()).max() <= 1e-12 * scale
    assert np.abs(up32.cpu().numpy() - u32.cpu().numpy()).max() <= 1e-5 * scale


def test_inverse_queued_from_the_assembled_matrix(ctx):
    """FESystem.prepare(coarse=...) queues the coarse operator and inverse from the assembled K on the
    side stream, beside the elimination: the inverse is bitwise the one the solver forms from K_D
    (constrained rows and columns are read as zero either way), and so is the solve."""
    _say('coarse inverse queued from K')
    pr = meshgen.cube_elasticity(n=12)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    fixed, g = s.dirichlet(pr.fixed_dofs, pr.fixed_values)
    dc = s.coarse_space(pr.fixed_dofs, target=64)
    f = s.ctx.tensor(pr.f)
    K = s.stiffness_bsr()
    KD1, b1 = s.prepare(K, fixed, g, f)
    KD2, b2 = s.prepare(K, fixed, g, f, coarse=dc)
    assert getattr(KD1, 'coarse_inverse', None) is None and KD2.coarse_inverse[0] is dc
    sol1 = s.solver(KD1, 'two_level', 'mixed', coarse=dc)
    sol2 = s.solver(KD2, 'two_level', 'mixed', coarse=dc)
    assert sol2.Ainv64 is KD2.coarse_inverse[1]
    u1, i1 = sol1.solve(b1, tol=1e-8)
    u2, i2 = sol2.solve(b2, tol=1e-8)
    torch.cuda.synchronize()
    n, ld = dc.n_coarse, dc.ld  # the FP32 copy holds the n x n inverse (its padding is not written)
    assert torch.equal(sol1.Ainv64, sol2.Ainv64)
    assert torch.equal(sol1.Ainv32.view(ld, ld)[:n, :n], sol2.Ainv32.view(ld, ld)[:n, :n])
    assert torch.equal(b1, b2) and torch.equal(KD1.vals, KD2.vals)
    assert i1['status'] == 0 and i1['iterations'] == i2['iterations'] and torch.equal(u1, u2)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_two_level_pcg_matches_direct_and_oracle_iterations(ctx, precision):
    _say(f'two-level pcg {precision}')
    pr = meshgen.cube_elasticity(n=10)
    s, KD, b, dc = _setup(pr, 64)
    tol = 1e-10 if precision == 'fp64' else 1e-8
    u, info = s.solver(KD, 'two_level', precision, coarse=dc).solve(b, tol=tol)
    _, ib = s.solver(KD, 'block_jacobi', precision).solve(b, tol=tol)
    _say(f'  two-level {info["iterations"]} it (relres {info["relres"]:.2e}), block Jacobi {ib["iterations"]} it')
    assert info['status'] == 0 and info['relres'] <= tol
    assert info['iterations'] < 0.5 * ib['iterations']
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials)
    ref = vz.reduced_solve(K, pr.f, pr.fixed_dofs)
    assert np.linalg.norm(u.cpu().numpy() - ref) <= 1e-5 * np.linalg.norm(ref)
    if precision == 'fp64':
        # the oracle's textbook two-level PCG takes about as many iterations
        A = _host_matrix(KD)
        binv = oc.block_jacobi_inverse(A, 3)
        bh = b.cpu().numpy()
        _, it_o = oc.pcg(A, bh, lambda r: oc.apply_two_level(dc.host, A, binv, r), tol=tol)
        _say(f'  oracle two-level PCG {it_o} it')
        assert abs(info['iterations'] - it_o) <= 0.1 * it_o + 3


def test_two_level_2d_and_determinism(ctx):
    _say('two-level 2D')
    pr = meshgen.cantilever_2d()
    s, KD, b, dc = _setup(pr, 16)
    sol = s.solver(KD, 'two_level', 'fp64', coarse=dc)
    u1, i1 = sol.solve(b, tol=1e-10)
    u2, i2 = sol.solve(b, tol=1e-10)
    assert i1['status'] == 0 and i1['iterations'] == i2['iterations']
    assert np.array_equal(u1.cpu().numpy(), u2.cpu().numpy())
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=pr.plane_mode)
    ref = vz.reduced_solve(K, pr.f, pr.fixed_dofs)
    assert np.linalg.norm(u1.cpu().numpy() - ref) <= 1e-6 * np.linalg.norm(ref)
