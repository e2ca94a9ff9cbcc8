"""GPU parity of the HIP path (element kernels, fused and two-phase assembly, SpMV, Dirichlet
elimination, PCG) against the CPU oracle and the committed golden vectors.

Tolerances: element and assembled matrix entries within 1e-12 relative to max(|entry|, row max)
(the reference's FP64 matmul and scipy summation orders differ from the kernels'); sparsity pattern
and DOF numbering bit-exact; solutions: true FP64 residual at the requested tolerance and the error
against a direct FP64 solve reported and bounded.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from finite_elements_amd import _hip, device, meshgen
from finite_elements_amd.system import FESystem
from oracle import vectorized as vec
from tests.conftest import rowscaled_err

pytestmark = pytest.mark.gpu

DOF = {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}


def _perturbed(kind, rng):
    if kind == 'tet_elastic':
        c, t = meshgen.kuhn_tets(5)
        c = c + rng.uniform(-0.03, 0.03, c.shape)
        mats = {'E': rng.uniform(5e10, 3e11, t.shape[0]), 'nu': rng.uniform(0.2, 0.45, t.shape[0]),
                'rho': rng.uniform(1e3, 8e3, t.shape[0])}
    else:
        c, t = meshgen.structured_tri(23, 17, 2.0, 1.5)
        c = c + rng.uniform(-0.01, 0.01, c.shape)
        if kind == 'tri_elastic':
            mats = {'E': rng.uniform(5e10, 3e11, t.shape[0]), 'nu': rng.uniform(0.2, 0.45, t.shape[0]),
                    'thickness': rng.uniform(0.005, 0.5, t.shape[0]), 'rho': rng.uniform(1e3, 8e3, t.shape[0])}
        else:
            mats = {'mu': 4e-7 * np.pi * rng.uniform(1, 1000, t.shape[0])}
    return c, t, mats


CASES = [('tri_elastic', 1), ('tri_elastic', 2), ('tet_elastic', 0), ('tri_magnetic', 0)]


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize('kind,plane', CASES)
@pytest.mark.parametrize('what', ['stiffness', 'mass'])
def test_element_kernels(ctx, kind, plane, what):
    if kind == 'tri_magnetic' and what == 'mass':
        pytest.skip('the reference has no magnetic mass matrix')
    _say(f'element kernels {kind} plane={plane} {what}')
    rng = np.random.default_rng(11)
    c, t, mats = _perturbed(kind, rng)
    want = vec.element_matrices(kind, c, t, mats, plane, what)  # (E, m*m)
    dm = device.DeviceMesh(ctx, kind, c, t, mats, plane)
    got = device.element_matrices(dm, what, sym=False).cpu().numpy().T
    scale = np.abs(want).max(axis=1, keepdims=True)
    assert np.max(np.abs(got - want) / scale) < 1e-12
    m = int(round(np.sqrt(want.shape[1])))
    iu = np.triu_indices(m)
    gs = device.element_matrices(dm, what, sym=True).cpu().numpy().T
    assert np.array_equal(gs, got.reshape(-1, m, m)[:, iu[0], iu[1]])


@pytest.mark.parametrize('kind,plane', CASES)
def test_assembly_matches_scipy(ctx, kind, plane):
    _say(f'assembly {kind} plane={plane}')
    rng = np.random.default_rng(12)
    c, t, mats = _perturbed(kind, rng)
    K = vec.assemble_csr(kind, c, t, mats, plane)
    s = FESystem(kind, c, t, mats, plane_mode=plane, with_gather=True)
    assert np.array_equal(s.pattern.rowptr, K.indptr) and np.array_equal(s.pattern.colind, K.indices)
    fused = s.stiffness(_hip.FE_LAYOUT_CSR).cpu().numpy()
    two = s.assemble_two_phase('stiffness').cpu().numpy()
    got = K.copy()
    got.data = fused
    err_f = rowscaled_err(K, got)
    got.data = two
    err_t = rowscaled_err(K, got)
    _say(f'  fused err {err_f:.2e} two-phase err {err_t:.2e}')
    assert err_f < 1e-12 and err_t < 1e-12
    # both sum the same element values in ascending element order (the two kernels may contract
    # multiply-adds differently, so agreement is to the last bits, not bitwise)
    assert np.max(np.abs(fused - two)) <= 1e-15 * np.abs(fused).max()
    # deterministic run to run
    assert np.array_equal(fused, s.stiffness(_hip.FE_LAYOUT_CSR).cpu().numpy())
    assert np.array_equal(two, s.assemble_two_phase('stiffness').cpu().numpy())
    # the BSR and plane layouts hold the same numbers permuted
    for layout in (_hip.FE_LAYOUT_BSR, _hip.FE_LAYOUT_PLANES):
        bsr = s.stiffness_bsr(layout)
        assert np.array_equal(bsr.csr_values().cpu().numpy(), fused)
    if kind != 'tri_magnetic':
        M = vec.assemble_csr(kind, c, t, mats, plane, what='mass')
        gm = M.copy()
        gm.data = s.mass(_hip.FE_LAYOUT_CSR).cpu().numpy()
        assert rowscaled_err(M, gm) < 1e-12


@pytest.mark.parametrize('threads', [64, 128, 256])
@pytest.mark.parametrize('max_inc', [32, 64, 128, 192])
def test_fused_schedules_agree(ctx, max_inc, threads):
    _say(f'fused schedule max_inc={max_inc} threads={threads}')
    c, t = meshgen.kuhn_tets(6)
    mats = {'E': np.linspace(1e10, 2e11, t.shape[0]), 'nu': 0.3}
    s = FESystem('tet_elastic', c, t, mats)
    ref_vals = s.stiffness().cpu().numpy()
    ref_planes = s.stiffness(_hip.FE_LAYOUT_PLANES).cpu().numpy()
    s.set_schedule(max_inc, threads)
    assert np.array_equal(s.stiffness().cpu().numpy(), ref_vals)
    assert np.array_equal(s.stiffness(_hip.FE_LAYOUT_PLANES).cpu().numpy(), ref_planes)


@pytest.mark.parametrize('kind', ['tet_elastic', 'tri_elastic', 'tri_magnetic'])
@pytest.mark.parametrize('max_inc,threads', [(48, 64), (96, 128), (192, 256), (384, 512)])
def test_persistent_assembly_equals_one_range_per_workgroup(ctx, kind, max_inc, threads):
    """The persistent fused assembly (headers + element node lists present) writes the same bits as
    the one-range-per-workgroup kernel (node lists withheld), stiffness and mass, every layout."""
    _say(f'persistent assembly {kind} max_inc={max_inc} threads={threads}')
    rng = np.random.default_rng(21)
    c, t, mats = _perturbed(kind, rng)
    s = FESystem(kind, c, t, mats, plane_mode=_hip.FE_PLANE_STRESS if kind == 'tri_elastic' else 0)
    s.set_schedule(max_inc, threads)
    g = s.graph
    assert g.welem_conn is not None and g.max_nodes < threads and g.max_elem <= threads
    whats = ['stiffness'] + ([] if kind == 'tri_magnetic' else ['mass'])
    layouts = [_hip.FE_LAYOUT_CSR, _hip.FE_LAYOUT_BSR, _hip.FE_LAYOUT_PLANES]
    got = {(w, l): (s.stiffness(l) if w == 'stiffness' else s.mass(l)).cpu().numpy()
           for w in whats for l in layouts}
    saved = g.welem_conn
    g.welem_conn = None
    try:
        for w in whats:
            for l in layouts:
                ref = (s.stiffness(l) if w == 'stiffness' else s.mass(l)).cpu().numpy()
                assert np.array_equal(got[(w, l)], ref), (w, l)
    finally:
        g.welem_conn = saved


def _perturbed_large(kind, rng):
    """Meshes with hundreds to thousands of node ranges at these schedules, so that every persistent
    workgroup walks several ranges (both LDS slots, the register prefetch and the round loop)."""
    if kind == 'tet_elastic':
        c, t = meshgen.kuhn_tets(24)
        c = c + rng.uniform(-0.01, 0.01, c.shape)
    else:
        c, t = meshgen.structured_tri(200, 150, 2.0, 1.5)
        c = c + rng.uniform(-0.002, 0.002, c.shape)
    n = t.shape[0]
    if kind == 'tri_magnetic':
        return c, t, {'mu': 4e-7 * np.pi * rng.uniform(1, 1000, n)}
    mats = {'E': rng.uniform(5e10, 3e11, n), 'nu': rng.uniform(0.2, 0.45, n), 'rho': rng.uniform(1e3, 8e3, n)}
    if kind == 'tri_elastic':
        mats['thickness'] = rng.uniform(0.005, 0.5, n)
    return c, t, mats


@pytest.mark.parametrize('kind', ['tet_elastic', 'tri_elastic', 'tri_magnetic'])
@pytest.mark.parametrize('threads,producers,max_inc', [(512, 2, 64), (512, 4, 256), (1024, 4, 96),
                                                       (1024, 6, 384), (1024, 8, 512)])
def test_wave_specialised_assembly_equals_one_range_per_workgroup(ctx, kind, threads, producers, max_inc):
    """The wave-specialised persistent assembly (producer and consumer wavefronts, two LDS slots)
    writes the same bits as the one-range-per-workgroup kernel: stiffness and mass, every layout."""
    _say(f'wave-specialised assembly {kind} threads={threads} producers={producers} max_inc={max_inc}')
    rng = np.random.default_rng(23)
    c, t, mats = _perturbed_large(kind, rng)
    s = FESystem(kind, c, t, mats, plane_mode=_hip.FE_PLANE_STRAIN if kind == 'tri_elastic' else 0)
    s.set_schedule(max_inc, threads, producers)
    g = s.graph
    # more ranges than the grid has workgroups (one of 1024 threads or two of 512 per CU on 256 CUs):
    # the workgroups walk several ranges
    assert g.producers == producers and g.n_blk > 256 * (1024 // threads), g.n_blk
    whats = ['stiffness'] + ([] if kind == 'tri_magnetic' else ['mass'])
    layouts = [_hip.FE_LAYOUT_CSR, _hip.FE_LAYOUT_BSR, _hip.FE_LAYOUT_PLANES]
    got = {(w, l): (s.stiffness(l) if w == 'stiffness' else s.mass(l)).cpu().numpy()
           for w in whats for l in layouts}
    g.producers = 0
    saved = g.welem_conn
    g.welem_conn = None
    try:
        for w in whats:
            for l in layouts:
                ref = (s.stiffness(l) if w == 'stiffness' else s.mass(l)).cpu().numpy()
                assert np.array_equal(got[(w, l)], ref), (w, l)
    finally:
        g.welem_conn = saved
        g.producers = producers
    # a schedule that does not fit the producers' registers is refused, not run another way
    s.set_schedule(max_inc, threads, producers)
    g.max_elem = 64 * producers + 1
    with pytest.raises(ValueError, match='wave-specialised'):
        s.stiffness(_hip.FE_LAYOUT_PLANES)


def test_wave_specialised_assembly_full_size(ctx):
    """At 64^3 (823,875 DOF, 12,481 ranges of at most 512 incidences) the wave-specialised kernel
    writes the same bits as the one-range-per-workgroup kernel."""
    _say('wave-specialised assembly at 64^3')
    pr = meshgen.cube_elasticity(n=64)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    s.set_schedule(512, 1024, 8)
    a = s.stiffness(_hip.FE_LAYOUT_PLANES)
    g = s.graph
    g.producers = 0
    saved = g.welem_conn
    g.welem_conn = None
    try:
        b = s.stiffness(_hip.FE_LAYOUT_PLANES)
    finally:
        g.welem_conn = saved
        g.producers = 8
    assert torch.equal(a, b)


def test_autotuner_picks_a_valid_schedule(ctx):
    _say('assembly autotuner')
    c, t = meshgen.kuhn_tets(8)
    s = FESystem('tet_elastic', c, t, {'E': 1e11, 'nu': 0.3})
    ref_vals = s.stiffness().cpu().numpy()
    best, times = s.tune_assembly(reps=2)
    assert best in times and len(times) >= 3
    assert np.array_equal(s.stiffness().cpu().numpy(), ref_vals)


def test_spmv(ctx):
    _say('spmv')
    rng = np.random.default_rng(13)
    for kind, plane in CASES:
        c, t, mats = _perturbed(kind, rng)
        K = vec.assemble_csr(kind, c, t, mats, plane)
        s = FESystem(kind, c, t, mats, plane_mode=plane)
        A = s.stiffness_bsr()
        x = rng.standard_normal(s.n)
        y = device.spmv(A, ctx.tensor(x)).cpu().numpy()
        want = K @ x
        assert np.max(np.abs(y - want)) <= 1e-12 * np.abs(K).max() * np.abs(x).sum() / 10
        yb = device.spmv(s.stiffness_bsr(_hip.FE_LAYOUT_BSR), ctx.tensor(x)).cpu().numpy()
        assert np.array_equal(yb, y)
        y32 = device.spmv(A.to_f32(), ctx.tensor(x.astype(np.float32))).cpu().numpy()
        assert np.linalg.norm(y32 - want) <= 1e-5 * np.linalg.norm(want) + 1e-6 * np.abs(K).max()
        # the solvers' iteration copies on the padded pattern (16-byte vector loads)
        yp = device.spmv(A.padded(torch.float64), ctx.tensor(x)).cpu().numpy()
        assert np.max(np.abs(yp - want)) <= 1e-12 * np.abs(K).max() * np.abs(x).sum() / 10
        yp32 = device.spmv(A.padded(torch.float32), ctx.tensor(x.astype(np.float32))).cpu().numpy()
        assert np.linalg.norm(yp32 - want) <= 1e-5 * np.linalg.norm(want) + 1e-6 * np.abs(K).max()
        rp = ctx.tensor(s.pattern.rowptr)
        ci = ctx.tensor(s.pattern.colind)
        yc = device.spmv_csr(ctx, rp, ci, s.stiffness(_hip.FE_LAYOUT_CSR), ctx.tensor(x)).cpu().numpy()
        np.testing.assert_allclose(yc, y, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def _solve_case(name):
    if name == 'cantilever':
        return meshgen.cantilever_2d()
    if name == 'magnetic':
        return meshgen.magnetostatic_2d(n=60)
    if name == 'cube':
        return meshgen.cube_elasticity(n=8, sigma=1.0)
    raise KeyError(name)


@pytest.mark.parametrize('name', ['cantilever', 'magnetic', 'cube'])
@pytest.mark.parametrize('precond,precision', [('jacobi', 'fp64'), ('block_jacobi', 'fp64'),
                                               ('jacobi', 'mixed'), ('block_jacobi', 'mixed')])
def test_pcg_solution(ctx, name, precond, precision):
    _say(f'pcg {name} {precond} {precision}')
    pr = _solve_case(name)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=pr.plane_mode)
    K = vec.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, pr.plane_mode)
    u_ref = vec.reduced_solve(K, pr.f, pr.fixed_dofs)
    tol = 1e-10
    u, Kd, info = s.solve(pr.f, pr.fixed_dofs, tol=tol, precond=precond, precision=precision)
    u = u.cpu().numpy()
    _say(f'  {info}')
    assert info['status'] == 0
    # true FP64 residual of the eliminated system, recomputed on the CPU
    free = np.ones(s.n, bool)
    free[pr.fixed_dofs] = False
    r = (pr.f - K @ u)[free]
    assert np.linalg.norm(r) <= 1.01 * tol * np.linalg.norm(pr.f[free])
    assert np.all(u[pr.fixed_dofs] == 0.0)
    err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    _say(f'  relative error vs direct FP64 solve: {err:.2e}')
    assert err < 1e-5


def test_pcg_deterministic(ctx):
    _say('pcg determinism')
    pr = meshgen.cube_elasticity(n=6)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    u1, _, i1 = s.solve(pr.f, pr.fixed_dofs, tol=1e-9, precond='block_jacobi', precision='mixed')
    u2, _, i2 = s.solve(pr.f, pr.fixed_dofs, tol=1e-9, precond='block_jacobi', precision='mixed')
    assert i1['iterations'] == i2['iterations']
    assert np.array_equal(u1.cpu().numpy(), u2.cpu().numpy())


def test_pcg_failure_statuses(ctx):
    _say('pcg failure statuses')
    pr = meshgen.cantilever_2d(10, 3)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=2)
    _, _, info = s.solve(pr.f, pr.fixed_dofs, tol=1e-12, max_iter=3)
    assert info['status'] == _hip.FE_ERR_NOT_CONVERGED and info['iterations'] >= 3
    # a negative-definite operator breaks CG down (p.Ap <= 0) and is reported, not hung
    K = s.stiffness_bsr()
    fixed, g = s.dirichlet(pr.fixed_dofs)
    KD, b = s.prepare(K, fixed, g, ctx.tensor(pr.f))
    KD.vals.neg_()
    M = device.precond_setup(KD, 'jacobi').abs()
    solver = s.solver(KD, 'jacobi', 'fp64', M64=M)
    _, info = solver.solve(b, tol=1e-10)
    assert info['status'] == _hip.FE_ERR_BREAKDOWN
    with pytest.raises(ValueError):
        s.solver(KD, 'ilu', 'fp64')


def test_zero_rhs_and_exact_initial_guess(ctx):
    _say('pcg zero rhs')
    pr = meshgen.cantilever_2d(10, 3)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=2)
    u, _, info = s.solve(np.zeros(s.n), pr.fixed_dofs, tol=1e-10)
    assert info['status'] == 0 and info['iterations'] == 0
    assert float(u.abs().max()) == 0.0
