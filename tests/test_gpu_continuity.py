"""Continuity conditions on the GPU path (finite_elements/conditions.py:196-241, analysis.py:175-199):
``FiniteElementAnalysis`` with periodic / anti-periodic links between the left and right sides of a
magnetostatic square, against the oracle's Lagrange formulation.

* small (12 x 12, with a chain of two links): ``create_matrix`` / ``create_source_matrix`` against
  the oracle's (pattern bit-exact, values 1e-12 row-scaled), and ``solve()`` -- potentials and every
  multiplier -- against the oracle's spsolve of the augmented system, for plain, block-Jacobi and
  two-level PCG, FP64 and mixed;
* full size (450^2 P1 magnetostatics, 203,401 DOF, 449 anti-periodic links): ``solve()`` with the
  headline two-level mixed PCG against the oracle's direct solve of the augmented system.

The folded matrix T^T K T is formed on the device by ``fe_bsr_fold``.
"""

import numpy as np
import pytest

from finite_elements_amd import conditions, loads, mesh, meshgen
from finite_elements_amd.analysis import FiniteElementAnalysis
from oracle import reference as ref
from oracle import vectorized as vz
from tests.conftest import rowscaled_err

pytestmark = pytest.mark.gpu
MU0 = 4e-7 * np.pi


def _say(m):
    print(m, flush=True)


def side_links(coords, value):
    x, y = coords[:, 0], coords[:, 1]
    x0, x1, y0, y1 = x.min(), x.max(), y.min(), y.max()
    left = np.nonzero((x == x0) & (y > y0) & (y < y1))[0]
    right = np.nonzero((x == x1) & (y > y0) & (y < y1))[0]
    left, right = left[np.argsort(y[left])], right[np.argsort(y[right])]
    assert np.array_equal(y[left], y[right])
    return left, right, np.full(left.size, float(value))


def analysis(coords, conn, mats, f, bnd, links):
    am = mesh.ArrayMesh(coords, conn, 'tri_magnetic', mats)
    nodes = am.nodes
    nl = [loads.NodeLoad(nodes[i], float(f[i]), 1) for i in np.nonzero(f)[0]]
    bcs = [conditions.NodeBoundaryCondition(nodes[i], 0.0, 1) for i in bnd]
    cc = [conditions.ContinuityCondition(nodes[a], nodes[b], float(v)) for a, b, v in zip(*links)]
    return FiniteElementAnalysis(am, [], [], nl, [], cc, bcs, [], [])


def small_case(value, chain):
    coords, conn = meshgen.structured_tri(12, 12)
    rng = np.random.default_rng(11)
    mats = {'mu': MU0 * rng.uniform(1.0, 100.0, conn.shape[0])}
    n = coords.shape[0]
    y = coords[:, 1]
    bnd = np.nonzero((y == 0.0) | (y == 1.0))[0]
    f = np.zeros(n)
    f[rng.choice(n, 8, replace=False)] = rng.standard_normal(8)
    s, m, v = side_links(coords, value)
    if chain:
        s, m, v = np.append(s, s[3] + 1), np.append(m, s[3]), np.append(v, 0.5)
    return coords, conn, mats, f, bnd, (s, m, v)


@pytest.mark.parametrize('value,chain', [(1.0, False), (-1.0, True)])
def test_small_matrix_and_source(ctx, value, chain):
    coords, conn, mats, f, bnd, links = small_case(value, chain)
    an = analysis(coords, conn, mats, f, bnd, links)
    pr = ref.problem_from_arrays('tri_magnetic', coords, conn, mats, fixed_dofs=bnd, f=f,
                                 continuity=list(zip(*links)))
    want = ref.create_matrix(pr).tocsr()
    got = an.create_matrix().tocsr()
    want.sort_indices()
    got.sort_indices()
    assert got.shape == want.shape
    assert np.array_equal(got.indptr, want.indptr) and np.array_equal(got.indices, want.indices)
    assert rowscaled_err(want, got) <= 1e-12
    assert np.array_equal(an.create_source_matrix(), ref.create_source_matrix(pr))


@pytest.mark.parametrize('value,chain', [(1.0, False), (-1.0, True)])
@pytest.mark.parametrize('precond,precision', [('jacobi', 'fp64'), ('block_jacobi', 'mixed'),
                                               ('two_level', 'mixed')])
def test_small_solve(ctx, value, chain, precond, precision):
    coords, conn, mats, f, bnd, links = small_case(value, chain)
    n = coords.shape[0]
    an = analysis(coords, conn, mats, f, bnd, links)
    pr = ref.problem_from_arrays('tri_magnetic', coords, conn, mats, fixed_dofs=bnd, f=f,
                                 continuity=list(zip(*links)))
    want = np.asarray(ref.solve(pr))
    got = np.asarray(an.solve(tol=1e-12, precond=precond, precision=precision).result_vector)
    assert got.size == want.size == n + bnd.size + links[0].size
    u, uw = got[:n], want[:n]
    s, m, v = links
    assert np.abs(u - uw).max() <= 1e-9 * np.abs(uw).max()
    assert np.abs(u[s] - v * u[m]).max() <= 1e-14 * np.abs(uw).max()
    fs = np.abs(f).max()
    assert np.abs(got[n:] - want[n:]).max() <= 1e-8 * fs
    info = an.last_solve_info
    _say(f'  value {value} chain {chain} {precond}/{precision}: {info["iterations"]} it, relres '
         f'{info["relres"]:.2e}')


@pytest.mark.timeout(300)
def test_magnetic450_antiperiodic(ctx):
    """Full size: the 450^2 magnetostatic square with A = 0 on top and bottom and anti-periodic
    left/right sides, through ``FiniteElementAnalysis.solve`` (two-level mixed PCG to 1e-8)."""
    pr = meshgen.magnetostatic_2d(n=450)
    coords, conn = pr.coords, pr.conn
    y = coords[:, 1]
    bnd = np.nonzero((y == -1.0) | (y == 1.0))[0]
    links = side_links(coords, -1.0)
    n = coords.shape[0]
    _say(f'  {n} DOF, {bnd.size} Dirichlet rows, {links[0].size} anti-periodic links; oracle KKT solve')
    K = vz.assemble_csr('tri_magnetic', coords, conn, pr.materials)
    uw, lbw, lcw = vz.kkt_solve(K, pr.f, bnd, continuity=np.stack(links, axis=1))
    an = analysis(coords, conn, pr.materials, pr.f, bnd, links)
    res = an.solve(tol=1e-8, precond='two_level', precision='mixed')
    x = np.asarray(res.result_vector)
    info = an.last_solve_info
    u = x[:n]
    err = float(np.linalg.norm(u - uw) / np.linalg.norm(uw))
    s, m, v = links
    # the constrained system's true residual, on the oracle's K: T^T (f - K u) on the free masters
    r = pr.f - K @ u
    rr = r.copy()
    np.add.at(rr, m, v * r[s])
    free = np.ones(n, dtype=bool)
    free[bnd] = False
    free[s] = False
    relres = float(np.linalg.norm(rr[free]) / np.linalg.norm((pr.f + np.bincount(m, v * pr.f[s], n))[free]))
    # multipliers are residual entries: their error is relres-sized relative to the load scale (the
    # continuity multipliers here are ~1e-5 of max|f|, so their own relative error is larger)
    fs = np.abs(pr.f).max()
    cc_err = float(np.abs(x[n + bnd.size:] - lcw).max())
    bc_err = float(np.abs(x[n:n + bnd.size] - lbw).max())
    _say(f'  two-level mixed: {info["iterations"]} it, GPU relres {info["relres"]:.2e}, CPU relres '
         f'{relres:.2e}, error vs oracle {err:.2e}; multiplier errors / max|f|: continuity '
         f'{cc_err / fs:.2e} (relative to max|lambda| {cc_err / np.abs(lcw).max():.2e}), Dirichlet '
         f'{bc_err / fs:.2e}')
    assert info['status'] == 0 and info['relres'] <= 1e-8
    assert relres <= 1e-8 + 1e-12
    assert np.all(u[s] == v * u[m])
    assert err <= 1e-5
    assert cc_err <= 1e-7 * fs and bc_err <= 1e-7 * fs
