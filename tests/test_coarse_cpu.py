"""Host side of the two-level preconditioner: aggregates, the orthonormal rigid-body prolongator and the
coarse pair map (finite_elements_amd/coarse.py) against the oracle's assembled P^T A P."""

import numpy as np
import scipy.sparse as sp

from finite_elements_amd import meshgen
from finite_elements_amd.coarse import CoarseSpace, rigid_modes
from oracle import coarse as oc, vectorized as vz


def _cube(n=6):
    pr = meshgen.cube_elasticity(n=n)
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials).tocsr()
    B = sp.csr_matrix((np.ones(K.nnz), K.indices, K.indptr), shape=K.shape)
    # node graph from the scalar pattern
    nn = pr.coords.shape[0]
    G = sp.csr_matrix(K.shape)
    rows = np.repeat(np.arange(K.shape[0]), np.diff(K.indptr)) // 3
    cols = K.indices // 3
    Gn = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(nn, nn))
    Gn.sum_duplicates()
    Gn.sort_indices()
    return pr, K, Gn.indptr, Gn.indices


def test_rigid_modes_are_zero_energy():
    pr, K, bptr, bcol = _cube(4)
    Z = rigid_modes(pr.coords - pr.coords.mean(axis=0), 3).reshape(-1, 6)
    # an unconstrained stiffness maps every rigid-body mode to zero force
    assert np.abs(K @ Z).max() <= 1e-6 * np.abs(K).max()


def test_prolongator_orthonormal_and_constrained():
    pr, K, bptr, bcol = _cube(6)
    cs = CoarseSpace(pr.coords, bptr, bcol, 3, pr.fixed_dofs, target=27)
    P = oc.prolongator(cs)
    PtP = (P.T @ P).toarray()
    keep = np.diag(PtP) > 0
    np.testing.assert_allclose(PtP[np.ix_(keep, keep)], np.eye(keep.sum()), atol=1e-12)
    assert np.abs(P[pr.fixed_dofs]).max() == 0.0
    assert cs.n_agg == len(np.unique(cs.node_agg))
    assert np.all(np.diff(cs.agg_ptr) > 0)
    # every block of the pattern sits in exactly one coarse pair, pairs sorted, rows consistent
    assert cs.pair_q.size == bcol.size and np.array_equal(np.sort(cs.pair_q), np.arange(bcol.size))
    rows = np.repeat(np.arange(len(bptr) - 1), np.diff(bptr))
    assert np.array_equal(cs.pair_row, rows[cs.pair_q])
    for p in range(cs.n_pairs):
        qs = cs.pair_q[cs.pair_ptr[p]:cs.pair_ptr[p + 1]]
        assert np.all(cs.node_agg[rows[qs]] == cs.pair_ij[p, 0])
        assert np.all(cs.node_agg[bcol[qs]] == cs.pair_ij[p, 1])


def test_two_level_oracle_reduces_iterations():
    """The oracle's two-level PCG needs far fewer iterations than block Jacobi (the point of it)."""
    pr, K, bptr, bcol = _cube(8)
    cs = CoarseSpace(pr.coords, bptr, bcol, 3, pr.fixed_dofs, target=27)
    KD = vz.dirichlet_matrix(K, pr.fixed_dofs)
    b = pr.f.copy()
    b[pr.fixed_dofs] = 0.0
    binv = oc.block_jacobi_inverse(KD, 3)
    _, it_bj = oc.pcg(KD, b, lambda r: np.einsum('nij,nj->ni', binv, r.reshape(-1, 3)).reshape(-1))
    Ac, P = oc.coarse_operator(cs, KD)
    L = np.linalg.cholesky(Ac)

    def m2(r):
        u = np.einsum('nij,nj->ni', binv, r.reshape(-1, 3)).reshape(-1)
        return u + P @ np.linalg.solve(L.T, np.linalg.solve(L, P.T @ r))
    x, it_2 = oc.pcg(KD, b, m2)
    ref = vz.reduced_solve(K, pr.f, pr.fixed_dofs)
    assert np.linalg.norm(x - ref) <= 1e-6 * np.linalg.norm(ref)
    assert it_2 < 0.6 * it_bj, (it_2, it_bj)


def _moment_restrict(cs, r):
    """P^T r the way the moment kernels form it: w = (sum r, sum x^ x r) per aggregate over free
    DOFs, then r_c = T^T w."""
    d = cs.dof
    free = ((cs.fixed_bits[:, None] >> np.arange(d)[None, :]) & 1) == 0
    rv = np.where(free, r.reshape(-1, d), 0.0)
    x = cs.xhat
    if d == 3:
        m = np.concatenate([rv, np.cross(x, rv)], axis=1)
    else:
        m = np.concatenate([rv, (x[:, 0] * rv[:, 1] - x[:, 1] * rv[:, 0])[:, None]], axis=1)
    w = np.zeros((cs.n_agg, cs.nc))
    np.add.at(w, cs.node_agg, m)
    return np.einsum('gkc,gk->gc', cs.T, w).reshape(-1)


def _moment_prolong(cs, z):
    """P z the way the moment kernels form it: y = T z per aggregate, u = t + omega x x^ on free DOFs."""
    d = cs.dof
    y = np.einsum('gkc,gc->gk', cs.T, z.reshape(cs.n_agg, cs.nc))[cs.node_agg]
    x = cs.xhat
    if d == 3:
        u = y[:, :3] + np.cross(y[:, 3:], x)
    else:
        u = y[:, :2] + y[:, 2:3] * np.stack([-x[:, 1], x[:, 0]], axis=1)
    free = ((cs.fixed_bits[:, None] >> np.arange(d)[None, :]) & 1) == 0
    return np.where(free, u, 0.0).reshape(-1)


def test_moment_form_of_prolongator():
    """P = Z(x^) T: restriction by rigid-body moments and prolongation by rigid motions reproduce
    P^T r and P z of the assembled prolongator (3D, and 2D plane elasticity with rotations)."""
    rng = np.random.default_rng(3)
    pr, K, bptr, bcol = _cube(6)
    cs = CoarseSpace(pr.coords, bptr, bcol, 3, pr.fixed_dofs, target=27)
    assert cs.xhat.shape == pr.coords.shape and np.abs(cs.xhat).max() <= 1.0 + 1e-15
    P = oc.prolongator(cs)
    r = rng.standard_normal(P.shape[0])
    z = rng.standard_normal(P.shape[1])
    np.testing.assert_allclose(_moment_restrict(cs, r), P.T @ r, rtol=0, atol=1e-12 * np.abs(r).sum())
    np.testing.assert_allclose(_moment_prolong(cs, z), P @ z, rtol=0, atol=1e-12 * np.abs(z).sum())
    # 2D: a plate with its left edge clamped
    n = 9
    g = np.stack(np.meshgrid(np.linspace(0, 2, n), np.linspace(0, 1, n), indexing='ij'), -1).reshape(-1, 2)
    nn = g.shape[0]
    fixed = np.flatnonzero(np.repeat(g[:, 0] == 0, 2))
    ptr = np.arange(nn + 1)
    cs2 = CoarseSpace(g, ptr, np.arange(nn), 2, fixed, target=9)
    P2 = oc.prolongator(cs2)
    r2 = rng.standard_normal(P2.shape[0])
    z2 = rng.standard_normal(P2.shape[1])
    np.testing.assert_allclose(_moment_restrict(cs2, r2), P2.T @ r2, rtol=0, atol=1e-12 * np.abs(r2).sum())
    np.testing.assert_allclose(_moment_prolong(cs2, z2), P2 @ z2, rtol=0, atol=1e-12 * np.abs(z2).sum())


def test_block_diagonal_space_of_a_batch():
    """Aggregates of a batch stay inside their mesh, every mesh's block fits max_block, and P^T A P
    of the block-diagonal system is block diagonal with the layout the batched solver reads."""
    from finite_elements_amd.batch import concatenate
    plates = [pr for pr in meshgen.plates(count=60, seed=7) if pr.meta['nx'] <= 80][:5]
    kind, plane, coords, conn, mats, fixed, vals, f, offs = concatenate(plates)
    K = vz.assemble_csr(kind, coords, conn, mats, plane_mode=plane).tocsr()
    nn = coords.shape[0]
    rows = np.repeat(np.arange(K.shape[0]), np.diff(K.indptr)) // 2
    Gn = sp.csr_matrix((np.ones(rows.size), (rows, K.indices // 2)), shape=(nn, nn))
    Gn.sum_duplicates()
    Gn.sort_indices()
    cs = CoarseSpace(coords, Gn.indptr, Gn.indices, 2, fixed, node_offsets=offs, max_block=36)
    assert cs.agg_seg.size == len(plates) + 1 and cs.max_block <= 36
    for s in range(len(plates)):
        nodes = np.arange(offs[s], offs[s + 1])
        assert np.all(cs.seg_of_agg[cs.node_agg[nodes]] == s)
        assert set(np.unique(cs.node_agg[nodes])) == set(range(cs.agg_seg[s], cs.agg_seg[s + 1]))
    n_s = np.diff(cs.agg_seg) * cs.nc
    assert np.array_equal(cs.blk_ofs, np.concatenate([[0], np.cumsum(n_s * n_s)]))
    Ac, P = oc.coarse_operator(cs, vz.dirichlet_matrix(K, fixed))
    for s in range(len(plates)):
        c0, c1 = cs.agg_seg[s] * cs.nc, cs.agg_seg[s + 1] * cs.nc
        assert not Ac[c0:c1, :c0].any() and not Ac[c0:c1, c1:].any()
