"""Host construction of the padded pattern the solvers' SpMV walks with 16-byte loads."""

import numpy as np
import scipy.sparse as sp

from finite_elements_amd import meshgen
from finite_elements_amd.device import pad_pattern


def _node_pattern(conn, n_nodes):
    """Node-block pattern (bptr, bcol) of a mesh: nodes sharing an element, columns sorted."""
    npe = conn.shape[1]
    r = np.repeat(conn, npe, axis=1).ravel()
    c = np.tile(conn, (1, npe)).ravel()
    g = sp.csr_matrix((np.ones(r.size), (r, c)), shape=(n_nodes, n_nodes))
    g.sum_duplicates()
    g.sort_indices()
    return g.indptr.astype(np.int32), g.indices.astype(np.int32)


def test_pad_pattern_structure_and_product():
    rng = np.random.default_rng(5)
    coords, conn = meshgen.kuhn_tets(4)
    n = coords.shape[0]
    bptr, bcol = _node_pattern(conn, n)
    pptr, pcol, src = pad_pattern(bptr, bcol, 4)
    deg, pdeg = np.diff(bptr), np.diff(pptr)
    # every row starts at a multiple of 4 and holds the smallest multiple of 4 >= its degree
    assert np.all(pptr % 4 == 0) and np.all(pdeg == (deg + 3) // 4 * 4)
    # the original blocks keep their order at the head of each row; padding points at the row
    for i in range(n):
        head = slice(pptr[i], pptr[i] + deg[i])
        assert np.array_equal(pcol[head], bcol[bptr[i]:bptr[i + 1]])
        assert np.array_equal(src[head], np.arange(bptr[i], bptr[i + 1]))
        tail = slice(pptr[i] + deg[i], pptr[i + 1])
        assert np.all(pcol[tail] == i) and np.all(src[tail] == -1)
    # a block matrix gathered onto the padded pattern (zeros in the padding) has the same product
    d = 3
    vals = rng.standard_normal((bcol.size, d, d))
    pvals = np.where(src[:, None, None] >= 0, vals[np.maximum(src, 0)], 0.0)
    A = sp.bsr_matrix((vals, bcol, bptr), shape=(n * d, n * d))
    P = sp.bsr_matrix((pvals, pcol, pptr), shape=(n * d, n * d))
    x = rng.standard_normal(n * d)
    np.testing.assert_allclose(P @ x, A @ x, rtol=1e-13, atol=1e-13)


def test_pad_pattern_aligned_rows_unchanged():
    bptr = np.array([0, 4, 8, 12], dtype=np.int32)
    bcol = np.array([0, 1, 2, 3] * 3, dtype=np.int32)
    pptr, pcol, src = pad_pattern(bptr, bcol, 4)
    assert np.array_equal(pptr, bptr) and np.array_equal(pcol, bcol)
    assert np.array_equal(src, np.arange(12))
