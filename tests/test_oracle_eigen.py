"""Pins of the eigenvalue oracle (oracle/eigen.py) on small meshes: the sparse eigsh calls against
dense generalized eigh (brute force), and mechanical facts of the unconstrained pencil (six zero
eigenvalues = rigid-body modes for a free 3D body; mass-orthonormal eigenvectors)."""

import numpy as np
import pytest

from finite_elements_amd import meshgen
from oracle import eigen, vectorized as vz


def _pencil(n=3, seed=5000):
    pr = meshgen.cube_modal(n=n, seed=seed)
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, what='stiffness')
    M = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, what='mass')
    return pr, K, M


def test_smallest_clamped_matches_dense():
    pr, K, M = _pencil(3)
    vals, vecs = eigen.smallest_clamped(K, M, pr.fixed_dofs, 8)
    dense = eigen.dense_reduced(K, M, pr.fixed_dofs)
    np.testing.assert_allclose(vals, dense[:8], rtol=1e-10)
    # M-orthonormal, K-diagonal
    Mv = M @ vecs
    np.testing.assert_allclose(vecs.T @ Mv, np.eye(8), atol=1e-9)
    np.testing.assert_allclose(vecs.T @ (K @ vecs), np.diag(vals), rtol=1e-9, atol=1e-6 * vals.max())


def test_largest_matches_dense_and_rigid_modes():
    pr, K, M = _pencil(2)
    vals, _ = eigen.largest(K, M, 6)
    dense = eigen.dense_reduced(K, M)
    np.testing.assert_allclose(vals, dense[-6:], rtol=1e-10)
    # an unconstrained solid has exactly six rigid-body (zero-energy) modes
    scale = dense[-1]
    assert np.all(np.abs(dense[:6]) < 1e-9 * scale)
    assert dense[6] > 1e-6 * scale
