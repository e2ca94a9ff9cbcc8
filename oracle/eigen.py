"""Eigenvalue oracle for the modal path (test infrastructure only: imported by tests/ alone).

* ``largest`` restates finite_elements/analysis.py:775-784: ``scipy.sparse.linalg.eigsh(A=K, M=M,
  which='LM', k=k)`` on the unconstrained stiffness and mass matrices (ascending output).
* ``smallest_clamped`` is the well-posed form of the reference's ``'smallest'`` branch
  (analysis.py:788-794 inverts an unconstrained, singular K): the k lowest eigenpairs of the pencil
  restricted to the free DOFs, ``eigsh(K_ff, k, M_ff, sigma=0)`` (shift-invert about 0).
* ``dense_reduced`` is the brute-force check of both on small systems (``scipy.linalg.eigh``).
"""

import numpy as np
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg


def largest(K, M, k):
    vals, vecs = scipy.sparse.linalg.eigsh(A=sp.csc_matrix(K), M=sp.csc_matrix(M), which='LM', k=k)
    return vals, vecs


def free_dofs(n, fixed_dofs):
    mask = np.ones(n, dtype=bool)
    mask[np.asarray(fixed_dofs, dtype=np.int64)] = False
    return np.nonzero(mask)[0]


def smallest_clamped(K, M, fixed_dofs, k):
    free = free_dofs(K.shape[0], fixed_dofs)
    Kf = sp.csc_matrix(K)[free][:, free]
    Mf = sp.csc_matrix(M)[free][:, free]
    vals, vecs = scipy.sparse.linalg.eigsh(Kf, k=k, M=Mf, sigma=0, which='LM')
    order = np.argsort(vals)
    full = np.zeros((K.shape[0], k))
    full[free] = vecs[:, order]
    return vals[order], full


def dense_reduced(K, M, fixed_dofs=()):
    free = free_dofs(K.shape[0], fixed_dofs)
    Kd = np.asarray(sp.csr_matrix(K)[free][:, free].todense())
    Md = np.asarray(sp.csr_matrix(M)[free][:, free].todense())
    return scipy.linalg.eigh(Kd, Md, eigvals_only=True)
