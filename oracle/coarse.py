"""Two-level preconditioner oracle (test infrastructure only): P as a scipy matrix, A_c = P^T A P,
the correction P A_c^-1 P^T r and the full two-level application M r = M_bj r + P A_c^-1 P^T r, all
in FP64 from the host description of the coarse space (finite_elements_amd/coarse.py)."""

import numpy as np
import scipy.sparse as sp


def prolongator(cs):
    n = cs.P.shape[0]
    d, nc = cs.dof, cs.nc
    rows = np.repeat((np.arange(n)[:, None] * d + np.arange(d)[None, :]).reshape(-1), nc)
    cols = np.broadcast_to(cs.node_agg.astype(np.int64)[:, None, None] * nc + np.arange(nc)[None, None, :],
                           cs.P.shape).reshape(-1)
    return sp.csr_matrix((cs.P.reshape(-1), (rows, cols)), shape=(n * d, cs.n_coarse))


def coarse_operator(cs, A):
    """Dense P^T A P with the decoupled unit diagonal on zero coarse columns."""
    P = prolongator(cs)
    Ac = np.asarray((P.T @ A @ P).todense())
    zero = np.abs(P).sum(axis=0).A1 == 0
    Ac[zero, zero] = 1.0
    return Ac, P


def block_jacobi_inverse(A, dof):
    """[n_nodes][d][d] inverses of the diagonal blocks of A (scipy, FP64)."""
    A = A.tocsr()
    n = A.shape[0] // dof
    out = np.zeros((n, dof, dof))
    for i in range(n):
        out[i] = np.linalg.inv(A[dof * i:dof * i + dof, dof * i:dof * i + dof].toarray())
    return out


def apply_two_level(cs, A, binv, r):
    Ac, P = coarse_operator(cs, A)
    d = cs.dof
    u = np.einsum('nij,nj->ni', binv, r.reshape(-1, d)).reshape(-1)
    return u + P @ np.linalg.solve(Ac, P.T @ r)


def pcg(A, b, apply_m, tol=1e-8, maxit=20000):
    """Textbook PCG on the host; returns (x, iterations) at true relres <= tol."""
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    for it in range(1, maxit + 1):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(b - A @ x) <= tol * bn:
            return x, it
        z = apply_m(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxit
