"""The "strong CPU" baseline of the solve path (CPU baseline only; never on the product path).

The reference's own path is a per-element Python loop and a natural-order SuperLU solve
(finite_elements/analysis.py:324-339, 661, 819-822), which stops fitting in memory long before the
64^3 cube.  This is what a careful CPU implementation of the same eliminated system does instead:

* vectorised numpy element matrices and scipy COO->CSR (oracle.vectorized.assemble_csr);
* Dirichlet elimination by keeping the free rows and columns (K_ff u_f = f_f - K_fc g);
* plain Jacobi-preconditioned CG in FP64 to ||b - K u|| / ||b|| <= tol, with the sparse products
  split into row blocks run by a thread pool (scipy's CSR product releases the GIL) and the dots
  done by numpy's multithreaded BLAS.

It is the CPU counterpart of the GPU's plain FP64 Jacobi-PCG baseline: same algorithm, same
iteration count up to rounding.  Nothing here is parity-pinned: it is timed, not compared.
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import vectorized as vz


class ThreadedCsr:
    """y = K x with K's rows split into ``threads`` contiguous blocks of equal nonzero count."""

    def __init__(self, K, threads):
        K = K.tocsr()
        self.n = K.shape[0]
        nnz_cut = np.linspace(0, K.nnz, threads + 1)
        cuts = np.searchsorted(K.indptr, nnz_cut).clip(0, self.n)
        cuts[0], cuts[-1] = 0, self.n
        self.cuts = np.unique(cuts)
        self.blocks = [K[a:b] for a, b in zip(self.cuts[:-1], self.cuts[1:])]
        self.pool = ThreadPoolExecutor(len(self.blocks))

    def matvec(self, x, y):
        def part(i):
            y[self.cuts[i]:self.cuts[i + 1]] = self.blocks[i] @ x
        list(self.pool.map(part, range(len(self.blocks))))
        return y

    def close(self):
        self.pool.shutdown()


def jacobi_cg(K, b, tol=1e-8, max_iter=100000, threads=1, iter_cap=None, progress=None):
    """Jacobi-preconditioned CG from zero; returns (x, iterations, true relres, seconds per
    iteration).  iter_cap stops early (a bounded timing sample): the relres is then not reached.
    progress(it, relres_estimate) is called every 250 iterations."""
    op = ThreadedCsr(K, threads)
    dinv = 1.0 / K.diagonal()
    n = b.size
    x = np.zeros(n)
    r = b.copy()
    z = dinv * r
    p = z.copy()
    q = np.empty(n)
    rz = float(r @ z)
    bb = float(b @ b)
    it = 0
    cap = max_iter if iter_cap is None else min(max_iter, iter_cap)
    t0 = time.perf_counter()
    while it < cap and float(r @ r) > tol * tol * bb:
        op.matvec(p, q)
        alpha = rz / float(p @ q)
        x += alpha * p
        r -= alpha * q
        np.multiply(dinv, r, out=z)
        rz_new = float(r @ z)
        p *= rz_new / rz
        p += z
        rz = rz_new
        it += 1
        if progress is not None and it % 250 == 0:
            progress(it, float(np.sqrt(float(r @ r) / bb)))
    per_it = (time.perf_counter() - t0) / max(it, 1)
    relres = float(np.linalg.norm(b - op.matvec(x, q)) / np.sqrt(bb))
    op.close()
    return x, it, relres, per_it


def solve_cube(pr, tol=1e-8, threads=1, iter_cap=None, progress=None):
    """Assemble and solve one Problem; returns a dict of timings."""
    t0 = time.perf_counter()
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, pr.plane_mode)
    t1 = time.perf_counter()
    free = np.ones(K.shape[0], dtype=bool)
    free[np.asarray(pr.fixed_dofs, dtype=np.int64)] = False
    Kff = K[free][:, free].tocsr()
    b = np.asarray(pr.f, dtype=np.float64)[free]
    t2 = time.perf_counter()
    _, its, relres, per_it = jacobi_cg(Kff, b, tol=tol, threads=threads, iter_cap=iter_cap,
                                       progress=progress)
    t3 = time.perf_counter()
    return {'dofs': int(K.shape[0]), 'assemble_s': t1 - t0, 'eliminate_s': t2 - t1,
            'cg_s': t3 - t2, 'iterations': its, 'relres': relres, 's_per_iteration': per_it}
