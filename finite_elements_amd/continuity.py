"""Continuity conditions u(node1) = value * u(node2) by elimination (finite_elements/conditions.py:196-241,
analysis.py:175-199).

The reference adds one Lagrange row per condition to its indefinite system: row R holds
``u1 - value * u2 = 0`` and the multiplier enters the equations of node1 (+1) and node2 (-value).
Here each condition eliminates node1 (the *slave*) in favour of node2 (its *master*): with
u = T w, where T copies w at every node that is not a slave and sets u[s] = c(s) w[rep(s)] at a
slave (rep(s) the end of its chain of masters, c(s) the product of the values along it), the
system solved is T^T K T w = T^T f -- symmetric positive definite on the same node numbering,
with the slave rows made identity rows (the slaves enter the Dirichlet mask at value 0).

``ContinuityFold`` builds, on the host and once per pattern and condition set:

* the folded node graph: a non-slave row a holds the masters of every column coupled to any node
  that folds into a; a slave row holds only its diagonal block;
* the fold map for ``fe_bsr_fold``: per folded block, the source blocks and their coefficients
  c(i) c(j), sorted by source block, so the device sums in a fixed order.

After the solve, ``expand`` sets the slaves and ``multipliers`` recovers the reference's multipliers
from the residual r = f - K u of the unfolded system: taking conditions leaves-first,
lambda_i = r[s_i] + sum of value_j lambda_j over the conditions j whose master is s_i, and each
Dirichlet row p gets r[p] + sum of value_j lambda_j over the conditions whose master is p.

A condition applies to every component of its nodes (the reference's conditions are scalar, for
the magnetic vector potential).  Rejected: a node that is the slave of two conditions, a cycle of
conditions, and a slave that also carries a Dirichlet condition (over-constrained).
"""

from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _hip
from .device import BsrMatrix, PaddedGraph, _vp


def resolve_chains(n_nodes: int, slaves, masters, values):
    """rep (n,) int64 and c (n,) float64: every node's root master and coefficient (identity for
    nodes that are no slave).  Raises ValueError on a repeated slave or a cycle."""
    slaves = np.asarray(slaves, dtype=np.int64)
    masters = np.asarray(masters, dtype=np.int64)
    values = np.asarray(values, dtype=np.float64)
    if slaves.size == 0:
        return np.arange(n_nodes, dtype=np.int64), np.ones(n_nodes)
    if slaves.min() < 0 or masters.min() < 0 or max(slaves.max(), masters.max()) >= n_nodes:
        raise ValueError('continuity condition node index out of range')
    if np.unique(slaves).size != slaves.size:
        raise ValueError('a node is the first node of two continuity conditions')
    if np.any(slaves == masters):
        raise ValueError('a continuity condition links a node to itself')
    parent = np.arange(n_nodes, dtype=np.int64)
    coef = np.ones(n_nodes)
    parent[slaves] = masters
    coef[slaves] = values
    rep = parent.copy()
    c = coef.copy()
    # pointer doubling: after k rounds every chain of length < 2^k is resolved
    for _ in range(int(np.ceil(np.log2(max(slaves.size, 1) + 1))) + 1):
        nxt = rep[rep]
        if np.array_equal(nxt, rep):
            break
        c = c * c[rep]
        rep = nxt
    if not np.array_equal(rep[rep], rep):
        raise ValueError('the continuity conditions form a cycle')
    return rep, c


class FoldedGraph:
    """Device node graph of a folded pattern (what the solver, SpMV and Dirichlet kernels read)."""

    def __init__(self, ctx, bptr: np.ndarray, bcol: np.ndarray, dof: int):
        self.ctx = ctx
        self.dof = dof
        self.n_nodes = int(bptr.size - 1)
        self.nnzb = int(bptr[-1])
        self.pattern = _HostPattern(bptr, bcol)
        self.bptr = ctx.tensor(bptr)
        self.bcol = ctx.tensor(bcol)
        self._padded = None

    def padded(self) -> PaddedGraph:
        if self._padded is None:
            self._padded = PaddedGraph(self)
        return self._padded


class _HostPattern:
    def __init__(self, bptr, bcol):
        self.bptr = bptr
        self.bcol = bcol


class ContinuityFold:
    """Elimination of the continuity conditions (slaves, masters, values) on a node graph."""

    def __init__(self, bptr, bcol, n_nodes: int, dof: int, slaves: Sequence[int],
                 masters: Sequence[int], values: Sequence[float]):
        bptr = np.asarray(bptr, dtype=np.int64)
        bcol = np.asarray(bcol, dtype=np.int64)
        self.n_nodes = int(n_nodes)
        self.dof = int(dof)
        self.slaves = np.asarray(slaves, dtype=np.int64)
        self.masters = np.asarray(masters, dtype=np.int64)
        self.values = np.asarray(values, dtype=np.float64)
        self.rep, self.c = resolve_chains(self.n_nodes, self.slaves, self.masters, self.values)
        n = self.n_nodes
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(bptr))
        a, b = self.rep[rows], self.rep[bcol]
        key = a * n + b
        perm = np.argsort(key, kind='stable')  # by folded block, then by source block
        skey = key[perm]
        diag = self.slaves * n + self.slaves  # a slave keeps its diagonal block (no contributions)
        uniq = np.unique(np.concatenate([skey, diag]))
        self.bptr = np.zeros(n + 1, dtype=np.int32)
        self.bptr[1:] = np.cumsum(np.bincount(uniq // n, minlength=n))
        self.bcol = (uniq % n).astype(np.int32)
        self.nnzb = int(uniq.size)
        self.seg_ptr = np.append(np.searchsorted(skey, uniq, side='left'), skey.size).astype(np.int32)
        self.src_block = perm.astype(np.int32)
        self.coef = (self.c[rows] * self.c[bcol])[perm]
        self._dev = None

    # ---- DOF-level maps ----------------------------------------------------------------------
    def slave_dofs(self) -> np.ndarray:
        d = self.dof
        return (self.slaves[:, None] * d + np.arange(d)[None, :]).reshape(-1)

    def restrict(self, f: np.ndarray) -> np.ndarray:
        """T^T f (slave entries 0), summed in condition order."""
        d = self.dof
        f = np.asarray(f, dtype=np.float64).reshape(self.n_nodes, d)
        out = f.copy()
        s = self.slaves
        np.add.at(out, self.rep[s], self.c[s][:, None] * f[s])
        out[s] = 0.0
        return out.reshape(-1)

    def expand(self, u: torch.Tensor) -> torch.Tensor:
        """u with every slave set to c(s) u[rep(s)] (device gather)."""
        d = self.dof
        if self.slaves.size == 0:
            return u
        dev = self._device(u.device)
        v = u.view(self.n_nodes, d).clone()
        v[dev['s']] = dev['cs'][:, None] * v[dev['rs']]
        return v.view(-1)

    def multipliers(self, r: np.ndarray, bc_dofs: np.ndarray):
        """(lambda of every condition in the given order, lambda of every Dirichlet DOF) from the
        unfolded residual r = f - K u.  Node-level conditions give one multiplier per component
        (d per condition, condition-major)."""
        d = self.dof
        r = np.asarray(r, dtype=np.float64).reshape(self.n_nodes, d)
        add = np.zeros_like(r)
        lam = np.zeros((self.slaves.size, d))
        for i in self._leaves_first():
            s = self.slaves[i]
            lam[i] = r[s] + add[s]
            add[self.masters[i]] += self.values[i] * lam[i]
        flat_r, flat_add = r.reshape(-1), add.reshape(-1)
        bc = np.asarray(bc_dofs, dtype=np.int64)
        return lam.reshape(-1), flat_r[bc] + flat_add[bc]

    def _leaves_first(self):
        """Condition indices ordered so that every condition comes after those whose master is its
        slave."""
        k = self.slaves.size
        by_master = {}
        for j in range(k):
            by_master.setdefault(int(self.masters[j]), []).append(j)
        pending = np.array([len(by_master.get(int(s), ())) for s in self.slaves], dtype=np.int64)
        slave_cond = {int(s): i for i, s in enumerate(self.slaves)}
        ready = [i for i in range(k) if pending[i] == 0]
        order = []
        while ready:
            i = ready.pop()
            order.append(i)
            parent = slave_cond.get(int(self.masters[i]))
            if parent is not None:
                pending[parent] -= 1
                if pending[parent] == 0:
                    ready.append(parent)
        return order

    # ---- device ------------------------------------------------------------------------------
    def _device(self, device):
        if self._dev is None:
            def t(a, dt):
                return torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(device)
            self._dev = {'s': t(self.slaves, torch.int64), 'rs': t(self.rep[self.slaves], torch.int64),
                         'cs': t(self.c[self.slaves], torch.float64),
                         'seg': t(self.seg_ptr, torch.int32), 'src': t(self.src_block, torch.int32),
                         'coef': t(self.coef, torch.float64)}
        return self._dev

    def graph(self, ctx) -> FoldedGraph:
        g = getattr(self, '_graph', None)
        if g is None or g.ctx is not ctx:
            g = FoldedGraph(ctx, self.bptr, self.bcol, self.dof)
            self._graph = g
        return g

    def fold(self, K: BsrMatrix) -> BsrMatrix:
        """T^T K T on the folded graph (FP64, K's layout), by ``fe_bsr_fold``."""
        ctx = K.graph.ctx
        if K.precision != 0:
            raise ValueError('fold an FP64 matrix')
        if K.graph.n_nodes != self.n_nodes or K.dof != self.dof:
            raise ValueError('matrix and fold describe different systems')
        g = self.graph(ctx)
        out = BsrMatrix(g, torch.empty(g.nnzb * self.dof * self.dof, dtype=torch.float64,
                                       device=K.vals.device), self.dof, K.layout)
        dev = self._device(K.vals.device)
        a, b = K.struct(), out.struct()
        _hip.call('fe_bsr_fold', ctx.h, ctypes.byref(a), ctypes.byref(b), _vp(dev['seg']),
                  _vp(dev['src']), _vp(dev['coef']))
        return out
