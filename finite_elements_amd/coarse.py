"""Coarse space of the two-level preconditioner: aggregated rigid-body modes.

The nodes are grouped into aggregates of about ``target`` nodes by spatial bins. The bins follow the
mesh's own coordinates, so any mesh works; aggregates need not be connected. On every aggregate the
rigid-body modes are evaluated at its nodes:

* 3D elasticity: 3 translations and 3 rotations about the aggregate centroid;
* 2D elasticity: 2 translations and 1 rotation;
* scalar problems: the constant.

Rows of constrained DOFs are zeroed, and each aggregate's modes are orthonormalised by an nc x nc
transform T (from the eigen-decomposition of the modes' Gram matrix, rank revealed): P = Z T with
one d x nc block per node. Because Z at a node is a fixed function of the node's centred and scaled
coordinates x^, the solver evaluates P on the fly from x^ and T instead of reading it: P^T r sums the
rigid-body moments of r over an aggregate and applies T^T; P z applies T to z and evaluates the
rigid-body motion at the node. The coarse
operator A_c = P^T A P is assembled on the GPU per solve (``fe_coarse_operator``) from the block
pairs grouped here by coarse pair. The preconditioner is M = M_bj + P A_c^-1 P^T, block Jacobi plus
the coarse correction: SPD, fixed, and cheap to apply (``fe_pcg_coarse``).

Everything here depends on the mesh and the constrained DOFs only, not on the matrix values, and is
built once per mesh on the host.
"""

from __future__ import annotations

from typing import Optional

import numpy as np


def rigid_modes(coords: np.ndarray, dof: int) -> np.ndarray:
    """(n_nodes, dof, nc) rigid-body modes at the given (centred) coordinates."""
    n = coords.shape[0]
    if dof == 1:
        return np.ones((n, 1, 1))
    if dof == 2:
        Z = np.zeros((n, 2, 3))
        Z[:, 0, 0] = 1.0
        Z[:, 1, 1] = 1.0
        Z[:, 0, 2] = -coords[:, 1]
        Z[:, 1, 2] = coords[:, 0]
        return Z
    Z = np.zeros((n, 3, 6))
    Z[:, 0, 0] = Z[:, 1, 1] = Z[:, 2, 2] = 1.0
    x, y, z = coords[:, 0], coords[:, 1], coords[:, 2]
    # rotations about x, y, z: omega x r
    Z[:, 1, 3], Z[:, 2, 3] = -z, y
    Z[:, 0, 4], Z[:, 2, 4] = z, -x
    Z[:, 0, 5], Z[:, 1, 5] = -y, x
    return Z


def spatial_aggregates(coords: np.ndarray, target: int) -> np.ndarray:
    """Aggregate id per node: spatial bins holding about ``target`` nodes each."""
    n, dim = coords.shape
    lo = coords.min(axis=0)
    ext = coords.max(axis=0) - lo
    active = ext > 0
    k = max(1, int(active.sum()))
    # bins per active dimension so that n / prod(bins) ~ target, bins proportional to the extent
    vol = float(np.prod(ext[active])) if active.any() else 1.0
    width = (vol * target / max(n, 1)) ** (1.0 / k)
    bins = np.ones(dim, dtype=np.int64)
    for a in range(dim):
        if active[a]:
            bins[a] = max(1, int(round(ext[a] / width)))
    idx = np.zeros((n, dim), dtype=np.int64)
    for a in range(dim):
        if active[a]:
            idx[:, a] = np.minimum((coords[:, a] - lo[a]) / ext[a] * bins[a], bins[a] - 1).astype(np.int64)
    flat = np.zeros(n, dtype=np.int64)
    for a in reversed(range(dim)):
        flat = flat * bins[a] + idx[:, a]
    # renumber the non-empty bins 0..n_agg-1
    _, agg = np.unique(flat, return_inverse=True)
    return agg.astype(np.int64)


class CoarseSpace:
    """Host description of P: aggregates, the per-node prolongator blocks and the coarse pair map."""

    def __init__(self, coords: np.ndarray, bptr: np.ndarray, bcol: np.ndarray, dof: int,
                 fixed_dofs: Optional[np.ndarray] = None, target: Optional[int] = None,
                 node_offsets: Optional[np.ndarray] = None, max_block: Optional[int] = None):
        """node_offsets: node ranges of independent meshes (aggregates never cross them).  max_block:
        with node_offsets, a block-diagonal space (one coarse block per range, for the batched
        solver) whose blocks have at most max_block coarse DOF; a range's aggregates are enlarged
        beyond ``target`` nodes until its block fits."""
        coords = np.asarray(coords, dtype=np.float64)
        n = coords.shape[0]
        self.dof = dof
        self.nc = {1: 1, 2: 3, 3: 6}[dof]
        if target is None:
            # aggregates of at least 1000 nodes (3D) / 81 (2D), and for one large system a coarse
            # dimension near 1200 (the dense coarse inverse grows with its cube, the iteration
            # count falls with it: 64^3 tet cube 1,296 coarse DOF, 450^2 magnetostatics ~1,200)
            base = 1000 if coords.shape[1] == 3 else 81
            target = base if node_offsets is not None else max(base, -(-n * self.nc // 1200))
        self.target = target
        # aggregates never cross the given node ranges (independent meshes of a batch)
        offs = np.array([0, n]) if node_offsets is None else np.asarray(node_offsets, dtype=np.int64)
        agg = np.empty(n, dtype=np.int64)
        base = 0
        agg_seg = [0]
        for a, b in zip(offs[:-1], offs[1:]):
            if b > a:
                t = target
                loc = spatial_aggregates(coords[a:b], t)
                while max_block is not None and (int(loc.max()) + 1) * self.nc > max_block:
                    t = int(t * 1.25) + 1
                    loc = spatial_aggregates(coords[a:b], t)
                agg[a:b] = loc + base
                base += int(loc.max()) + 1
            agg_seg.append(base)
        self.n_agg = base
        # block-diagonal layout (batched solver): aggregates of range s are [agg_seg[s], agg_seg[s+1])
        if max_block is not None and node_offsets is not None:
            self.agg_seg = np.asarray(agg_seg, dtype=np.int32)
            counts = np.diff(self.agg_seg).astype(np.int64) * self.nc
            self.seg_of_agg = np.repeat(np.arange(counts.size, dtype=np.int32), np.diff(self.agg_seg))
            self.blk_ofs = np.zeros(counts.size + 1, dtype=np.int64)
            self.blk_ofs[1:] = np.cumsum(counts * counts)
            self.max_block = int(counts.max()) if counts.size else 0
        else:
            self.agg_seg = self.seg_of_agg = self.blk_ofs = None
            self.max_block = 0
        self.node_agg = agg.astype(np.int32)
        order = np.argsort(agg, kind='stable')
        self.agg_nodes = order.astype(np.int32)
        self.agg_ptr = np.zeros(self.n_agg + 1, dtype=np.int32)
        self.agg_ptr[1:] = np.cumsum(np.bincount(agg, minlength=self.n_agg))
        fixed = np.zeros(n * dof, dtype=bool)
        if fixed_dofs is not None and len(fixed_dofs):
            fixed[np.asarray(fixed_dofs, dtype=np.int64)] = True
        fixed = fixed.reshape(n, dof)
        # prolongator P = Z(x^) T, orthonormal per aggregate: T = V_kept diag(lambda_kept)^-1/2 from
        # the eigen-decomposition of the masked modes' Gram matrix Z^T Z
        P = np.zeros((n, dof, self.nc))
        xhat = np.zeros_like(coords)
        T = np.zeros((self.n_agg, self.nc, self.nc))
        self.rank = np.zeros(self.n_agg, dtype=np.int32)
        for g in range(self.n_agg):
            nodes = order[self.agg_ptr[g]:self.agg_ptr[g + 1]]
            x = coords[nodes] - coords[nodes].mean(axis=0)
            scale = max(np.abs(x).max(), 1e-300)
            xhat[nodes] = x / scale
            Z = rigid_modes(xhat[nodes], dof)
            Z[fixed[nodes]] = 0.0
            M = Z.reshape(-1, self.nc)
            lam, V = np.linalg.eigh(M.T @ M)
            keep = lam > 1e-20 * max(lam.max(), 1e-300)
            T[g][:, keep] = V[:, keep] / np.sqrt(lam[keep])
            self.rank[g] = int(keep.sum())
            P[nodes] = (M @ T[g]).reshape(len(nodes), dof, self.nc)
        self.P = P
        self.T = T
        self.xhat = xhat
        # per node, bit e set when DOF e is constrained
        self.fixed_bits = (fixed.astype(np.uint8) << np.arange(dof, dtype=np.uint8)[None, :]).sum(
            axis=1).astype(np.uint8)
        self.n_coarse = self.n_agg * self.nc
        # coarse pair map: every block (i, j) of the pattern contributes to pair (agg i, agg j)
        bptr = np.asarray(bptr, dtype=np.int64)
        bcol = np.asarray(bcol, dtype=np.int64)
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(bptr))
        key = agg[rows] * self.n_agg + agg[bcol]
        perm = np.argsort(key, kind='stable')
        keys = key[perm]
        uniq, start = np.unique(keys, return_index=True)
        self.n_pairs = uniq.size
        self.pair_ptr = np.append(start, keys.size).astype(np.int32)
        self.pair_ij = np.stack([uniq // self.n_agg, uniq % self.n_agg], axis=1).astype(np.int32)
        self.pair_q = perm.astype(np.int32)
        self.pair_row = rows[perm].astype(np.int32)
