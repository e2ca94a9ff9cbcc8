"""Array-level finite-element system on the device.

``FESystem`` owns one mesh's device data (SoA coordinates, connectivity, materials), its pattern and
node graph, and runs the path: fused assembly -> Dirichlet elimination -> preconditioner -> PCG.  The
reference-shaped ``FiniteElementAnalysis`` (analysis.py) lowers loads and conditions on the host and
drives this class; benchmarks drive it directly with arrays.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _hip, device
from .device import BsrMatrix, Context, DeviceGraph, DeviceMesh, Pattern, PcgSolver


class FESystem:
    """Device-resident mesh + pattern of one element kind.

    kind: 'tri_elastic' | 'tet_elastic' | 'tri_magnetic'; coords (N, dim); conn (E, npe);
    materials: dict of per-element arrays or scalars (E, nu, thickness, rho, mu).
    """

    def __init__(self, kind: str, coords, conn, materials: dict, plane_mode: int = 0,
                 device_index: int = 0, with_gather: bool = False, max_inc: Optional[int] = None,
                 ctx: Optional[Context] = None):
        """ctx: the handle (and stream) the system's work runs on; default: the device's main one."""
        device.require_gpu()
        self.ctx = Context.get(device_index) if ctx is None else ctx
        self.kind = kind
        _, self.npe, self.dof = device.KIND_INFO[kind]
        coords = np.asarray(coords, dtype=np.float64)
        conn = np.asarray(conn, dtype=np.int32)
        self.n_nodes = coords.shape[0]
        self.n_elem = conn.shape[0]
        self.pattern = Pattern(conn, self.dof, self.n_nodes, sym=True, with_gather=with_gather,
                               max_inc=max_inc)
        self.graph = DeviceGraph(self.ctx, self.pattern)
        self.graph.attach_connectivity(conn)
        self.mesh = DeviceMesh(self.ctx, kind, coords, conn, materials, plane_mode)
        self._coords, self._conn = coords, conn
        if with_gather:
            self.gather = self.ctx.tensor(self.pattern.gather)
            self.seg_ptr = self.ctx.tensor(self.pattern.seg_ptr)
        else:
            self.gather = self.seg_ptr = None

    @property
    def n(self):
        return self.n_nodes * self.dof

    def set_materials(self, materials: dict):
        """Replace the per-element material arrays (same mesh and pattern)."""
        self.mesh = DeviceMesh(self.ctx, self.kind, self._coords, self._conn, materials,
                               self.mesh.plane_mode)

    # ---------------------------------------------------------------------------------------------
    def set_schedule(self, max_inc: int, threads: int = 256, producers: int = 0):
        """Re-tile the fused assembly: row-block schedule with at most ``max_inc`` node-element
        incidences per workgroup, ``threads`` lanes per workgroup (64, 128, 256, 512 or 1024) and,
        when ``producers`` > 0, the wave-specialised kernel with that many producer wavefronts
        (1024 threads: 4, 6 or 8; 512 threads: 2 or 4)."""
        blk, biggest = Pattern.schedule_for(self.pattern.inc_ptr, max_inc)
        self.graph.set_schedule(blk, biggest, threads, producers)

    def tune_assembly(self, candidates=None, reps: int = 5, layout: int = _hip.FE_LAYOUT_PLANES):
        """Autotuner of the fused assembly: time every (threads, max_inc, producers) candidate with
        HIP events on the device stream and keep the fastest.  Returns (best, {candidate: ms})."""
        if candidates is None:
            base = int(np.diff(self.pattern.inc_ptr).max()) if self.n_nodes else 1
            # one-range-per-workgroup and persistent forms, and the wave-specialised form (a
            # schedule whose element or node lists outgrow its producers, or whose slots outgrow
            # the LDS, is refused by the launch and skipped)
            cands = [(256, 128, 0), (256, 192, 0), (256, 256, 0), (512, 256, 0), (512, 384, 0),
                     (512, 512, 0), (1024, 512, 0), (1024, 768, 0), (1024, 1024, 0), (1024, 1536, 0),
                     (512, 128, 2), (512, 192, 4), (512, 256, 4), (1024, 384, 6), (1024, 384, 8),
                     (1024, 512, 8)]
            candidates = [c for c in cands if c[1] >= base]
        out = self.stiffness(layout)
        times = {}
        for cand in candidates:
            threads, mi, prods = (tuple(cand) + (0,))[:3]
            try:
                self.set_schedule(mi, threads, prods)
                self.stiffness(layout, out=out)
            except (ValueError, RuntimeError):
                continue
            start = torch.cuda.Event(enable_timing=True)
            stop = torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(reps):
                self.stiffness(layout, out=out)
            stop.record()
            stop.synchronize()
            times[(threads, mi, prods)] = start.elapsed_time(stop) / reps
        best = min(times, key=times.get)
        self.set_schedule(best[1], best[0], best[2])
        return best, times

    def stiffness(self, layout: int = _hip.FE_LAYOUT_BSR, out=None) -> torch.Tensor:
        return device.assemble_fused(self.mesh, self.graph, 'stiffness', layout, out)

    def mass(self, layout: int = _hip.FE_LAYOUT_BSR, out=None) -> torch.Tensor:
        return device.assemble_fused(self.mesh, self.graph, 'mass', layout, out)

    def stiffness_bsr(self, layout: int = _hip.FE_LAYOUT_PLANES) -> BsrMatrix:
        return BsrMatrix(self.graph, self.stiffness(layout), self.dof, layout)

    def mass_bsr(self, layout: int = _hip.FE_LAYOUT_PLANES) -> BsrMatrix:
        return BsrMatrix(self.graph, self.mass(layout), self.dof, layout)

    def element_matrices(self, what: str = 'stiffness', sym: bool = False) -> torch.Tensor:
        return device.element_matrices(self.mesh, what, sym)

    def element_fields(self, u):
        """Per-element strain, stress and energy of the solution u (magnetic: the field B)."""
        u_t = u if isinstance(u, torch.Tensor) else self.ctx.tensor(np.asarray(u, dtype=np.float64))
        return device.element_fields(self.mesh, u_t)

    def assemble_two_phase(self, what: str = 'stiffness') -> torch.Tensor:
        """Element kernel (symmetric storage) + sorted segmented reduction -> CSR values."""
        if self.gather is None:
            raise ValueError('FESystem was built without the two-phase contribution map '
                             '(with_gather=False)')
        ke = self.element_matrices(what, sym=True)
        return device.assemble_gather(self.ctx, ke, self.gather, self.seg_ptr, self.pattern.nnz)

    # ---------------------------------------------------------------------------------------------
    def dirichlet(self, fixed_dofs: np.ndarray, values: Optional[np.ndarray] = None):
        """Device mask and value vector for fixed DOFs (indices into the interleaved DOF vector)."""
        mask = np.zeros(self.n, dtype=np.uint8)
        g = np.zeros(self.n, dtype=np.float64)
        fixed_dofs = np.asarray(fixed_dofs, dtype=np.int64)
        mask[fixed_dofs] = 1
        if values is not None:
            g[fixed_dofs] = np.asarray(values, dtype=np.float64)
        return self.ctx.tensor(mask), self.ctx.tensor(g)

    def prepare(self, K: BsrMatrix, fixed: torch.Tensor, g: torch.Tensor, f: torch.Tensor,
                coarse=None):
        """Dirichlet elimination: returns (K_D, b) with K_D SPD on the free DOFs.  ``coarse``: the
        coarse space (``coarse_space``) of a two-level solve of K_D; its operator and inverse are then
        queued on the side stream from K itself, to run beside the elimination, and handed to the
        solver built for K_D."""
        ready = torch.cuda.Event()
        ready.record(self.ctx.stream)  # K is assembled
        kg = device.spmv(K, g)
        b = device.dirichlet_rhs(self.ctx, fixed, g, f, kg)
        KD = device.apply_dirichlet(K, fixed)
        if coarse is not None:
            KD.coarse_inverse = coarse.start_inverse(K, ready)
        return KD, b

    def coarse_space(self, fixed_dofs, target: Optional[int] = None, node_offsets=None,
                     max_block: Optional[int] = None, graph=None):
        """Device coarse space of aggregated rigid-body modes for these constrained DOFs (cached).
        With node_offsets (the meshes of a batch) and max_block, a block-diagonal space with one coarse
        block of at most max_block DOF per mesh, for the batched solver.  ``graph``: the node graph
        of the matrix it will serve when that is not the mesh's own (a folded pattern)."""
        from .coarse import CoarseSpace
        from .device import DeviceCoarse
        fixed_dofs = np.asarray(fixed_dofs, dtype=np.int64)
        offs = None if node_offsets is None else np.asarray(node_offsets, dtype=np.int64)
        pat = self.pattern if graph is None else graph.pattern
        key = (fixed_dofs.tobytes(), target, None if offs is None else offs.tobytes(), max_block,
               id(pat))
        cache = getattr(self, '_coarse_cache', None)
        if cache is not None and cache[0] == key:
            return cache[1]
        cs = CoarseSpace(self._coords, pat.bptr, pat.bcol, self.dof, fixed_dofs, target,
                         node_offsets=offs, max_block=max_block)
        dc = DeviceCoarse(self.ctx, cs)
        self._coarse_cache = (key, dc)
        return dc

    def solver(self, KD: BsrMatrix, precond: str = 'jacobi', precision: str = 'fp64',
               M64: Optional[torch.Tensor] = None, check_every: int = 16,
               ru_delta: float = 0.1, segments=None, graph: bool = False,
               coarse=None) -> PcgSolver:
        """PCG for the Dirichlet-applied KD.  precond 'jacobi', 'block_jacobi' or 'two_level'
        (block Jacobi -- or the learned blocks M64 -- plus the coarse correction of ``coarse``, a
        ``DeviceCoarse`` from ``coarse_space``).  M64 may be a callable returning the blocks: it is
        called after the coarse inverse has been started on the side stream, so that both run
        together."""
        if precision not in ('fp64', 'mixed'):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if precond not in ('jacobi', 'block_jacobi', 'two_level'):
            raise ValueError("precond must be 'jacobi', 'block_jacobi' or 'two_level' (or a learned "
                             "block preconditioner passed as M64)")
        if precond == 'two_level' and coarse is None:
            raise ValueError("precond='two_level' needs coarse=FESystem.coarse_space(fixed_dofs)")
        return PcgSolver(KD, precond='block_jacobi' if precond == 'two_level' else precond,
                         mixed=(precision == 'mixed'), M64=M64, check_every=check_every,
                         ru_delta=ru_delta, segments=segments, graph=graph,
                         coarse=coarse if precond == 'two_level' else None)

    def solve(self, f, fixed_dofs, values=None, tol: float = 1e-8, precond: str = 'jacobi',
              precision: str = 'fp64', max_iter: Optional[int] = None):
        """Assemble K, eliminate Dirichlet DOFs and solve K u = f; returns (u, K, info)."""
        K = self.stiffness_bsr()
        fixed, g = self.dirichlet(fixed_dofs, values)
        f_t = f if isinstance(f, torch.Tensor) else self.ctx.tensor(np.asarray(f, dtype=np.float64))
        KD, b = self.prepare(K, fixed, g, f_t)
        solver = self.solver(KD, precond, precision)
        u, info = solver.solve(b, tol=tol, max_iter=max_iter)
        return u, K, info
