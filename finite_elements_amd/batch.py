"""Many independent meshes (design variants, e.g. uneven plates) assembled and solved together.

The meshes are concatenated into one block-diagonal system (disjoint node ranges, one segment per
mesh).  One fused-assembly launch builds every stiffness matrix (row-block schedule picked by the
autotuner of ``FESystem.tune_assembly``), and one batched PCG (``fe_pcg_batched``) solves all of them:
each segment iterates with its own scalars, stops on its own true FP64 residual, and the chunks of
iterations can be replayed from a captured HIP graph.  Each mesh's result is what the reference would
compute for that mesh alone with ``FiniteElementAnalysis.solve`` (finite_elements/analysis.py:798-830),
on the eliminated system.
"""

from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np
import torch

from .meshgen import Problem
from .system import FESystem


def concatenate(problems: Sequence[Problem]):
    """(coords, conn, materials, fixed_dofs, fixed_values, f, node_offsets) of the combined system."""
    if not problems:
        raise ValueError('no problems to batch')
    kind = problems[0].kind
    plane_mode = problems[0].plane_mode
    for pr in problems:
        if pr.kind != kind or pr.plane_mode != plane_mode:
            raise ValueError('a batch needs one element kind and one plane mode')
    d = problems[0].dof
    nodes = np.array([pr.coords.shape[0] for pr in problems], dtype=np.int64)
    offs = np.zeros(len(problems) + 1, dtype=np.int64)
    offs[1:] = np.cumsum(nodes)
    coords = np.concatenate([pr.coords for pr in problems])
    conn = np.concatenate([pr.conn.astype(np.int64) + o for pr, o in zip(problems, offs[:-1])])
    keys = problems[0].materials.keys()
    mats = {k: np.concatenate([np.asarray(pr.materials[k], dtype=np.float64) for pr in problems])
            for k in keys}
    fixed = np.concatenate([np.asarray(pr.fixed_dofs, dtype=np.int64) + d * o
                            for pr, o in zip(problems, offs[:-1])])
    vals = np.concatenate([np.zeros(len(pr.fixed_dofs)) if pr.fixed_values is None
                           else np.asarray(pr.fixed_values, dtype=np.float64) for pr in problems])
    f = np.concatenate([np.zeros(pr.coords.shape[0] * d) if pr.f is None else np.asarray(pr.f)
                        for pr in problems])
    return kind, plane_mode, coords, conn.astype(np.int32), mats, fixed, vals, f, offs


class MeshBatch:
    """A batch of independent meshes of one element kind on one GPU."""

    def __init__(self, problems: Sequence[Problem], device_index: int = 0, tune: bool = True,
                 ctx=None):
        """ctx: the handle and stream the batch runs on (default: the device's main one)."""
        (kind, plane_mode, coords, conn, mats, fixed, vals, f, offs) = concatenate(problems)
        self.problems = list(problems)
        self.node_offsets = offs
        self.sysm = FESystem(kind, coords, conn, mats, plane_mode=plane_mode, device_index=device_index,
                             ctx=ctx)
        self.dof = self.sysm.dof
        self.fixed_dofs = fixed
        self.fixed_values = vals
        self.f = f
        self.tune_times = None
        self.schedule = None  # (threads, max_inc, producers) picked by the autotuner
        if tune:
            best, self.tune_times = self.sysm.tune_assembly()
            self.schedule = best
        self._fixed = None
        self._solver = None
        self.check_every = 16  # iterations per chunk (one status poll, one graph replay)

    @property
    def n_meshes(self) -> int:
        return len(self.problems)

    def assemble(self):
        """Stiffness of every mesh (one fused launch) with the Dirichlet conditions eliminated."""
        s = self.sysm
        if self._fixed is None:
            self._fixed = s.dirichlet(self.fixed_dofs, self.fixed_values)
            self._f = s.ctx.tensor(self.f)
        fixed, g = self._fixed
        K = s.stiffness_bsr()
        KD, b = s.prepare(K, fixed, g, self._f)
        return K, KD, b

    def coarse_space(self, max_block: int = 120):
        """Block-diagonal coarse space of the batch: aggregated rigid-body modes, one coarse block of
        at most max_block DOF per mesh (built once; the batched two-level solver's)."""
        return self.sysm.coarse_space(self.fixed_dofs, node_offsets=self.node_offsets, max_block=max_block)

    def solve(self, tol: float = 1e-8, precision: str = 'fp64', precond: str = 'block_jacobi',
              graph: bool = True, max_iter: Optional[int] = None, check_every: Optional[int] = None):
        """Assemble and solve every mesh; returns (u on the device, info with per-mesh arrays).

        check_every: iterations per chunk (default: ``self.check_every``, set by ``tune_solver``).

        precond 'block_jacobi', 'jacobi' or 'two_level' (block Jacobi plus each mesh's own coarse
        correction over its aggregated rigid-body modes).  precision 'mixed' iterates in FP32 inside
        FP64 reliable updates and finishes every mesh whose FP32 iteration fails (with block Jacobi:
        long thin plates, whose condition number times the FP32 unit roundoff is near one) with FP64
        PCG; 'fp64' iterates in FP64 throughout."""
        K, KD, b = self.assemble()
        coarse = self.coarse_space() if precond == 'two_level' else None
        ce = self.check_every if check_every is None else int(check_every)
        solver = self.sysm.solver(KD, precond, precision, segments=self.node_offsets, graph=graph,
                                  check_every=ce, coarse=coarse)
        u, info = solver.solve(b, tol=tol, max_iter=max_iter)
        return u, info

    def tune_solver(self, candidates=(8, 16, 32), tol: float = 1e-8, precision: str = 'fp64',
                    precond: str = 'two_level', graph: bool = True, reps: int = 2,
                    max_iter: Optional[int] = None):
        """Autotune the batched solve's chunk schedule: time the whole solve (assembly, set-up and
        iterations) for every chunk length -- iterations per launch group, between two status polls
        and one HIP-graph replay; longer chunks poll less, shorter ones stop closer to the last
        plate's convergence -- after one warm-up each, keep the fastest in ``self.check_every``.
        Every timed solve must converge.  Returns (best, {check_every: seconds})."""
        import time
        times = {}
        for ce in candidates:
            self.solve(tol=tol, precision=precision, precond=precond, graph=graph, max_iter=max_iter,
                       check_every=ce)
            best = None
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _, info = self.solve(tol=tol, precision=precision, precond=precond, graph=graph,
                                     max_iter=max_iter, check_every=ce)
                torch.cuda.synchronize()
                el = time.perf_counter() - t0
                if info['status'] != 0:
                    raise RuntimeError(f'tune_solver: the solve with check_every={ce} did not converge')
                best = el if best is None else min(best, el)
            times[int(ce)] = best
        self.check_every = min(times, key=times.get)
        return self.check_every, times

    def split(self, u: torch.Tensor) -> List[np.ndarray]:
        """Per-mesh solution vectors (host) from the batched solution."""
        host = u.detach().cpu().numpy()
        d = self.dof
        return [host[d * a:d * b] for a, b in zip(self.node_offsets[:-1], self.node_offsets[1:])]


def aspect_ratio(pr: Problem) -> float:
    """Largest over smallest extent of a mesh's bounding box (over the dimensions it spans)."""
    ext = np.ptp(np.asarray(pr.coords, dtype=np.float64), axis=0)
    ext = ext[ext > 0]
    return float(ext.max() / ext.min()) if ext.size else 1.0


class SplitBatch:
    """A batch split by an a-priori precision class: meshes whose FP32 iteration is expected to
    hold (bounding-box aspect ratio below ``fp64_aspect``) iterate in FP32 inside FP64 reliable
    updates, the long thin ones in FP64.  The two classes are two ``MeshBatch``es on two streams of
    their own, solved by two concurrent calls (one host thread each; ctypes releases the GIL), so
    the FP64 class's iterations run beside the FP32 class's.  On the 4096 plates (threshold 4: 2,063
    mixed, 2,033 FP64) no mixed plate falls back and the pair takes 249 ms against 295 ms for all
    plates in FP64 in one batch.  Results are per mesh in the order given."""

    def __init__(self, problems: Sequence[Problem], fp64_aspect: float = 4.0, device_index: int = 0,
                 tune: bool = True, schedule=None):
        """schedule: (threads, max_inc[, producers]) of the fused assembly to use in both classes
        instead of autotuning each (e.g. the schedule a MeshBatch of all the meshes picked)."""
        from .device import Context
        self.problems = list(problems)
        asp = np.array([aspect_ratio(pr) for pr in self.problems])
        self.index = {'mixed': np.flatnonzero(asp < fp64_aspect), 'fp64': np.flatnonzero(asp >= fp64_aspect)}
        main = Context.get(device_index)
        self.batches = {}
        for k, (name, idx) in enumerate(self.index.items()):
            if idx.size:
                ctx = main.side(10 + k)
                # torch's work (allocations, the autotuner's events) on the batch's own stream
                with torch.cuda.stream(ctx.stream):
                    mb = MeshBatch([self.problems[i] for i in idx], device_index,
                                   tune and schedule is None, ctx=ctx)
                    if schedule is not None:
                        threads, max_inc, prods = (tuple(schedule) + (0,))[:3]
                        mb.sysm.set_schedule(int(max_inc), int(threads), int(prods))
                        mb.schedule = (int(threads), int(max_inc), int(prods))
                    self.batches[name] = mb
                ctx.stream.synchronize()

    @property
    def n_meshes(self) -> int:
        return len(self.problems)

    def coarse_space(self, max_block: int = 120):
        """Each class's block-diagonal coarse space (built once)."""
        for mb in self.batches.values():
            with torch.cuda.stream(mb.sysm.ctx.stream):
                mb.coarse_space(max_block)
            mb.sysm.ctx.stream.synchronize()

    def solve(self, tol: float = 1e-8, precond: str = 'two_level', graph: bool = True,
              max_iter: Optional[int] = None):
        """Solve both classes concurrently; returns ({class: u}, info) with info['segments'] holding
        per-mesh arrays in the given order and info['precision'] each mesh's class."""
        import threading
        torch.cuda.synchronize()
        results, errors = {}, []

        def run(name, mb):
            try:
                with torch.cuda.stream(mb.sysm.ctx.stream):
                    results[name] = mb.solve(tol=tol, precision=name, precond=precond, graph=graph,
                                             max_iter=max_iter)
                    mb.sysm.ctx.stream.synchronize()
            except Exception as exc:  # re-raised after both calls returned
                errors.append(exc)

        items = list(self.batches.items())
        threads = [threading.Thread(target=run, args=it) for it in items[1:]]
        for th in threads:
            th.start()
        run(*items[0])
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        n = len(self.problems)
        seg = {}
        prec = np.empty(n, dtype=object)
        for name, (_, info) in results.items():
            idx = self.index[name]
            prec[idx] = name
            for key, arr in info['segments'].items():
                if key not in seg:
                    seg[key] = np.zeros(n, dtype=arr.dtype)
                seg[key][idx] = arr
        status = 0
        for _, info in results.values():
            if info['status'] != 0:
                status = info['status']
        out = {'segments': seg, 'precision': prec, 'status': status,
               'fallbacks': int(sum(info['fallbacks'] for _, info in results.values())),
               'iterations': int(seg['iterations'].max())}
        return {name: u for name, (u, _) in results.items()}, out

    def split(self, u) -> List[np.ndarray]:
        """Per-mesh solution vectors (host), in the given order, from ``solve``'s first result."""
        out = [None] * len(self.problems)
        for name, mb in self.batches.items():
            for i, part in zip(self.index[name], mb.split(u[name])):
                out[i] = part
        return out

