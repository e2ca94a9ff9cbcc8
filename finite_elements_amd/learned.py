"""Learned correction to the block-Jacobi preconditioner (no counterpart in the reference).

A per-node graph network (3 message-passing layers, width 32, FP32 weights) reads 8 scale-free node
features of the assembled stiffness and outputs a positive scaling per DOF; the preconditioner block
of node i is ``diag(s_i) K_ii^-1 diag(s_i)`` -- SPD for any weights, so PCG stays valid, and exactly
block Jacobi when the network outputs zero.  Forward, loss gradient, backward and Adam all run in the
HIP kernels of libfe_hip.so (``fe_gnn_*``); PyTorch only holds the buffers.

Training data are residual/error pairs ``(r_k, e_k)`` that plain block-Jacobi PCG produces on
synthetic lognormal-field cubes: ``x*`` from a tight solve, ``x_k`` after k iterations,
``e_k = x* - x_k``, ``r_k = b - K x_k``.  The loss per pair is the energy fraction an optimally damped
preconditioned Richardson step leaves: ``1 - (r.z)^2 / ((z.Kz)(r.e))`` with ``z = M r``.

Plain PCG is the judge and the fallback: ``solve_with_fallback`` re-solves with block Jacobi when the
learned solve does not converge.
"""

from __future__ import annotations

import ctypes
import json
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _hip, device, meshgen
from .device import BsrMatrix

WIDTH = 32
N_FEATURES = 8
N_LAYERS = 3


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def num_params(dof: int) -> int:
    n = _hip.load().fe_gnn_num_params(dof)
    if n < 0:
        raise ValueError('dof must be 1..3')
    return n


def init_params(dof: int, seed: int = 0) -> np.ndarray:
    """Scaled-uniform initial weights, zero biases and a zero head (the network starts as block Jacobi)."""
    rng = np.random.default_rng(seed)
    out = []
    for layer in range(N_LAYERS):
        fin = N_FEATURES if layer == 0 else WIDTH
        lim = np.sqrt(3.0 / fin)
        out.append(rng.uniform(-lim, lim, WIDTH * fin))  # W1
        out.append(rng.uniform(-lim, lim, WIDTH * fin) * 0.5)  # W2
        out.append(np.zeros(WIDTH))
    out.append(np.zeros(dof * WIDTH))
    out.append(np.zeros(dof))
    p = np.concatenate(out).astype(np.float32)
    assert p.size == num_params(dof)
    return p


class LearnedPreconditioner:
    """Network weights + the device-side forward / training loop."""

    def __init__(self, dof: int = 3, params: Optional[np.ndarray] = None, seed: int = 0,
                 meta: Optional[dict] = None):
        self.dof = dof
        self.params_host = init_params(dof, seed) if params is None else np.asarray(params, np.float32)
        if self.params_host.size != num_params(dof):
            raise ValueError(f'expected {num_params(dof)} parameters, got {self.params_host.size}')
        self.meta = dict(meta or {})
        self._dev = {}

    # ---- persistence ------------------------------------------------------------------------------
    def save(self, path: str):
        self.sync_to_host()
        np.savez(path, params=self.params_host, dof=np.int64(self.dof),
                 meta=np.frombuffer(json.dumps(self.meta).encode(), dtype=np.uint8))

    @classmethod
    def load(cls, path: str) -> 'LearnedPreconditioner':
        z = np.load(path, allow_pickle=False)
        meta = json.loads(bytes(z['meta']).decode()) if 'meta' in z.files else {}
        return cls(int(z['dof']), params=z['params'], meta=meta)

    # ---- device buffers ---------------------------------------------------------------------------
    def params(self, ctx) -> torch.Tensor:
        t = self._dev.get(ctx.device)
        if t is None:
            t = ctx.tensor(self.params_host)
            self._dev[ctx.device] = t
        return t

    def sync_to_host(self):
        for t in self._dev.values():
            self.params_host = t.detach().cpu().numpy().astype(np.float32)
            break

    # ---- forward ----------------------------------------------------------------------------------
    def features(self, KD: BsrMatrix, fixed: torch.Tensor) -> torch.Tensor:
        ctx = KD.graph.ctx
        N = KD.graph.n_nodes
        feat = torch.empty(N * N_FEATURES, dtype=torch.float32, device=ctx.torch_device)
        scratch = torch.empty(N, dtype=torch.float64, device=ctx.torch_device)
        s = KD.struct()
        _hip.call('fe_gnn_features', ctx.h, ctypes.byref(s), _vp(fixed), _vp(feat), _vp(scratch))
        return feat

    def forward(self, KD: BsrMatrix, fixed: torch.Tensor, binv: torch.Tensor,
                feat: Optional[torch.Tensor] = None):
        """Returns (M, o, feat, acts, aggs); M is the learned block preconditioner [N][d][d] (FP64)."""
        ctx = KD.graph.ctx
        g = KD.graph
        N, d = g.n_nodes, KD.dof
        if feat is None:
            feat = self.features(KD, fixed)
        acts = torch.empty(N_LAYERS * N * WIDTH, dtype=torch.float32, device=ctx.torch_device)
        aggs = torch.empty(N * (N_FEATURES + 2 * WIDTH), dtype=torch.float32, device=ctx.torch_device)
        o = torch.empty(N * d, dtype=torch.float32, device=ctx.torch_device)
        M = torch.empty(N * d * d, dtype=torch.float64, device=ctx.torch_device)
        _hip.call('fe_gnn_forward', ctx.h, N, d, _vp(g.bptr), _vp(g.bcol), _vp(self.params(ctx)),
                  _vp(feat), _vp(acts), _vp(aggs), _vp(binv), _vp(o), _vp(M))
        return M, o, feat, acts, aggs

    def preconditioner(self, sysm, KD: BsrMatrix, fixed: torch.Tensor) -> torch.Tensor:
        """Learned block preconditioner for a Dirichlet-applied K (inference, FP64 blocks)."""
        binv = device.precond_setup(KD, 'block_jacobi')
        M, _, _, _, _ = self.forward(KD, fixed, binv)
        return M

    # ---- training ---------------------------------------------------------------------------------
    def loss_and_grad(self, sample: 'TrainingSystem', grad: torch.Tensor) -> float:
        """Accumulate dL/dparams of one system (mean over its pairs) into ``grad``; returns the mean loss."""
        ctx = sample.KD.graph.ctx
        g = sample.KD.graph
        N, d = g.n_nodes, self.dof
        M, o, feat, acts, aggs = self.forward(sample.KD, sample.fixed, sample.binv, sample.feat)
        nwork = ctypes.c_int64()
        _hip.call('fe_gnn_loss_workspace', N * d, ctypes.byref(nwork))
        work = torch.empty(nwork.value, dtype=torch.float64, device=ctx.torch_device)
        loss = torch.zeros(1, dtype=torch.float64, device=ctx.torch_device)
        dlo = torch.zeros(N * d, dtype=torch.float32, device=ctx.torch_device)
        s = sample.KD.struct()
        zcs = sample.zcs if sample.zcs is not None else [None] * len(sample.pairs)
        for (r, e), zc in zip(sample.pairs, zcs):
            _hip.call('fe_gnn_loss_grad', ctx.h, ctypes.byref(s), _vp(sample.binv), _vp(M), _vp(o),
                      _vp(r), _vp(e), _vp(zc), _vp(work), _vp(loss), _vp(dlo), self.loss_kind)
        dlo.mul_(1.0 / len(sample.pairs))
        nb = ctypes.c_int64()
        _hip.call('fe_gnn_backward_workspace', N, ctypes.byref(nb))
        bwork = torch.empty(nb.value, dtype=torch.float32, device=ctx.torch_device)
        _hip.call('fe_gnn_backward', ctx.h, N, d, _vp(g.bptr), _vp(g.bcol), _vp(self.params(ctx)),
                  _vp(feat), _vp(acts), _vp(aggs), _vp(dlo), _vp(bwork), _vp(grad))
        return float(loss.item()) / len(sample.pairs)

    # loss of fe_gnn_loss_grad: 0 = energy left after an optimally damped step, 1 = undamped step
    LOSSES = {'damped': 0, 'undamped': 1}
    loss_kind = 0

    def train(self, samples: Sequence['TrainingSystem'], epochs: int = 20, lr: float = 3e-3,
              seed: int = 0, log=None, loss: str = 'damped') -> List[float]:
        """Adam over the systems (one step per system), returns the mean loss per epoch.  loss:
        'damped' (1 - (r.z)^2/((z.Az)(r.e))) or 'undamped' (|e - z|_A^2 / |e|_A^2)."""
        if loss not in self.LOSSES:
            raise ValueError("loss must be 'damped' or 'undamped'")
        self.loss_kind = self.LOSSES[loss]
        ctx = samples[0].KD.graph.ctx
        P = self.params(ctx)
        n = P.numel()
        m = torch.zeros_like(P)
        v = torch.zeros_like(P)
        grad = torch.zeros_like(P)
        rng = np.random.default_rng(seed)
        history, step = [], 0
        for ep in range(epochs):
            losses = []
            for idx in rng.permutation(len(samples)):
                grad.zero_()
                losses.append(self.loss_and_grad(samples[idx], grad))
                step += 1
                _hip.call('fe_gnn_adam', ctx.h, n, _vp(P), _vp(grad), _vp(m), _vp(v), lr, 0.9, 0.999,
                          1e-8, step, 1.0)
            history.append(float(np.mean(losses)))
            if log:
                log(f'epoch {ep}: loss {history[-1]:.5f}')
        self.sync_to_host()
        self.meta.update({'epochs': epochs, 'lr': lr, 'loss_history': history,
                          'n_systems': len(samples), 'loss': loss})
        return history


class TrainingSystem:
    """One training matrix with its residual/error pairs, resident on the device.  zcs: per pair, the
    network-independent part of the preconditioned residual (the coarse correction of the two-level
    preconditioner), or None for plain block Jacobi."""

    def __init__(self, KD: BsrMatrix, fixed: torch.Tensor, pairs, binv: torch.Tensor, feat=None,
                 zcs=None):
        self.KD = KD
        self.fixed = fixed
        self.pairs = pairs
        self.binv = binv
        self.feat = feat
        self.zcs = zcs


def make_training_systems(n: int = 32, seeds: Sequence[int] = range(1000, 1010),
                          snapshots: Sequence[int] = (10, 40, 160), sigma: float = 1.0,
                          tol: float = 1e-10, log=None, sysm=None, two_level: bool = False,
                          coarse_target: Optional[int] = None, probes: int = 0) -> List[TrainingSystem]:
    """Residual/error pairs from plain PCG on lognormal-field cubes (one system per seed).

    two_level: the pairs come from two-level PCG (block Jacobi + coarse correction) and every pair
    carries its coarse correction, so the network learns the smoother of that preconditioner.
    probes: extra pairs per system with a random (rough) error e, r = K e, so the loss also sees
    the top of the spectrum that the smooth PCG errors miss."""
    from .system import FESystem
    out = []
    for seed in seeds:
        pr = meshgen.cube_elasticity(n=n, sigma=sigma, seed=seed)
        if sysm is None:
            sysm = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
        else:
            sysm.set_materials(pr.materials)
        K = sysm.stiffness_bsr()
        fixed, g = sysm.dirichlet(pr.fixed_dofs)
        KD, b = sysm.prepare(K, fixed, g, sysm.ctx.tensor(pr.f))
        coarse = sysm.coarse_space(pr.fixed_dofs, target=coarse_target) if two_level else None
        solver = sysm.solver(KD, 'two_level' if two_level else 'block_jacobi', 'fp64', coarse=coarse)
        xstar, info = solver.solve(b, tol=tol)
        pairs, zcs = [], []
        for k in snapshots:
            xk, _ = solver.solve(b, tol=tol, max_iter=k)
            rk = b - device.spmv(solver.A64, xk)
            pairs.append((rk, xstar - xk))
            if two_level:
                zcs.append(coarse.apply(solver.Ainv64, rk, torch.zeros_like(rk)))
        if probes:
            gen = torch.Generator(device='cpu').manual_seed(int(seed))
            for _ in range(probes):
                e = torch.randn(b.numel(), generator=gen, dtype=torch.float64).to(b.device)
                e[fixed.bool()] = 0.0
                r = device.spmv(solver.A64, e)
                pairs.append((r, e))
                if two_level:
                    zcs.append(coarse.apply(solver.Ainv64, r, torch.zeros_like(r)))
        binv = device.precond_setup(KD, 'block_jacobi')
        ts = TrainingSystem(solver.A64, fixed, pairs, binv, zcs=zcs if two_level else None)
        out.append(ts)
        if log:
            log(f'training system seed {seed}: {info["iterations"]} it to {info["relres"]:.1e}')
    return out


def solve_with_fallback(sysm, KD: BsrMatrix, b: torch.Tensor, fixed: torch.Tensor,
                        model: LearnedPreconditioner, precision: str = 'mixed', tol: float = 1e-8,
                        max_iter: Optional[int] = None, coarse=None):
    """Learned-preconditioned PCG (two-level when ``coarse`` is given) judged by its true FP64
    residual; the same PCG without the network re-solves when it fails (``info['fallback']``), then
    plain FP64 PCG.  Raises ``analysis.SolveError`` when no attempt converges."""
    from .analysis import solve_reduced
    pc = 'two_level' if coarse is not None else 'block_jacobi'
    return solve_reduced(sysm, KD, b, fixed, pc, precision, tol=tol, max_iter=max_iter, model=model,
                         coarse=coarse)
