"""Modal analysis on the GPU: batched LOBPCG (``fe_lobpcg``) for K phi = lambda M phi.

Replaces ``FiniteElementAnalysis.modal_analysis`` (finite_elements/analysis.py:741-796):

* ``order='largest'`` -- the reference calls ``eigsh(K, M, which='LM', k)`` on the unconstrained K and
  M; here LOBPCG iterates for the largest eigenvalues of the same unconstrained pencil.
* ``order='smallest'`` -- the reference inverts the dense (unconstrained, hence singular) K and calls
  ``eigs``; here LOBPCG with the block-Jacobi preconditioner of K iterates for the smallest eigenvalues
  of the pencil restricted to the free DOFs (the analysis's Dirichlet conditions remove the
  constrained DOFs), which is the well-posed form of that request.

Both return ``(eigvals, eigvecs.T)`` like the reference, eigenvalues ascending.  Many variants of one
mesh (same pattern, different K values) are solved in one batched call by ``lobpcg``.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Union

import numpy as np
import torch

from . import _hip, device
from .device import BsrMatrix


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _planes64(A: BsrMatrix) -> BsrMatrix:
    if A.layout != _hip.FE_LAYOUT_PLANES or A.precision != 0:
        A = A.convert(_hip.FE_LAYOUT_PLANES, torch.float64)
    return A


def _bsr64(A: BsrMatrix) -> torch.Tensor:
    """The FP64 values in the block-contiguous layout the eigen-solve's block SpMM reads."""
    if A.layout == _hip.FE_LAYOUT_BSR and A.precision == 0:
        return A.vals
    return A.convert(_hip.FE_LAYOUT_BSR, torch.float64).vals


def lobpcg(K: Union[BsrMatrix, Sequence[BsrMatrix]], M: Union[BsrMatrix, Sequence[BsrMatrix]],
           n_wanted: int, block: Optional[int] = None, fixed: Optional[torch.Tensor] = None,
           precond: Union[str, None, torch.Tensor, Sequence[torch.Tensor]] = 'block_jacobi',
           largest: bool = False, tol: float = 1e-6, max_iter: int = 2000, X0=None, seed: int = 0,
           check_every: int = 8, coarse=None, iteration_layout: int = _hip.FE_LAYOUT_BSR,
           streams: int = 1):
    """Lowest (or largest) eigenpairs of one or several variants K_v phi = lambda M_v phi.

    K, M: one matrix or a list of variants on one pattern (FP64; converted to the plane layout).  A
    single M is shared by all variants.  fixed: uint8 DOF mask of constrained DOFs (their rows and
    columns of K and M must already be decoupled, e.g. by ``device.apply_dirichlet``); the iteration
    stays in the free subspace.  precond: 'block_jacobi' (blocks of each K_v), None, or explicit
    FP64 blocks [n_brows][d][d] (one tensor or one per variant).  coarse: a ``DeviceCoarse`` of the
    constrained mesh (``FESystem.coarse_space``): the preconditioner becomes two-level, T + P A_c^-1 P^T
    with A_c = P^T K_v P per variant (smallest eigenvalues only).  iteration_layout: the value
    layout the block SpMM reads (FE_LAYOUT_BSR, or FE_LAYOUT_PLANES; same results bit for bit).
    X0: initial block [nvar][n][block] (host array); by default a standard normal block drawn on the
    device from ``seed``.  streams: 1 to 4; above 1 the variants are split into that many groups
    iterated by concurrent calls on streams of their own (one host thread each), so that one
    group's latency-bound Rayleigh-Ritz launches (one workgroup per variant) overlap the others'
    bandwidth-bound kernels.  A variant's iteration is the same except that the Gram partial sums
    are cut into more row chunks when a call holds fewer variants (rounding-level differences).
    Returns (eigvals [nvar][n_wanted] numpy, eigvecs [nvar][n][n_wanted] torch FP64, info).
    """
    Ks = [K] if isinstance(K, BsrMatrix) else list(K)
    Ms = [M] if isinstance(M, BsrMatrix) else list(M)
    nvar = len(Ks)
    if nvar == 0:
        raise ValueError('lobpcg: no matrices')
    if len(Ms) not in (1, nvar):
        raise ValueError('lobpcg: give one M or one per variant')
    g = Ks[0].graph
    ctx = g.ctx
    d = Ks[0].dof
    n = g.n_nodes * d
    if block is None:
        block = min(32, max(4, ((n_wanted + max(4, n_wanted // 4) + 3) // 4) * 4))
    if not 1 <= n_wanted <= block:
        raise ValueError('lobpcg: need 1 <= n_wanted <= block')
    Ks = [_planes64(A) for A in Ks]
    Ms = [_planes64(A) for A in Ms]
    for A in Ks + Ms:
        if A.graph is not g or A.dof != d:
            raise ValueError('lobpcg: all matrices must share one pattern')
    nv_vals = Ks[0].vals.numel()
    # the iteration reads K and M block-contiguous; the set-up below (block Jacobi, the coarse
    # operators) reads the plane layout
    if iteration_layout not in (_hip.FE_LAYOUT_BSR, _hip.FE_LAYOUT_PLANES):
        raise ValueError('lobpcg: iteration_layout must be FE_LAYOUT_BSR or FE_LAYOUT_PLANES')
    vals = _bsr64 if iteration_layout == _hip.FE_LAYOUT_BSR else (lambda A: A.vals)
    Kall = vals(Ks[0]) if nvar == 1 else torch.stack([vals(A) for A in Ks])
    Mall = vals(Ms[0]) if len(Ms) == 1 else torch.stack([vals(A) for A in Ms])
    k_stride = nv_vals if nvar > 1 else 0
    m_stride = nv_vals if len(Ms) > 1 else 0
    # preconditioner blocks
    T = None
    t_stride = 0
    if isinstance(precond, str):
        if precond != 'block_jacobi':
            raise ValueError("precond must be 'block_jacobi', None or explicit blocks")
        Ts = [device.precond_setup(A, 'block_jacobi') for A in Ks]
        T = Ts[0] if nvar == 1 else torch.stack(Ts)
        t_stride = Ts[0].numel() if nvar > 1 else 0
    elif isinstance(precond, torch.Tensor):
        T = precond.to(torch.float64).contiguous()
        t_stride = 0
    elif precond is not None:
        Ts = [t.to(torch.float64).contiguous() for t in precond]
        T = torch.stack(Ts)
        t_stride = Ts[0].numel()
    # initial block
    if X0 is None:
        # standard normal start drawn on the device (seeded): 64 variants of the 48^3 cube need
        # 5 GB of it, which a host generator takes seconds to produce
        gen = torch.Generator(device=ctx.torch_device)
        gen.manual_seed(int(seed))
        X = torch.randn((nvar, n, block), generator=gen, dtype=torch.float64, device=ctx.torch_device)
    else:
        X0 = np.asarray(X0, dtype=np.float64).reshape(nvar, n, block)
        X = torch.as_tensor(X0, device=ctx.torch_device).contiguous().clone()
    if fixed is not None:
        X[:, fixed.bool(), :] = 0.0
    lam = torch.zeros(nvar * block, dtype=torch.float64, device=ctx.torch_device)
    res = torch.zeros_like(lam)
    iters = torch.zeros(nvar, dtype=torch.int32, device=ctx.torch_device)
    vstat = torch.zeros(nvar, dtype=torch.int32, device=ctx.torch_device)
    if streams not in (1, 2, 3, 4):
        raise ValueError('lobpcg: streams must be 1, 2, 3 or 4')
    if coarse is not None:
        if largest:
            raise ValueError('lobpcg: the two-level preconditioner approximates K^-1 (smallest modes)')
        # per-variant coarse inverses A_c^-1 = (P^T K_v P)^-1, FP64
        Ainv = torch.stack([coarse.operator(A, invert=True)[0] for A in Ks])
        cs = coarse.struct()
    # variant groups [v0, v1), each a call of its own on a handle of its own
    ng = max(1, min(streams, nvar))
    bounds = [(nvar * k + ng - 1) // ng for k in range(ng + 1)]
    groups = []
    for gi, (v0, v1) in enumerate(zip(bounds[:-1], bounds[1:])):
        gctx = ctx if gi == 0 else ctx.side(gi - 1)
        part = lambda t, stride: t if (stride == 0 or nvar == 1) else t[v0:v1]  # noqa: E731
        kg, mg = part(Kall, k_stride), part(Mall, m_stride)
        tg = None if T is None else part(T, t_stride)
        opts = _hip.FeLobpcgOpts(v1 - v0, block, n_wanted, int(bool(largest)), int(max_iter),
                                 int(check_every), float(tol), k_stride, m_stride, t_stride)
        ks = _hip.FeBsr(g.n_nodes, g.nnzb, d, 0, iteration_layout, 0, _vp(g.bptr), _vp(g.bcol), _vp(kg))
        ms = _hip.FeBsr(g.n_nodes, g.nnzb, d, 0, iteration_layout, 0, _vp(g.bptr), _vp(g.bcol), _vp(mg))
        nbytes = ctypes.c_int64()
        if coarse is None:
            _hip.call('fe_lobpcg_workspace', gctx.h, ctypes.byref(ks), ctypes.byref(opts),
                      ctypes.byref(nbytes))
        else:
            _hip.call('fe_lobpcg_coarse_workspace', gctx.h, ctypes.byref(ks), ctypes.byref(cs),
                      ctypes.byref(opts), ctypes.byref(nbytes))
        work = torch.empty(nbytes.value, dtype=torch.uint8, device=ctx.torch_device)
        sl = slice(v0 * block, v1 * block)
        groups.append(dict(ctx=gctx, opts=opts, ks=ks, ms=ms, T=tg, X=X[v0:v1], lam=lam[sl],
                           res=res[sl], iters=iters[v0:v1], vstat=vstat[v0:v1], work=work,
                           Ainv=None if coarse is None else Ainv[v0:v1],
                           stats=_hip.FeLobpcgStats()))

    def run(gr):
        a = (ctypes.byref(gr['ks']), ctypes.byref(gr['ms']), _vp(gr['T']), _vp(fixed))
        tail = (_vp(gr['X']), _vp(gr['lam']), _vp(gr['res']), _vp(gr['iters']), _vp(gr['vstat']),
                _vp(gr['work']), gr['work'].numel(), ctypes.byref(gr['opts']), ctypes.byref(gr['stats']))
        if coarse is None:
            _hip.call('fe_lobpcg', gr['ctx'].h, *a, *tail)
        else:
            _hip.call('fe_lobpcg_coarse', gr['ctx'].h, *a, ctypes.byref(cs), _vp(gr['Ainv']),
                      coarse.ld * coarse.ld, *tail)

    if len(groups) == 1:
        run(groups[0])
    else:
        # everything the groups read is ready before any stream starts; ctypes releases the GIL
        # for the calls, so the host loops (each polling its own stream) run side by side
        torch.cuda.synchronize(ctx.device)
        import threading
        errors = []

        def guarded(gr):
            try:
                run(gr)
            except Exception as exc:  # re-raised below, after every call returned
                errors.append(exc)

        ths = [threading.Thread(target=guarded, args=(gr,)) for gr in groups[1:]]
        for th in ths:
            th.start()
        guarded(groups[0])
        for th in ths:
            th.join()
        torch.cuda.synchronize(ctx.device)
        if errors:
            raise errors[0]
    stats = groups[0]['stats']
    for gr in groups[1:]:
        q = gr['stats']
        stats.max_iterations = max(stats.max_iterations, q.max_iterations)
        stats.unconverged += q.unconverged
        stats.failed += q.failed
        if q.status != 0 and (stats.status == 0 or q.status == _hip.FE_ERR_BREAKDOWN):
            stats.status = q.status
    lam_h = lam.cpu().numpy().reshape(nvar, block)
    res_h = res.cpu().numpy().reshape(nvar, block)
    info = {'iterations': iters.cpu().numpy(), 'variant_status': vstat.cpu().numpy(),
            'max_iterations': stats.max_iterations, 'unconverged': stats.unconverged,
            'failed': stats.failed, 'status': stats.status, 'resid': res_h[:, :n_wanted],
            'block': block}
    order = np.argsort(-lam_h[:, :n_wanted], axis=1) if largest else None
    vals = lam_h[:, :n_wanted]
    vecs = X[:, :, :n_wanted]
    if largest:
        # ascending like eigsh
        idx = torch.as_tensor(order[:, ::-1].copy(), device=ctx.torch_device)
        vals = np.take_along_axis(vals, order[:, ::-1], axis=1)
        vecs = torch.gather(vecs, 2, idx[:, None, :].expand(-1, n, -1))
    return vals, vecs, info


def sym_eig(A: torch.Tensor):
    """Batched symmetric eigen-decomposition of [n_mat][r][r] FP64 matrices on the GPU (Jacobi).

    Returns (w [n_mat][r], V [n_mat][r][r]) with eigenvalues ascending and eigenvectors in columns."""
    from .device import Context
    if A.dim() == 2:
        A = A[None]
    ctx = Context.get(A.device.index or 0)
    A = A.to(torch.float64).contiguous()
    nmat, r, _ = A.shape
    w = torch.empty(nmat, r, dtype=torch.float64, device=A.device)
    V = torch.empty_like(A)
    _hip.call('fe_sym_eig', ctx.h, nmat, r, _vp(A), _vp(w), _vp(V))
    order = torch.argsort(w, dim=1)
    w = torch.gather(w, 1, order)
    V = torch.gather(V, 2, order[:, None, :].expand(-1, r, -1))
    return w, V


def modal_analysis(analysis, order: str, k: int, method: str = 'lobpcg_gpu', tol: float = 1e-7,
                   max_iter: int = 5000, block: Optional[int] = None, seed: int = 0,
                   two_level: bool = True):
    """(eigvals, eigvecs.T) of the analysis's K, M (finite_elements/analysis.py:741-796).  For the
    smallest modes of a constrained elasticity problem the iteration is preconditioned by the
    two-level method (block Jacobi + rigid-body coarse correction) unless ``two_level`` is False."""
    if order not in ('largest', 'smallest'):
        raise ValueError("Order parameter should be either 'largest' or 'smallest'")
    if method != 'lobpcg_gpu':
        raise ValueError("finite_elements_amd computes modes on the GPU only (method='lobpcg_gpu')")
    sysm = analysis.system()
    analysis._plane_mode()
    K = sysm.stiffness_bsr()
    Mm = sysm.mass_bsr()
    if order == 'largest':
        vals, vecs, info = lobpcg(K, Mm, k, block=block, precond=None, largest=True, tol=tol,
                                  max_iter=max_iter, seed=seed)
    else:
        pos, _ = analysis.boundary_condition_dofs()
        fixed, _ = sysm.dirichlet(pos)
        KD = device.apply_dirichlet(K, fixed)
        MD = device.apply_dirichlet(Mm, fixed)
        coarse = sysm.coarse_space(pos) if (two_level and len(pos) and sysm.dof > 1) else None
        vals, vecs, info = lobpcg(KD, MD, k, block=block, fixed=fixed, precond='block_jacobi',
                                  tol=tol, max_iter=max_iter, seed=seed, coarse=coarse)
    analysis.last_modal_info = info
    if info['status'] != 0:
        import warnings
        warnings.warn(f'LOBPCG did not converge: status {info["status"]}, '
                      f'max residual {np.max(info["resid"]):.2e}')
    return vals[0], vecs[0].cpu().numpy().T
