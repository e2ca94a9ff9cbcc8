"""The allocation-free coarse correction (fe_coarse_apply_noalloc, for callers that cannot allocate
device memory) against the workspace entry point it mirrors (fe_coarse_apply), through the C ABI:
two calls in a row on one handle, the second on a smaller coarse space, in FP64 and FP32, for a
one-system coarse space and a block-diagonal (batched) one.  The same kernels run in both, so the
results must agree bitwise.  Also checks that a coarse space above the stated maximum is refused."""

import ctypes

import numpy as np
import pytest
import torch

from finite_elements_amd import _hip, meshgen
from finite_elements_amd.batch import MeshBatch
from finite_elements_amd.system import FESystem

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _single(n):
    pr = meshgen.cube_elasticity(n=n)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    fixed, g = s.dirichlet(pr.fixed_dofs)
    KD, _ = s.prepare(s.stiffness_bsr(), fixed, g, s.ctx.tensor(pr.f))
    dc = s.coarse_space(pr.fixed_dofs, target=27)
    return s, KD, dc


def _batched(count):
    mb = MeshBatch(meshgen.plates(count=count, seed=11), tune=False)
    s = mb.sysm
    fixed, g = s.dirichlet(mb.fixed_dofs, mb.fixed_values)
    KD, _ = s.prepare(s.stiffness_bsr(), fixed, g, s.ctx.tensor(np.zeros(s.n)))
    dc = mb.coarse_space()
    return s, KD, dc


def _compare(s, KD, dc, seed):
    """Both entry points on the same r and u, both precisions; returns the coarse DOF count."""
    lib = _hip.load()
    ctx = s.ctx
    inv64, inv32 = dc.operator(KD, invert=True)
    c = dc.struct()
    rng = np.random.default_rng(seed)
    for prec, dt in ((0, torch.float64), (1, torch.float32)):
        ainv = inv64 if (prec == 0 or dc.n_seg) else inv32
        r = torch.as_tensor(rng.standard_normal(s.n), dtype=dt, device=ctx.torch_device)
        u0 = torch.as_tensor(rng.standard_normal(s.n), dtype=dt, device=ctx.torch_device)
        u_ws, u_na = u0.clone(), u0.clone()
        el = 8 if prec == 0 else 4
        work = torch.empty(2 * ((dc.n_coarse * el + 255) // 256) * 256, dtype=torch.uint8,
                           device=ctx.torch_device)
        _hip.call('fe_coarse_apply', ctx.h, ctypes.byref(c), prec, _vp(ainv), _vp(r), _vp(u_ws),
                  _vp(work))
        rc = lib.fe_coarse_apply_noalloc(ctx.h, ctypes.byref(c), prec, _vp(ainv), _vp(r), _vp(u_na))
        assert rc == 0, lib.fe_last_error().decode()
        ctx.synchronize()
        assert not torch.equal(u_ws, u0), 'the correction changed nothing'
        assert torch.equal(u_ws, u_na), f'precision {prec}: results differ'
    return dc.n_coarse


def test_noalloc_single_system_two_calls(ctx):
    """A 12^3 cube's coarse space, then a smaller 6^3 one, on the same handle and scratch array."""
    _say('fe_coarse_apply_noalloc: one system, 12^3 then 6^3')
    n1 = _compare(*_single(12), seed=1)
    n2 = _compare(*_single(6), seed=2)
    _say(f'  coarse DOF {n1} then {n2}: bitwise equal to fe_coarse_apply')
    assert n2 < n1


def test_noalloc_block_diagonal_two_calls(ctx):
    """The per-plate block-diagonal coarse space of 48 uneven plates, then of 12."""
    _say('fe_coarse_apply_noalloc: block-diagonal, 48 plates then 12')
    n1 = _compare(*_batched(48), seed=3)
    n2 = _compare(*_batched(12), seed=4)
    _say(f'  coarse DOF {n1} then {n2}: bitwise equal to fe_coarse_apply')
    assert n2 < n1


def test_noalloc_rejects_oversized_space(ctx):
    """A coarse space above FE_COARSE_NOALLOC_MAX is refused with FE_ERR_ARG before any launch (the
    struct is a copy whose aggregate count is enlarged: nothing is read from the device)."""
    _say('fe_coarse_apply_noalloc: oversized coarse space refused')
    s, KD, dc = _batched(12)
    c = dc.struct()
    c.n_agg = _hip.FE_COARSE_NOALLOC_MAX // c.nc + 1
    lib = _hip.load()
    x = torch.zeros(s.n, dtype=torch.float64, device=s.ctx.torch_device)
    rc = lib.fe_coarse_apply_noalloc(s.ctx.h, ctypes.byref(c), 0, _vp(x), _vp(x), _vp(x))
    assert rc == _hip.FE_ERR_ARG
    assert 'allocation-free maximum' in lib.fe_last_error().decode()
