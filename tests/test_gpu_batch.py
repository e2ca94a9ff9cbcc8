"""Batched uneven plates: one block-diagonal assembly and one segmented PCG for many meshes.

Each plate's solution is checked against the oracle (scipy direct solve of that plate's eliminated
system, oracle/vectorized.py), and the batched path against solving every plate on its own: the
segments must not interact, so the per-plate results are bitwise identical.
"""

import numpy as np
import pytest
import scipy.sparse.linalg

from finite_elements_amd import meshgen
from finite_elements_amd.batch import MeshBatch
from finite_elements_amd.system import FESystem
from oracle import vectorized as vz

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _small_plates(count, seed, nx_max=60, ny_max=20):
    """The plate generator's draws, capped in size so the oracle's direct solves stay quick."""
    out = []
    for pr in meshgen.plates(count=count, seed=seed):
        if pr.meta['nx'] <= nx_max and pr.meta['ny'] <= ny_max:
            out.append(pr)
    return out


def _oracle_solution(pr):
    K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=pr.plane_mode)
    return vz.reduced_solve(K, pr.f, pr.fixed_dofs, pr.fixed_values), K


def test_batch_matches_oracle_per_plate(ctx):
    _say('batch vs oracle')
    plates = _small_plates(400, 3)[:16]
    assert len(plates) >= 8
    mb = MeshBatch(plates)
    u, info = mb.solve(tol=1e-10, precision='fp64', graph=True)
    seg = info['segments']
    _say(f'  {len(plates)} plates, iterations {seg["iterations"].min()}..{seg["iterations"].max()}, '
         f'status {seg["status"]}, relres {seg["relres"]}')
    assert np.all(seg['status'] == 0)
    assert np.all(seg['relres'] <= 1e-10)
    for pr, up in zip(plates, mb.split(u)):
        ref, K = _oracle_solution(pr)
        free = np.ones(up.size, bool)
        free[pr.fixed_dofs] = False
        res = np.linalg.norm((pr.f - K @ up)[free]) / np.linalg.norm(pr.f[free])
        assert res <= 2e-10, res
        assert np.linalg.norm(up - ref) <= 1e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_batch_equals_individual_solves(ctx, precision):
    """Segments do not interact: every plate of the batch is bitwise the plate solved alone."""
    _say(f'batch vs individual ({precision})')
    plates = _small_plates(200, 7)[:5]
    mb = MeshBatch(plates, tune=False)
    u, info = mb.solve(tol=1e-8, precision=precision, graph=False)
    for g, (pr, up) in enumerate(zip(plates, mb.split(u))):
        s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials, plane_mode=pr.plane_mode)
        fixed, gv = s.dirichlet(pr.fixed_dofs, pr.fixed_values)
        KD, b = s.prepare(s.stiffness_bsr(), fixed, gv, s.ctx.tensor(pr.f))
        x, inf1 = s.solver(KD, 'block_jacobi', precision).solve(b, tol=1e-8)
        assert inf1['iterations'] == info['segments']['iterations'][g]
        assert np.array_equal(x.cpu().numpy(), up)


def test_batch_mixed_precision_reaches_tolerance(ctx):
    """FP32 iterations inside FP64 reliable updates: every plate reaches true FP64 relres 1e-8."""
    _say('batch mixed precision')
    plates = _small_plates(400, 5)[:16]
    mb = MeshBatch(plates, tune=False)
    u, info = mb.solve(tol=1e-8, precision='mixed', graph=True)
    seg = info['segments']
    _say(f'  iterations {seg["iterations"]}, status {seg["status"]}, relres {seg["relres"]}')
    assert np.all(seg['status'] == 0) and np.all(seg['relres'] <= 1e-8)
    for pr, up in zip(plates, mb.split(u)):
        ref, K = _oracle_solution(pr)
        free = np.ones(up.size, bool)
        free[pr.fixed_dofs] = False
        res = np.linalg.norm((pr.f - K @ up)[free]) / np.linalg.norm(pr.f[free])
        assert res <= 2e-8, res


def test_batch_graph_replay_equals_stream_launches(ctx):
    _say('batch graph vs stream')
    plates = _small_plates(200, 11)[:6]
    mb = MeshBatch(plates, tune=False)
    u1, i1 = mb.solve(tol=1e-8, graph=False)
    u2, i2 = mb.solve(tol=1e-8, graph=True)
    assert np.array_equal(u1.cpu().numpy(), u2.cpu().numpy())
    assert np.array_equal(i1['segments']['iterations'], i2['segments']['iterations'])


def test_batch_4096_plates(ctx):
    """The full 4096-plate draw (24.6M triangles): every plate reaches true FP64 relative residual
    1e-8 in one batched solve; a sample of plates against the oracle."""
    _say('4096 plates')
    plates = meshgen.plates(count=4096, seed=3)
    mb = MeshBatch(plates)
    _say(f'  {mb.sysm.n_elem} elements, {mb.sysm.n} DOF, nnz {mb.sysm.pattern.nnz}, schedule {mb.schedule}')
    u, info = mb.solve(tol=1e-8, precision='mixed', graph=True)
    seg = info['segments']
    _say(f'  iterations {seg["iterations"].min()}..{seg["iterations"].max()}, '
         f'max relres {seg["relres"].max():.2e}, FP64 fallbacks {info["fallbacks"]}')
    assert np.all(seg['status'] == 0)
    assert np.all(seg['relres'] <= 1e-8)
    parts = mb.split(u)
    rng = np.random.default_rng(0)
    for g in rng.choice(len(plates), 6, replace=False):
        pr = plates[g]
        ref, K = _oracle_solution(pr)
        free = np.ones(parts[g].size, bool)
        free[pr.fixed_dofs] = False
        res = np.linalg.norm((pr.f - K @ parts[g])[free]) / np.linalg.norm(pr.f[free])
        assert res <= 2e-8, (g, res)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_batched_two_level(ctx, precision):
    """Block-diagonal coarse space of a batch: every plate's coarse block and its inverse against the
    oracle's P^T A P of the whole batch (which must be block diagonal), each plate's solution against
    its direct solve, and fewer iterations than block Jacobi on the longest plates."""
    from oracle import coarse as oc
    _say(f'batched two-level {precision}')
    plates = _small_plates(400, 5)[:12]
    assert len(plates) >= 6
    mb = MeshBatch(plates)
    dc = mb.coarse_space(max_block=60)
    cs = dc.host
    assert dc.n_seg == len(plates) and cs.max_block <= 60
    # aggregates stay inside their plate
    offs = mb.node_offsets
    for s in range(len(plates)):
        nodes = np.flatnonzero(cs.seg_of_agg[cs.node_agg] == s)
        assert nodes.min() >= offs[s] and nodes.max() < offs[s + 1]
    if precision == 'fp64':
        import scipy.sparse as sp
        from finite_elements_amd import _hip
        K, KD, b = mb.assemble()
        d = KD.dof
        g = KD.graph.pattern
        bv = KD.convert(_hip.FE_LAYOUT_BSR).vals.cpu().numpy().reshape(-1, d, d)
        Ah = sp.bsr_matrix((bv, g.bcol, g.bptr), shape=(g.n_nodes * d,) * 2).tocsr()
        ref, P = oc.coarse_operator(cs, Ah)
        Ac, _ = dc.operator(KD, invert=False)
        Ac = Ac.cpu().numpy()
        inv64, inv32 = dc.operator(KD, invert=True)
        inv64 = inv64.cpu().numpy()
        for s in range(len(plates)):
            c0, c1 = cs.agg_seg[s] * cs.nc, cs.agg_seg[s + 1] * cs.nc
            n = c1 - c0
            blk = Ac[cs.blk_ofs[s]:cs.blk_ofs[s + 1]].reshape(n, n)
            want = ref[c0:c1, c0:c1]
            assert np.abs(blk - want).max() <= 1e-12 * np.abs(want).max()
            # nothing couples two plates
            assert not ref[c0:c1, :c0].any() and not ref[c0:c1, c1:].any()
            binv = inv64[cs.blk_ofs[s]:cs.blk_ofs[s + 1]].reshape(n, n)
            err = np.abs(binv - np.linalg.inv(want)).max() / np.abs(np.linalg.inv(want)).max()
            assert err <= 1e-12 * np.linalg.cond(want)
    u, info = mb.solve(tol=1e-10, precision=precision, precond='two_level')
    seg = info['segments']
    _, info_bj = mb.solve(tol=1e-10, precision=precision, precond='block_jacobi')
    seg_bj = info_bj['segments']
    _say(f'  iterations two-level {seg["iterations"].max()} (sum {seg["iterations"].sum()}), '
         f'block Jacobi {seg_bj["iterations"].max()} (sum {seg_bj["iterations"].sum()}), '
         f'fallbacks {seg["fallback"].sum()}')
    assert np.all(seg['status'] == 0) and np.all(seg['relres'] <= 1e-10)
    assert seg['iterations'].max() < seg_bj['iterations'].max()
    for pr, up in zip(plates, mb.split(u)):
        ref_u, _ = _oracle_solution(pr)
        assert np.linalg.norm(up - ref_u) <= 1e-6 * np.linalg.norm(ref_u)


def test_solver_chunk_autotune(ctx):
    """The batched solver's chunk schedule is autotuned like the assembly's: every candidate chunk
    length solves every plate, the fastest is kept, and later solves use it."""
    _say('batched solver chunk autotune')
    plates = _small_plates(200, 11)[:10]
    mb = MeshBatch(plates, tune=False)
    mb.coarse_space(max_block=60)
    best, times = mb.tune_solver(candidates=(4, 8, 16), reps=1)
    _say(f'  {len(plates)} plates: seconds per solve {times}, best {best}')
    assert set(times) == {4, 8, 16} and best == min(times, key=times.get)
    assert mb.check_every == best
    _, info = mb.solve(tol=1e-8, precision='fp64', precond='two_level')
    seg = info['segments']
    assert np.all(seg['status'] == 0) and np.all(seg['relres'] <= 1e-8)


def test_split_batch_precision_classes(ctx):
    """SplitBatch: the long thin plates (aspect ratio >= 4) in FP64, the others in FP32 inside FP64
    reliable updates, two batches on two streams solved concurrently.  Every plate against its
    direct solve; the FP64 class bitwise equal to the same plates batched and solved alone in FP64
    (one launch per class, same per-segment arithmetic)."""
    from finite_elements_amd.batch import SplitBatch, aspect_ratio
    _say('split batch (precision classes, two streams)')
    plates = _small_plates(400, 7)[:16]
    asp = np.array([aspect_ratio(pr) for pr in plates])
    assert (asp < 4).any() and (asp >= 4).any()
    sb = SplitBatch(plates, fp64_aspect=4.0, tune=False)
    sb.coarse_space(max_block=60)
    u, info = sb.solve(tol=1e-10, precond='two_level')
    seg = info['segments']
    _say(f'  {len(plates)} plates: {int((info["precision"] == "mixed").sum())} mixed, '
         f'{int((info["precision"] == "fp64").sum())} FP64, iterations {seg["iterations"].max()}, '
         f'fallbacks {info["fallbacks"]}')
    assert info['status'] == 0 and np.all(seg['status'] == 0) and np.all(seg['relres'] <= 1e-10)
    assert list(info['precision']) == ['fp64' if a >= 4 else 'mixed' for a in asp]
    parts = sb.split(u)
    for pr, x in zip(plates, parts):
        ref, _ = _oracle_solution(pr)
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-6
    hi = [pr for pr, a in zip(plates, asp) if a >= 4]
    mb = MeshBatch(hi, tune=False)
    mb.coarse_space(max_block=60)
    u64, _ = mb.solve(tol=1e-10, precision='fp64', precond='two_level')
    for a, b in zip(mb.split(u64), [x for x, a in zip(parts, asp) if a >= 4]):
        assert np.array_equal(a, b)
