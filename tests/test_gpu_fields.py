"""GPU parity of the per-element field kernel (fe_element_fields: strain, stress, energy; magnetic
field) against the CPU oracle (oracle/vectorized.element_fields, pinned by the patch test, rigid
motions and the assembled energy), and the reference-named result properties built on it.

Tolerances: strain, stress and B within 1e-12 relative to the largest component of the element's
field (the kernel and numpy order the FP64 products differently); energies within 1e-11 relative;
summed energy against 1/2 u^T K u with the GPU SpMV of the assembled K within 1e-11."""

import numpy as np
import pytest

from finite_elements_amd import device, meshgen
from finite_elements_amd.system import FESystem
from oracle import vectorized as vz
from tests.test_gpu_analysis import build

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


def _case(kind, mode, rng):
    if kind == 'tet_elastic':
        c, t = meshgen.kuhn_tets(5)
        c = c + rng.uniform(-0.03, 0.03, c.shape)
        mats = {'E': rng.uniform(5e10, 3e11, t.shape[0]), 'nu': rng.uniform(0.2, 0.45, t.shape[0])}
    else:
        c, t = meshgen.structured_tri(23, 17, 2.0, 1.5)
        c = c + rng.uniform(-0.01, 0.01, c.shape)
        if kind == 'tri_elastic':
            mats = {'E': rng.uniform(5e10, 3e11, t.shape[0]), 'nu': rng.uniform(0.2, 0.45, t.shape[0]),
                    'thickness': rng.uniform(0.005, 0.5, t.shape[0])}
        else:
            mats = {'mu': 4e-7 * np.pi * rng.uniform(1, 1000, t.shape[0])}
    return c, t, mats


def _rel(got, want):
    scale = np.maximum(np.abs(want).max(axis=1, keepdims=True), 1e-300)
    return float(np.max(np.abs(got - want) / scale))


@pytest.mark.parametrize('kind,mode', [('tet_elastic', 0), ('tri_elastic', 1), ('tri_elastic', 2),
                                       ('tri_magnetic', 0)])
def test_fields_match_oracle(ctx, kind, mode):
    _say(f'element fields {kind} mode={mode}')
    rng = np.random.default_rng(21)
    c, t, mats = _case(kind, mode, rng)
    s = FESystem(kind, c, t, mats, plane_mode=mode)
    u = rng.standard_normal(s.n) * 1e-3
    f_gpu, s_gpu, e_gpu = s.element_fields(u)
    f_ref, s_ref, e_ref = vz.element_fields(kind, c, t, mats, u, mode)
    err = _rel(f_gpu.cpu().numpy().T, f_ref)
    _say(f'  field rel err {err:.2e}')
    assert err < 1e-12
    if kind == 'tri_magnetic':
        assert s_gpu is None and e_gpu is None
        return
    assert _rel(s_gpu.cpu().numpy().T, s_ref) < 1e-12
    e = e_gpu.cpu().numpy()
    np.testing.assert_allclose(e, e_ref, rtol=1e-11, atol=1e-11 * np.abs(e_ref).max())
    # the summed energy is 1/2 u^T K u of the assembled matrix (GPU SpMV)
    K = s.stiffness_bsr()
    ku = device.spmv(K, ctx.tensor(u)).cpu().numpy()
    total = 0.5 * u @ ku
    assert abs(e.sum() - total) <= 1e-11 * abs(total)


def test_energy_identity_at_size(ctx):
    """Size-independent check on the 32^3 lognormal cube: sum of element energies = 1/2 u^T K u."""
    _say('element fields energy identity 32^3')
    pr = meshgen.cube_elasticity(n=32)
    s = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
    rng = np.random.default_rng(2)
    u = ctx.tensor(rng.standard_normal(s.n) * 1e-4)
    _, _, e = s.element_fields(u)
    ku = device.spmv(s.stiffness_bsr(), u)
    total = 0.5 * float(u @ ku)
    assert abs(float(e.sum()) - total) <= 1e-11 * abs(total)


def test_result_properties(ctx):
    """ElasticityResults / MagneticResults dictionaries (reference names) from the kernel."""
    _say('result properties')
    from finite_elements_amd import results
    for key in ('ex2', 'tet1', 'mag'):
        an = build(key)
        res = an.solve(tol=1e-10)
        x = np.asarray(res.result_vector)
        kind, coords, conn, mats = an._mesh_arrays()
        mode = an._plane_mode()
        n = an.n_dofs
        f_ref, s_ref, e_ref = vz.element_fields(kind, coords, conn, mats, x[:n], mode)
        els = [el for g in an.mesh.elements_groups for el in g.elements]
        if kind == 'tri_magnetic':
            r = results.MagneticResults(an.mesh, list(x))
            got = np.array([r.magnetic_field_per_element[el] for el in els])
            assert _rel(got, f_ref) < 1e-12
            np.testing.assert_allclose(r.magnetic_field_norm(), np.hypot(*f_ref.T), rtol=1e-12)
            continue
        cls = results.ElasticityResults3D if kind == 'tet_elastic' else results.ElasticityResults2D
        r = cls(an.mesh, list(x), plane_strain=an.plane_strain, plane_stress=an.plane_stress)
        assert _rel(np.array([r.strain[el] for el in els]), f_ref) < 1e-12
        assert _rel(np.array([r.stress[el] for el in els]), s_ref) < 1e-12
        np.testing.assert_allclose([r.energy_per_element[el] for el in els], e_ref, rtol=1e-11,
                                   atol=1e-11 * np.abs(e_ref).max())
        assert abs(r.energy - e_ref.sum()) <= 1e-11 * abs(e_ref.sum())
        d = an.dimension
        assert r.displacements_per_element[els[0]] == [float(x[an.mesh.node_to_index[p] * d + c])
                                                        for p in els[0].points for c in range(d)]
        # component accessors of the 2D / 3D classes (the reference's index of each component)
        names = (['axial_strain_x', 'axial_strain_y', 'shear_strain_xy'] if d == 2 else
                 ['axial_strain_x', 'axial_strain_y', 'axial_strain_z', 'shear_strain_xy',
                  'shear_strain_yz', 'shear_strain_zx'])
        for k, name in enumerate(names):
            for which, ref_arr in (('strain', f_ref), ('stress', s_ref)):
                comp = getattr(r, name.replace('strain', which))()
                got = np.array([comp[el] for el in els])
                scale = np.maximum(np.abs(ref_arr).max(axis=1), 1e-300)
                assert np.max(np.abs(got - ref_arr[:, k]) / scale) < 1e-12
        u = x[:n].reshape(-1, d)
        assert r.displacement_per_node_x() == u[:, 0].tolist()
        assert r.displacement_per_node_y() == u[:, 1].tolist()
        if d == 3:
            assert r.displacement_per_node_z() == u[:, 2].tolist()
