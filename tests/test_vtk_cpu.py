"""VTK export (finite_elements/results.py:832-871) and the Gmsh MSH reader, on the CPU.

Fixtures in tests/golden/gmsh/ are pairs of files the reference ships (scripts/InputFiles): a mesh
in Gmsh's MSH 4.1 format and the VTK file Gmsh itself wrote for the same mesh.  Gmsh's VTK is an
independent statement of the mesh, so the MSH reader is pinned against it (points and the
triangle / tetrahedron connectivity), and the writer is pinned against its layout.
"""

import os

import numpy as np
import pytest

from finite_elements_amd import gmsh, vtk
from finite_elements_amd.mesh import ArrayMesh
from finite_elements_amd.results import ElasticityResults
from tests.conftest import GOLDEN

PAIRS = ['beam_2d_0.5', 'beam_2d_0.8', 'beam3d_0.5', 'beam3d_1']


def _path(name, ext):
    return os.path.join(GOLDEN, 'gmsh', f'{name}.{ext}')


@pytest.mark.parametrize('name', PAIRS)
def test_msh_reader_matches_gmsh_vtk(name):
    m = gmsh.read_msh(_path(name, 'msh'))
    v = vtk.read_vtk(_path(name, 'vtk'))
    assert v['dataset'] == 'UNSTRUCTURED_GRID'
    np.testing.assert_array_equal(m['coords'], v['points'])
    types = v['cell_types']
    for vtype, key in ((vtk.VTK_TRIANGLE, 'triangles'), (vtk.VTK_TETRA, 'tetrahedra')):
        cells = np.array([c for c, t in zip(v['cells'], types) if t == vtype], dtype=np.int64)
        cells = cells.reshape(-1, m[key].shape[1])
        np.testing.assert_array_equal(cells, m[key])
    assert (types == vtk.VTK_TETRA).any() == name.startswith('beam3d')


@pytest.mark.parametrize('name', ['beam_2d_0.5', 'beam3d_1'])
def test_writer_layout_matches_gmsh(name, tmp_path):
    """Header lines, section keywords and their order as in Gmsh's file, then the point data the
    reference appends."""
    m = gmsh.read_msh(_path(name, 'msh'))
    three_d = name.startswith('beam3d')
    cells = m['tetrahedra'] if three_d else m['triangles']
    coords = m['coords'] if three_d else m['coords'][:, :2]
    rng = np.random.default_rng(0)
    u = rng.standard_normal((coords.shape[0], coords.shape[1]))
    out = vtk.write_vtk(str(tmp_path / 'u.vtk'), coords, cells, u, title=name)
    with open(out) as fh:
        ours = [ln.rstrip('\n') for ln in fh]
    with open(_path(name, 'vtk')) as fh:
        theirs = [ln.rstrip('\n') for ln in fh]
    assert ours[0] == theirs[0] == '# vtk DataFile Version 2.0'
    assert ours[2] == theirs[2] == 'ASCII'
    assert ours[3] == theirs[3] == 'DATASET UNSTRUCTURED_GRID'
    assert ours[4].split()[0::2] == theirs[4].split()[0::2] == ['POINTS', 'double']

    def keys(lines):
        return [ln.split()[0] for ln in lines if ln and ln.split()[0] in
                ('POINTS', 'CELLS', 'CELL_TYPES', 'POINT_DATA', 'SCALARS', 'LOOKUP_TABLE', 'VECTORS')]
    assert keys(ours)[:3] == keys(theirs)[:3] == ['POINTS', 'CELLS', 'CELL_TYPES']
    assert keys(ours)[3:] == ['POINT_DATA', 'SCALARS', 'LOOKUP_TABLE', 'VECTORS']
    i = ours.index(f'POINT_DATA {coords.shape[0]}')
    assert ours[i + 1] == 'SCALARS Displacement_Magnitude float 1'
    assert ours[i + 2] == 'LOOKUP_TABLE default'
    assert ours[i + 3 + coords.shape[0]] == 'VECTORS Displacement_Vectors float'
    # a 2D vector gets a zero third component, as the reference writes it
    last = ours[i + 4 + coords.shape[0]].split()
    assert len(last) == 3 and (three_d or last[2] == '0')


@pytest.mark.parametrize('dim', [2, 3])
def test_vtk_round_trip(dim, tmp_path):
    rng = np.random.default_rng(dim)
    coords = rng.random((30, dim))
    cells = rng.integers(0, 30, size=(17, dim + 1))
    u = rng.standard_normal((30, dim))
    path = vtk.write_vtk(str(tmp_path / 'x.vtk'), coords, cells, u)
    back = vtk.read_vtk(path)
    np.testing.assert_array_equal(back['points'][:, :dim], coords)
    assert not back['points'][:, dim:].any()
    assert [list(c) for c in back['cells']] == cells.tolist()
    assert set(back['cell_types'].tolist()) == {vtk.VTK_TRIANGLE if dim == 2 else vtk.VTK_TETRA}
    np.testing.assert_array_equal(back['point_data']['Displacement_Magnitude'], np.linalg.norm(u, axis=1))
    np.testing.assert_array_equal(back['point_data']['Displacement_Vectors'][:, :dim], u)
    assert not back['point_data']['Displacement_Vectors'][:, dim:].any()


def test_results_generate_vtk_file(tmp_path):
    """ElasticityResults.generate_vtk_file: '.vtk' appended, node-order displacements from the
    interleaved result vector (multiplier tail ignored)."""
    m = gmsh.read_msh(_path('beam3d_1', 'msh'))
    mesh = ArrayMesh(m['coords'], m['tetrahedra'], 'tet_elastic', {'E': 2e11, 'nu': 0.3, 'rho': 7850.0})
    n = m['coords'].shape[0]
    x = np.arange(3 * n + 5, dtype=np.float64) * 1e-3
    res = ElasticityResults(mesh, list(x), plane_strain=None, plane_stress=None)
    path = res.generate_vtk_file(str(tmp_path / 'beam'))
    assert path.endswith('beam.vtk')
    back = vtk.read_vtk(path)
    np.testing.assert_array_equal(back['point_data']['Displacement_Vectors'], x[:3 * n].reshape(n, 3))
    np.testing.assert_array_equal(back['points'], m['coords'])
    assert len(back['cells']) == m['tetrahedra'].shape[0]
