"""Legacy-ASCII VTK export of a solved mesh (finite_elements/results.py:832-871).

The reference's ``ElasticityResults.generate_vtk_file`` copies the mesh file that Gmsh wrote
(``# vtk DataFile Version 2.0`` / title / ``ASCII`` / ``DATASET UNSTRUCTURED_GRID`` / ``POINTS n double``
/ ``CELLS m size`` / ``CELL_TYPES m``, as in scripts/InputFiles/*.vtk) and appends the displacement
field per node:

    POINT_DATA n
    SCALARS Displacement_Magnitude float 1
    LOOKUP_TABLE default
    |u_i|            (one line per node)
    VECTORS Displacement_Vectors float
    ux uy uz         (2D: ux uy 0)

``write_vtk`` writes the same sections from arrays: the points (3 coordinates, z = 0 in 2D), the
mesh's elements as cells (triangles: type 5, tetrahedra: type 10) and, when given, the same point
data, with values printed as Python's ``repr`` of the float (what ``str()`` of a float gives in the
reference).  Gmsh's file also lists the vertex and line entities of the geometry with a
``CellEntityIds`` cell field; a mesh built from arrays has no such entities, so only the elements
are written.  ``read_vtk`` parses the sections back (for tests and for Gmsh's own files).
"""

from __future__ import annotations

from typing import Optional

import numpy as np

VTK_VERTEX = 1
VTK_LINE = 3
VTK_TRIANGLE = 5
VTK_TETRA = 10
_CELL_TYPE = {3: VTK_TRIANGLE, 4: VTK_TETRA, 2: VTK_LINE, 1: VTK_VERTEX}


def _num(v: float) -> str:
    return repr(float(v))


def write_vtk(path: str, coords, cells, displacements=None, title: str = 'finite_elements_amd'):
    """Write an unstructured grid: coords (N, 2 or 3), cells (E, npe) 0-based node indices,
    displacements (N, 2 or 3) or None.  Returns the path written."""
    coords = np.asarray(coords, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.int64)
    n, dim = coords.shape
    npe = cells.shape[1] if cells.ndim == 2 else 0
    if npe not in _CELL_TYPE:
        raise ValueError(f'unsupported cell with {npe} nodes')
    if cells.size and (cells.min() < 0 or cells.max() >= n):
        raise ValueError('cell node index out of range')
    pts = np.zeros((n, 3))
    pts[:, :dim] = coords
    lines = ['# vtk DataFile Version 2.0', f'{title}, Created by finite_elements_amd', 'ASCII',
             'DATASET UNSTRUCTURED_GRID', f'POINTS {n} double']
    lines.extend(' '.join(_num(v) for v in p) for p in pts)
    lines.append('')
    m = cells.shape[0]
    lines.append(f'CELLS {m} {m * (npe + 1)}')
    lines.extend(f'{npe} ' + ' '.join(str(int(v)) for v in c) for c in cells)
    lines.append('')
    lines.append(f'CELL_TYPES {m}')
    lines.extend([str(_CELL_TYPE[npe])] * m)
    lines.append('')
    if displacements is not None:
        u = np.asarray(displacements, dtype=np.float64)
        if u.shape[0] != n:
            raise ValueError('one displacement per point expected')
        lines.append(f'POINT_DATA {n}')
        lines.append('SCALARS Displacement_Magnitude float 1')
        lines.append('LOOKUP_TABLE default')
        lines.extend(_num(v) for v in np.linalg.norm(u, axis=1))
        lines.append('VECTORS Displacement_Vectors float')
        if u.shape[1] == 2:
            lines.extend(f'{_num(a)} {_num(b)} 0' for a, b in u)
        else:
            lines.extend(' '.join(_num(v) for v in row) for row in u)
    with open(path, 'w', encoding='utf-8') as fh:
        fh.write('\n'.join(lines))
        fh.write('\n')
    return path


def read_vtk(path: str) -> dict:
    """Sections of a legacy-ASCII unstructured grid: title, points (N, 3), cells (list of index
    lists), cell_types (M,), and point_data {name: array} for SCALARS / VECTORS fields."""
    with open(path, encoding='utf-8') as fh:
        tokens_lines = [ln.rstrip('\n') for ln in fh]
    if not tokens_lines or not tokens_lines[0].startswith('# vtk DataFile'):
        raise ValueError('not a legacy VTK file')
    out = {'title': tokens_lines[1], 'format': tokens_lines[2].strip(), 'dataset': None,
           'point_data': {}, 'cell_data': {}}
    i = 3
    n_pts = 0
    section: Optional[str] = None

    def words(k):
        return tokens_lines[k].split()

    while i < len(tokens_lines):
        w = words(i)
        if not w:
            i += 1
            continue
        key = w[0]
        if key == 'DATASET':
            out['dataset'] = w[1]
            i += 1
        elif key == 'POINTS':
            n_pts = int(w[1])
            vals = []
            i += 1
            while len(vals) < 3 * n_pts:
                vals.extend(float(v) for v in words(i))
                i += 1
            out['points'] = np.asarray(vals).reshape(n_pts, 3)
        elif key == 'CELLS':
            m = int(w[1])
            cells = []
            i += 1
            while len(cells) < m:
                row = [int(v) for v in words(i)]
                i += 1
                if row:
                    cells.append(row[1:1 + row[0]])
            out['cells'] = cells
        elif key == 'CELL_TYPES':
            m = int(w[1])
            types = []
            i += 1
            while len(types) < m:
                types.extend(int(v) for v in words(i))
                i += 1
            out['cell_types'] = np.asarray(types, dtype=np.int64)
        elif key in ('POINT_DATA', 'CELL_DATA'):
            section = 'point_data' if key == 'POINT_DATA' else 'cell_data'
            count = int(w[1])
            i += 1
        elif key == 'SCALARS':
            name, ncomp = w[1], int(w[3]) if len(w) > 3 else 1
            i += 1
            if words(i) and words(i)[0] == 'LOOKUP_TABLE':
                i += 1
            vals = []
            while len(vals) < count * ncomp:
                vals.extend(float(v) for v in words(i))
                i += 1
            arr = np.asarray(vals)
            out[section][name] = arr if ncomp == 1 else arr.reshape(count, ncomp)
        elif key == 'VECTORS':
            name = w[1]
            i += 1
            vals = []
            while len(vals) < 3 * count:
                vals.extend(float(v) for v in words(i))
                i += 1
            out[section][name] = np.asarray(vals).reshape(count, 3)
        else:
            i += 1
    return out
