"""Minimal Gmsh MSH 4.1 ASCII reader (nodes, triangles, tetrahedra).

The reference reads meshes with volmdlr's ``GmshParser`` (scripts/Elasticity/beam2d_example_3.py:52-73)
and keeps the file's node order (``mesh.nodes = gmsh.nodes['all_nodes']``).  This reader returns the
nodes in file order and the connectivity as 0-based indices into that order.
"""

from __future__ import annotations

import numpy as np

from .mesh import ArrayMesh

_NPE = {1: 2, 2: 3, 4: 4, 15: 1}


def read_msh(path):
    """Return dict(coords (N,3), node_tags (N,), triangles (T,3), tetrahedra (K,4), lines (L,2))."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    it = iter(range(len(lines)))
    coords, tags = [], []
    elems = {1: [], 2: [], 4: []}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln == '$MeshFormat':
            version = lines[i + 1].split()[0]
            if not version.startswith('4'):
                raise ValueError(f'unsupported MSH version {version}')
            i += 2
        elif ln == '$Nodes':
            nblocks = int(lines[i + 1].split()[0])
            i += 2
            for _ in range(nblocks):
                _, _, param, nb = (int(v) for v in lines[i].split())
                if param:
                    raise ValueError('parametric nodes are not supported')
                i += 1
                tags.extend(int(lines[i + k]) for k in range(nb))
                i += nb
                coords.extend([float(v) for v in lines[i + k].split()[:3]] for k in range(nb))
                i += nb
        elif ln == '$Elements':
            nblocks = int(lines[i + 1].split()[0])
            i += 2
            for _ in range(nblocks):
                _, _, etype, nb = (int(v) for v in lines[i].split())
                i += 1
                if etype in elems:
                    for k in range(nb):
                        elems[etype].append([int(v) for v in lines[i + k].split()[1:1 + _NPE[etype]]])
                i += nb
        else:
            i += 1
    del it
    tags = np.asarray(tags, dtype=np.int64)
    index = {t: k for k, t in enumerate(tags.tolist())}

    def remap(lst, npe):
        if not lst:
            return np.zeros((0, npe), dtype=np.int32)
        return np.asarray([[index[t] for t in row] for row in lst], dtype=np.int32)

    return {'coords': np.asarray(coords, dtype=np.float64), 'node_tags': tags,
            'lines': remap(elems[1], 2), 'triangles': remap(elems[2], 3),
            'tetrahedra': remap(elems[4], 4)}


def array_mesh_from_msh(path, kind, materials):
    """ArrayMesh of the file's triangles ('tri_elastic' / 'tri_magnetic') or tetrahedra ('tet_elastic').

    Nodes not used by any element of that kind are kept (file order is preserved).
    """
    m = read_msh(path)
    if kind == 'tet_elastic':
        return ArrayMesh(m['coords'], m['tetrahedra'], kind, materials)
    return ArrayMesh(m['coords'][:, :2], m['triangles'], kind, materials)
