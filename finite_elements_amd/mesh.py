"""Array-backed mesh layer with the attribute names the analysis API expects.

The reference builds meshes from ``volmdlr.mesh`` objects (``Node2D``, ``TriangularElement2D``,
``TetrahedralElement``, ``ElementsGroup``, ``Mesh``; used at finite_elements/elements.py:7,77,391,715
and finite_elements/analysis.py:290,716).  That package is not available here, so this module supplies
the same names and attributes:

* ``Node2D`` / ``Node3D`` (aliases ``Point2D`` / ``Point3D``): value objects compared and hashed by their
  exact coordinates, indexable as ``node[0]`` and readable as ``node.x``.
* ``TriangularElement2D(points)``: ``points``, ``form_functions`` (``(a_i, b_i, c_i)`` with
  ``N_i = a_i + b_i x + c_i y`` and ``N_i(p_j) = delta_ij``), ``area`` (unsigned), ``center``.
* ``TetrahedralElement(points)``: ``points``, ``form_functions`` in cofactor form
  (``(alpha_i, beta_i, gamma_i, delta_i)`` with ``N_i = (alpha_i + beta_i x + gamma_i y + delta_i z)/(6V)``,
  ``V`` the unsigned volume), ``volume`` (unsigned).  The reference's tet B matrix divides these by ``6V``
  (finite_elements/elements.py:749), which therefore yields the true shape-function gradients.
* ``ElementsGroup(elements, name)`` and ``Mesh(elements_groups)`` with ``nodes`` and ``node_to_index``.
  The default node order is the order of first appearance while walking groups, elements and element
  points; callers may overwrite ``nodes`` / ``node_to_index`` (as finite_elements scripts do to keep Gmsh
  file order) and the analysis honours them verbatim.

``ArrayMesh`` is the large-mesh form: coordinates and connectivity live in numpy arrays and no per-node or
per-element Python object is created unless asked for.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np


class _NodeBase:
    __slots__ = ('_c',)

    def __getitem__(self, i):
        return self._c[i]

    def __iter__(self):
        return iter(self._c)

    def __len__(self):
        return len(self._c)

    def __eq__(self, other):
        if isinstance(other, _NodeBase):
            return self._c == other._c
        return NotImplemented

    def __hash__(self):
        return hash(self._c)

    def __repr__(self):
        return f'{self.__class__.__name__}{self._c}'

    @property
    def x(self):
        return self._c[0]

    @property
    def y(self):
        return self._c[1]

    def __sub__(self, other):
        return tuple(a - b for a, b in zip(self._c, other._c))


class Node2D(_NodeBase):
    """A 2D mesh node (value object, equality and hash by exact coordinates)."""
    __slots__ = ()

    def __init__(self, x, y, name: str = ''):
        self._c = (float(x), float(y))


class Node3D(_NodeBase):
    """A 3D mesh node (value object, equality and hash by exact coordinates)."""
    __slots__ = ()

    def __init__(self, x, y, z, name: str = ''):
        self._c = (float(x), float(y), float(z))

    @property
    def z(self):
        return self._c[2]


Point2D = Node2D
Point3D = Node3D


def _as_node(p):
    if isinstance(p, _NodeBase):
        return p
    p = tuple(p)
    return Node2D(*p) if len(p) == 2 else Node3D(*p)


class TriangularElement2D:
    """Linear (P1) triangle.  ``form_functions[i] = (a_i, b_i, c_i)``, ``N_i = a_i + b_i x + c_i y``."""

    dimension_geometry = 2
    n_points = 3

    def __init__(self, points: Sequence, name: str = ''):
        self.points = [_as_node(p) for p in points]
        self.name = name
        self._ff = None

    def _signed_double_area(self):
        (x1, y1), (x2, y2), (x3, y3) = (p._c for p in self.points)
        return (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)

    @property
    def area(self):
        return abs(self._signed_double_area()) / 2.0

    @property
    def form_functions(self):
        if self._ff is None:
            (x1, y1), (x2, y2), (x3, y3) = (p._c for p in self.points)
            det = self._signed_double_area()
            if det == 0.0:
                raise ZeroDivisionError('degenerate triangle')
            pts = ((x1, y1), (x2, y2), (x3, y3))
            ff = []
            for i in range(3):
                (xj, yj), (xk, yk) = pts[(i + 1) % 3], pts[(i + 2) % 3]
                ff.append(((xj * yk - xk * yj) / det, (yj - yk) / det, (xk - xj) / det))
            self._ff = ff
        return self._ff

    @property
    def center(self):
        return tuple(sum(p._c[k] for p in self.points) / 3.0 for k in range(2))


class TetrahedralElement:
    """Linear (P1) tetrahedron with cofactor-form ``form_functions`` and unsigned ``volume``."""

    dimension_geometry = 3
    n_points = 4

    def __init__(self, points: Sequence, name: str = ''):
        self.points = [_as_node(p) for p in points]
        self.name = name
        self._ff = None

    def _signed_six_volume(self):
        p0, p1, p2, p3 = (np.array(p._c) for p in self.points)
        return float(np.dot(np.cross(p1 - p0, p2 - p0), p3 - p0))

    @property
    def volume(self):
        return abs(self._signed_six_volume()) / 6.0

    @property
    def form_functions(self):
        if self._ff is None:
            vander = np.array([[1.0, *p._c] for p in self.points])
            inv = np.linalg.inv(vander)
            six_v = 6.0 * self.volume
            self._ff = [tuple(six_v * inv[:, i]) for i in range(4)]
        return self._ff

    @property
    def center(self):
        return tuple(sum(p._c[k] for p in self.points) / 4.0 for k in range(3))


class ElementsGroup:
    """A named list of elements (one material region in the reference's scripts)."""

    def __init__(self, elements: List, name: str = ''):
        self.elements = elements
        self.name = name


class Mesh:
    """Mesh of element groups with a node table.

    ``nodes`` defaults to first-appearance order over groups, elements and element points.
    """

    def __init__(self, elements_groups: List[ElementsGroup], name: str = ''):
        self.elements_groups = elements_groups
        self.name = name
        seen = {}
        for group in elements_groups:
            for element in group.elements:
                for p in element.points:
                    if p not in seen:
                        seen[p] = len(seen)
        self.nodes = list(seen.keys())
        self.node_to_index = seen


class ArrayMesh(Mesh):
    """Mesh held as arrays: ``coords`` (N, dim) float64 and ``connectivity`` (E, npe) int.

    ``element_kind`` names the finite-element class of every element ('tri_elastic', 'tet_elastic',
    'tri_magnetic') and ``materials`` holds per-element arrays (``E``, ``nu``, ``rho``, ``thickness``,
    ``mu``), any of which may be a scalar.  Node objects are only built if ``nodes`` is read.
    """

    def __init__(self, coords, connectivity, element_kind: str, materials: dict, name: str = ''):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.connectivity = np.ascontiguousarray(connectivity, dtype=np.int32)
        self.element_kind = element_kind
        self.materials = dict(materials)
        self.name = name
        self._nodes = None
        self._node_to_index = None
        self._groups = None

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    @property
    def n_elements(self):
        return self.connectivity.shape[0]

    @property
    def nodes(self):
        if self._nodes is None:
            cls = Node2D if self.coords.shape[1] == 2 else Node3D
            self._nodes = [cls(*c) for c in self.coords.tolist()]
        return self._nodes

    @nodes.setter
    def nodes(self, value):
        self._nodes = value

    @property
    def node_to_index(self):
        if self._node_to_index is None:
            self._node_to_index = {n: i for i, n in enumerate(self.nodes)}
        return self._node_to_index

    @node_to_index.setter
    def node_to_index(self, value):
        self._node_to_index = value

    def material(self, key, default=None):
        """Per-element array for ``key`` (scalars broadcast to shape (E,))."""
        v = self.materials.get(key, default)
        if v is None:
            return None
        v = np.asarray(v, dtype=np.float64)
        if v.ndim == 0:
            v = np.full(self.n_elements, float(v))
        return np.ascontiguousarray(v)

    @property
    def elements_groups(self):
        """Element objects are built on demand (slow for big meshes; the analysis does not need them)."""
        if self._groups is None:
            from finite_elements_amd import elements as fe_el
            nodes = self.nodes
            els = []
            E = self.material('E')
            nu = self.material('nu')
            rho = self.material('rho', 0.0)
            t = self.material('thickness', 1.0)
            mu = self.material('mu')
            for e, row in enumerate(self.connectivity.tolist()):
                pts = [nodes[i] for i in row]
                if self.element_kind == 'tri_elastic':
                    els.append(fe_el.ElasticityTriangularElement2D(
                        TriangularElement2D(pts), E[e], nu[e], rho[e], t[e]))
                elif self.element_kind == 'tet_elastic':
                    els.append(fe_el.ElasticityTetrahedralElement3D(
                        TetrahedralElement(pts), E[e], nu[e], rho[e]))
                elif self.element_kind == 'tri_magnetic':
                    els.append(fe_el.MagneticElement2D(TriangularElement2D(pts), mu[e]))
                else:
                    raise ValueError(f'unknown element kind {self.element_kind}')
            self._groups = [ElementsGroup(els, '')]
        return self._groups

    @elements_groups.setter
    def elements_groups(self, value):
        self._groups = value
