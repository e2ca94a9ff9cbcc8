"""Load and Dirichlet-condition lowering on arrays (the per-object loops of
finite_elements/analysis.py:28-110, 201-279, 407-577 restated over numpy arrays).

The reference turns every element load, element condition and edge condition into node objects, puts
them in a dictionary keyed by (node, dimension) -- the last write wins, a key keeps the position of
its first insertion -- and then writes one right-hand-side or constraint row per key.  Here the same
items are produced as arrays in the same order (node loads, then element loads element by element
and vertex by vertex, then edge loads; likewise for conditions) and collapsed by ``collapse``, which
reproduces the dictionary's semantics.  Element factors use the reference's arithmetic term by term
(``tri_node_factors``), so the values agree bitwise with the per-element formulas.

``magnet_sources`` lowers a ``MagnetLoad`` (finite_elements/loads.py:105-147, analysis.py:556-577):
the contour edges of the magnet's elements (edges used by exactly one of its elements, except those
with both ends among ``non_contour_nodes``) each add ``M . dl * length / 2`` to both end nodes, with
``dl = (-n_y, n_x)`` for ``n`` the edge's unit normal pointing into its element.
"""

from __future__ import annotations

from typing import Dict, Sequence

import numpy as np


def collapse(nodes, dims, values):
    """Keys (node, dim) in first-insertion order with the last written value (the reference's
    node_loads_to_dict / node_boundary_conditions_to_dict, analysis.py:28-47, 71-90)."""
    nodes = np.asarray(nodes, dtype=np.int64)
    dims = np.asarray(dims, dtype=np.int64)
    values = np.asarray(values, dtype=np.float64)
    if nodes.size == 0:
        return nodes, dims, values
    key = nodes * 4 + dims  # dims are 1..3
    _, first = np.unique(key, return_index=True)
    _, last_rev = np.unique(key[::-1], return_index=True)
    last = key.size - 1 - last_rev
    order = np.argsort(first, kind='stable')
    return nodes[first[order]], dims[first[order]], values[last[order]]


def tri_areas(coords, conn):
    """|signed double area| / 2 per triangle, as TriangularElement2D.area computes it."""
    p = np.asarray(coords, dtype=np.float64)[np.asarray(conn, dtype=np.int64)]
    (x1, y1), (x2, y2), (x3, y3) = ((p[:, i, 0], p[:, i, 1]) for i in range(3))
    return np.abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)) / 2.0


def tri_node_factors(coords, conn):
    """(E, 3) element-to-node load factors of triangles (finite_elements/elements.py:18-53):
    |detJ| * (a_i + b_i x2/2 + c_i y2/2 + b_i x3/2 + c_i y3/2), with the form-function coefficients
    (a_i, b_i, c_i) of vertex i, each term evaluated in the reference's order."""
    p = np.asarray(coords, dtype=np.float64)[np.asarray(conn, dtype=np.int64)]
    pts = [(p[:, i, 0], p[:, i, 1]) for i in range(3)]
    (x1, y1), (x2, y2), (x3, y3) = pts
    det_s = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    det_j = np.abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1))
    out = np.empty((p.shape[0], 3))
    for i in range(3):
        (xj, yj), (xk, yk) = pts[(i + 1) % 3], pts[(i + 2) % 3]
        a = (xj * yk - xk * yj) / det_s
        b = (yj - yk) / det_s
        c = (xk - xj) / det_s
        out[:, i] = det_j * (a + 0.5 * b * x2 + 0.5 * c * y2 + 0.5 * b * x3 + 0.5 * c * y3)
    return out


def element_load_items(coords, conn, elem_idx, value, dim):
    """Node-load items of one ElementsLoad (finite_elements/loads.py:24-35, analysis.py:407-431):
    value * A_e / sum(A) per element (the sum taken element by element), times the element's
    factors, vertex by vertex.  Returns (nodes, dims, values)."""
    elem_idx = np.asarray(elem_idx, dtype=np.int64)
    sub = np.asarray(conn, dtype=np.int64)[elem_idx]
    areas = tri_areas(coords, sub)
    total = np.cumsum(areas)[-1] if areas.size else 0.0  # sequential, as Python's sum()
    per = value * areas / total
    fac = tri_node_factors(coords, sub)
    vals = (per[:, None] * fac).reshape(-1)
    return sub.reshape(-1), np.full(vals.size, dim, dtype=np.int64), vals


def element_condition_items(coords, conn, elem_idx, values, dims):
    """Node-condition items of ElementBoundaryConditions (analysis.py:201-221): value * factor_i per
    vertex, condition by condition.  Returns (nodes, dims, values)."""
    elem_idx = np.asarray(elem_idx, dtype=np.int64)
    sub = np.asarray(conn, dtype=np.int64)[elem_idx]
    fac = tri_node_factors(coords, sub)
    vals = (np.asarray(values, dtype=np.float64)[:, None] * fac).reshape(-1)
    d = np.repeat(np.asarray(dims, dtype=np.int64), sub.shape[1])
    return sub.reshape(-1), d, vals


def magnet_sources(coords, conn, elem_idx, non_contour_nodes: Sequence[int], magnetization):
    """(nodes, values) of one MagnetLoad: per contour edge, in the order edges first appear (element
    by element, edges (v0, v1), (v1, v2), (v2, v0)), its two end nodes each with
    M . dl * length / 2 (finite_elements/analysis.py:556-577).  An edge is on the contour when exactly
    one of the magnet's elements uses it (either orientation) and not both of its ends are listed in
    non_contour_nodes (loads.py:129-147)."""
    coords = np.asarray(coords, dtype=np.float64)
    sub = np.asarray(conn, dtype=np.int64)[np.asarray(elem_idx, dtype=np.int64)]
    if sub.size == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(0)
    a = sub[:, [0, 1, 2]].reshape(-1)
    b = sub[:, [1, 2, 0]].reshape(-1)
    opp = sub[:, [2, 0, 1]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (int(coords.shape[0]) + 1) + hi
    _, first, counts = np.unique(key, return_index=True, return_counts=True)
    once = first[counts == 1]
    if len(non_contour_nodes):
        inner = np.zeros(coords.shape[0] + 1, dtype=bool)
        inner[np.asarray(non_contour_nodes, dtype=np.int64)] = True
        once = once[~(inner[a[once]] & inner[b[once]])]
    e = np.sort(once)  # first-appearance order of the contour edges
    pa, pb, po = coords[a[e], :2], coords[b[e], :2], coords[opp[e], :2]
    t = pb - pa
    length = np.hypot(t[:, 0], t[:, 1])
    n = np.stack([-t[:, 1], t[:, 0]], axis=1) / length[:, None]
    inward = np.einsum('ij,ij->i', n, po - pa) > 0
    n[~inward] *= -1.0
    dl = np.stack([-n[:, 1], n[:, 0]], axis=1)
    m = np.asarray(magnetization, dtype=np.float64)[:2]
    v = (dl @ m) * length / 2
    nodes = np.stack([a[e], b[e]], axis=1).reshape(-1)
    return nodes, np.repeat(v, 2)


class ElementIndex:
    """Element object -> row of the mesh's connectivity (elements group by group, the order of
    ``analysis.mesh_arrays``)."""

    def __init__(self, mesh):
        self._map: Dict[int, int] = {}
        self._mesh = mesh

    def __call__(self, elements) -> np.ndarray:
        out = np.empty(len(elements), dtype=np.int64)
        for k, el in enumerate(elements):
            idx = getattr(el, 'mesh_index', None)
            if idx is None:
                if not self._map:
                    e = 0
                    for group in self._mesh.elements_groups:
                        for obj in group.elements:
                            self._map[id(obj)] = e
                            e += 1
                idx = self._map.get(id(el))
                if idx is None:
                    raise ValueError('element is not part of the analysis mesh')
            out[k] = idx
        return out
