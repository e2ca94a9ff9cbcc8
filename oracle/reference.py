"""Per-element CPU restatement of the reference path (test infrastructure only; see oracle/__init__).

Node identity is a coordinate tuple (the reference keys its dictionaries by volmdlr points, which
compare by coordinates).  A ``Problem`` lists nodes in DOF order, elements as vertex tuples with a
material dict, and loads / conditions as tuples.  Functions follow the reference line by line:

  tri elasticity     finite_elements/elements.py:395-453 (B, D), 466-511 (Ke), 513-536 (Me)
  tet elasticity     finite_elements/elements.py:719-796 (B, D), 809-828 (Ke), 830-857 (Me)
  magnetic           finite_elements/elements.py:93-118 (Ke), 156-191 (load factors)
  load factors       finite_elements/elements.py:18-53
  DOF map            finite_elements/core.py:89-108
  row/col indices    finite_elements/analysis.py:714-735
  k/m_matrix_data    finite_elements/analysis.py:324-339, 357-365
  matrix_sparse      finite_elements/analysis.py:387-405
  conditions         finite_elements/analysis.py:28-68, 201-279, 509-543
  loads              finite_elements/analysis.py:71-110, 407-489; loads.py:24-35
  continuity rows    finite_elements/analysis.py:175-199, conditions.py:196-231
  create_matrix      finite_elements/analysis.py:617-663
  create_source      finite_elements/analysis.py:665-708, 737-739
  solve              finite_elements/analysis.py:798-830
  modal 'largest'    finite_elements/analysis.py:741-786
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np
import scipy.sparse as sparse
import scipy.sparse.linalg


# ------------------------------------------------------------------------------------------------
# geometry (the volmdlr quantities the reference reads)
# ------------------------------------------------------------------------------------------------
def tri_form_functions(pts):
    """(a_i, b_i, c_i) with N_i = a_i + b_i x + c_i y, N_i(p_j) = delta_ij: columns of inv([1 x y])."""
    a = np.array([[1.0, p[0], p[1]] for p in pts])
    inv = np.linalg.inv(a)
    return [tuple(inv[:, i]) for i in range(3)]


def tri_area(pts):
    (x1, y1), (x2, y2), (x3, y3) = pts
    return abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)) / 2


def tet_volume(pts):
    p = np.array(pts, dtype=np.float64)
    return abs(np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p[3] - p[0])) / 6


def tet_form_functions(pts):
    """Cofactor form (alpha_i, beta_i, gamma_i, delta_i) = 6V * columns of inv([1 x y z])."""
    a = np.array([[1.0, *p] for p in pts])
    inv = np.linalg.inv(a)
    six_v = 6 * tet_volume(pts)
    return [tuple(six_v * inv[:, i]) for i in range(4)]


# ------------------------------------------------------------------------------------------------
# element formulas
# ------------------------------------------------------------------------------------------------
def check_plane_flags(plane_strain, plane_stress):
    """finite_elements/elements.py:265-273."""
    if plane_strain and plane_stress:
        raise ValueError('just one of plane_strain or plane_stress can be True')
    if not plane_strain and not plane_stress:
        raise ValueError('one of plane_strain or plane_stress must be True')


def tri_b_matrix(pts):
    """finite_elements/elements.py:395-416."""
    y = [(pts[i][1] - pts[j][1]) for (i, j) in [(1, 2), (2, 0), (0, 1)]]
    x = [(pts[i][0] - pts[j][0]) for (i, j) in [(2, 1), (0, 2), (1, 0)]]
    det = (pts[0][0] - pts[2][0]) * (pts[1][1] - pts[2][1]) - \
        (pts[0][1] - pts[2][1]) * (pts[1][0] - pts[2][0])
    data = [y[0], 0, y[1], 0, y[2], 0,
            0, x[0], 0, x[1], 0, x[2],
            x[0], y[0], x[1], y[1], x[2], y[2]]
    return (1 / det) * np.array(data).reshape(3, 6)


def tri_d_plane_strain(E, nu):
    """finite_elements/elements.py:418-434."""
    a = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    b = E / (2 * (1 + nu))
    c = a + 2 * b
    return np.array([c, a, 0, a, c, 0, 0, 0, b]).reshape(3, 3)


def tri_d_plane_stress(E, nu):
    """finite_elements/elements.py:436-453."""
    a = (E * nu) / (1 - nu ** 2)
    b = E / (2 * (1 + nu))
    c = a + 2 * b
    return np.array([c, a, 0, a, c, 0, 0, 0, b]).reshape(3, 3)


def tri_elastic_ke(pts, E, nu, thickness, plane_strain, plane_stress):
    """Flat (36,) t*A*B^T D B (finite_elements/elements.py:466-511)."""
    check_plane_flags(plane_strain, plane_stress)
    b = tri_b_matrix(pts)
    d = tri_d_plane_strain(E, nu) if plane_strain else tri_d_plane_stress(E, nu)
    k = thickness * tri_area(pts) * (np.matmul(np.matmul(b.transpose(), d), b))
    return k.flatten()


TRI_MASS_PATTERN = np.array([2, 0, 1, 0, 1, 0,
                             0, 2, 0, 1, 0, 1,
                             1, 0, 2, 0, 1, 0,
                             0, 1, 0, 2, 0, 1,
                             1, 0, 1, 0, 2, 0,
                             0, 1, 0, 1, 0, 2], dtype=np.float64).reshape(6, 6)


def tri_mass(pts, rho, thickness):
    """finite_elements/elements.py:513-536."""
    return (((rho * tri_area(pts) * thickness) / 12) * TRI_MASS_PATTERN).flatten()


def tet_b_matrix(pts):
    """finite_elements/elements.py:719-751."""
    f = tet_form_functions(pts)
    a = [f[i][1] for i in range(4)]
    b = [f[i][2] for i in range(4)]
    c = [f[i][3] for i in range(4)]
    data = [a[0], 0, 0, a[1], 0, 0, a[2], 0, 0, a[3], 0, 0,
            0, b[0], 0, 0, b[1], 0, 0, b[2], 0, 0, b[3], 0,
            0, 0, c[0], 0, 0, c[1], 0, 0, c[2], 0, 0, c[3],
            b[0], a[0], 0, b[1], a[1], 0, b[2], a[2], 0, b[3], a[3], 0,
            0, c[0], b[0], 0, c[1], b[1], 0, c[2], b[2], 0, c[3], b[3],
            c[0], 0, a[0], c[1], 0, a[1], c[2], 0, a[2], c[3], 0, a[3]]
    return (1 / (6 * tet_volume(pts))) * np.array(data).reshape(6, 12)


def tet_d(E, nu):
    """finite_elements/elements.py:773-796."""
    ca = 1 - nu
    cb = (1 - 2 * nu) / 2
    data = [ca, nu, nu, 0, 0, 0,
            nu, ca, nu, 0, 0, 0,
            nu, nu, ca, 0, 0, 0,
            0, 0, 0, cb, 0, 0,
            0, 0, 0, 0, cb, 0,
            0, 0, 0, 0, 0, cb]
    coeff = (E / ((1 + nu) * (1 - 2 * nu)))
    return coeff * np.array(data).reshape(6, 6)


def tet_elastic_ke(pts, E, nu, plane_strain=True, plane_stress=False):
    """Flat (144,) V*B^T D B (finite_elements/elements.py:809-828); the 2D flags are still checked."""
    check_plane_flags(plane_strain, plane_stress)
    b = tet_b_matrix(pts)
    k = tet_volume(pts) * (np.matmul(np.matmul(b.transpose(), tet_d(E, nu)), b))
    return k.flatten()


def tet_mass_pattern():
    p = np.zeros((12, 12))
    for r in range(12):
        for c in range(12):
            if r % 3 == c % 3:
                p[r, c] = 2.0 if r == c else 1.0
    return p


TET_MASS_PATTERN = tet_mass_pattern()


def tet_mass(pts, rho):
    """finite_elements/elements.py:830-857."""
    return (((rho * tet_volume(pts)) / 20) * TET_MASS_PATTERN).flatten()


def magnetic_ke(pts, mu_total):
    """9-tuple of finite_elements/elements.py:93-118."""
    ff = tri_form_functions(pts)
    b = [ff[i][1] for i in range(3)]
    c = [ff[i][2] for i in range(3)]
    area = tri_area(pts)
    out = []
    for i in range(3):
        for j in range(3):
            out.append(1 / mu_total * (b[i] * b[j] + c[i] * c[j]) * area)
    return tuple(out)


def element_to_node_factors(pts):
    """finite_elements/elements.py:18-53 (and 156-191)."""
    (x_1, y_1), (x_2, y_2), (x_3, y_3) = pts
    det = abs((x_2 - x_1) * (y_3 - y_1) - (x_3 - x_1) * (y_2 - y_1))
    ff = tri_form_functions(pts)
    return tuple(det * (a + 0.5 * b * x_2 + 0.5 * c * y_2 + 0.5 * b * x_3 + 0.5 * c * y_3)
                 for (a, b, c) in ff)


# ------------------------------------------------------------------------------------------------
# problem and analysis
# ------------------------------------------------------------------------------------------------
@dataclass
class Problem:
    """kind: 'tri_elastic' | 'tet_elastic' | 'tri_magnetic'.  nodes: coordinate tuples in DOF order.
    elements: list of (vertex tuples, material dict with E/nu/thickness/rho or mu).
    node_loads / node_bcs: (node, value, dim); element_loads: (element indices, value, dim);
    element_bcs: (element index, value, dim); edge_loads / edge_bcs: ((start, end), value, dim)."""
    kind: str
    nodes: List[Tuple]
    elements: List[Tuple]
    node_loads: List = field(default_factory=list)
    element_loads: List = field(default_factory=list)
    edge_loads: List = field(default_factory=list)
    node_bcs: List = field(default_factory=list)
    element_bcs: List = field(default_factory=list)
    edge_bcs: List = field(default_factory=list)
    plane_strain: Optional[bool] = None
    plane_stress: Optional[bool] = None
    continuity: List = field(default_factory=list)  # (node1, node2, value): u(node1) = value u(node2)

    @property
    def dimension(self):
        return {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}[self.kind]

    def node_to_index(self):
        return {n: i for i, n in enumerate(self.nodes)}


def global_matrix_positions(dimension, nodes_number):
    """finite_elements/core.py:89-108."""
    positions = {}
    count = 0
    for i in range(nodes_number):
        for j in range(dimension):
            positions[(i, j + 1)] = count
            count += 1
    return positions


def element_matrix(problem, element):
    pts, mat = element
    if problem.kind == 'tri_magnetic':
        return list(magnetic_ke(pts, mat['mu']))
    if problem.kind == 'tri_elastic':
        return list(tri_elastic_ke(pts, mat['E'], mat['nu'], mat.get('thickness', 1.0),
                                   problem.plane_strain, problem.plane_stress))
    return list(tet_elastic_ke(pts, mat['E'], mat['nu'], problem.plane_strain, problem.plane_stress))


def element_mass(problem, element):
    pts, mat = element
    if problem.kind == 'tri_elastic':
        return list(tri_mass(pts, mat['rho'], mat.get('thickness', 1.0)))
    if problem.kind == 'tet_elastic':
        return list(tet_mass(pts, mat['rho']))
    raise AttributeError("'MagneticElement2D' object has no attribute 'elementary_mass_matrix'")


def get_row_col_indices(problem, element, positions, node_to_index):
    """finite_elements/analysis.py:714-735."""
    indexes = [node_to_index[p] for p in element[0]]
    d = problem.dimension
    row_ind, col = [], []
    for index in indexes:
        for i in range(d):
            row_ind.extend(len(indexes) * d * [positions[(index, i + 1)]])
            col.append(positions[(index, i + 1)])
    col_ind = col * (len(indexes) * d)
    return row_ind, col_ind


def matrix_data(problem, which='k'):
    """finite_elements/analysis.py:324-339 ('k') and 357-365 ('m')."""
    positions = global_matrix_positions(problem.dimension, len(problem.nodes))
    node_to_index = problem.node_to_index()
    row_ind, col_ind, data = [], [], []
    for element in problem.elements:
        data.extend(element_matrix(problem, element) if which == 'k' else element_mass(problem, element))
        r, c = get_row_col_indices(problem, element, positions, node_to_index)
        row_ind.extend(r)
        col_ind.extend(c)
    return data, row_ind, col_ind


def matrix_sparse(problem, which='k'):
    """Dict sum in element order, then csc_matrix (finite_elements/analysis.py:387-405)."""
    data, row, col = matrix_data(problem, which)
    dict_data = {}
    for i, v in enumerate(data):
        key = (row[i], col[i])
        if key in dict_data:
            dict_data[key] += v
        else:
            dict_data[key] = v
    d, r, c = [], [], []
    for key, value in dict_data.items():
        d.append(value)
        r.append(key[0])
        c.append(key[1])
    return sparse.csc_matrix((d, (r, c)))


def _last_write_wins(items):
    """finite_elements/analysis.py:28-47 / 71-90: key -> +value (assignment), first-insertion order."""
    out = {}
    for node, value, dim in items:
        out[(node, dim)] = + value
    return out


def collapsed_boundary_conditions(problem):
    """finite_elements/analysis.py:249-265: node BCs, then element BCs (value*factor_i), then edge
    BCs (value*0.5), collapsed per (node, dim)."""
    items = list(problem.node_bcs)
    for e_idx, value, dim in problem.element_bcs:
        pts = problem.elements[e_idx][0]
        factors = element_to_node_factors(pts)
        for p_index, point in enumerate(pts):
            items.append((point, value * factors[p_index], dim))
    for (start, end), value, dim in problem.edge_bcs:
        for point in (start, end):
            items.append((point, value * 0.5, dim))
    return [(k[0], v, k[1]) for k, v in _last_write_wins(items).items()]


def collapsed_node_loads(problem):
    """finite_elements/analysis.py:465-475 with loads.py:31-34 and analysis.py:407-447."""
    items = list(problem.node_loads)
    for elem_ids, value, dim in problem.element_loads:
        areas = [tri_area(problem.elements[e][0]) for e in elem_ids]
        total = sum(areas)
        per = [value * a / total for a in areas]
        for j, e in enumerate(elem_ids):
            pts = problem.elements[e][0]
            factors = element_to_node_factors(pts)
            for p_index, point in enumerate(pts):
                items.append((point, per[j] * factors[p_index], dim))
    for (start, end), value, dim in problem.edge_loads:
        for point in (start, end):
            items.append((point, value * 0.5, dim))
    return [(k[0], v, k[1]) for k, v in _last_write_wins(items).items()]


def magnet_sources(elements, non_contour_nodes, magnetization):
    """finite_elements/loads.py:129-147 and analysis.py:556-577, one magnet load: count the edges
    of the magnet's triangles (an edge is the same in either orientation; edges taken (v0, v1),
    (v1, v2), (v2, v0) element by element), keep those used once and not with both ends in
    non_contour_nodes, and give both ends of each M . dl * length / 2 with dl = (-n_y, n_x), n the
    unit normal of the edge pointing into its triangle.  elements: vertex tuples.  Returns
    [(node, value)] in the reference's order."""
    count = {}
    for pts in elements:
        for k in range(3):
            a, b, o = pts[k], pts[(k + 1) % 3], pts[(k + 2) % 3]
            key = frozenset((a, b))
            if key in count:
                count[key][1] += 1
            else:
                count[key] = [(a, b, o), 1]
    out = []
    inner = set(non_contour_nodes)
    for (a, b, o), c in count.values():
        if c != 1 or (a in inner and b in inner):
            continue
        tx, ty = b[0] - a[0], b[1] - a[1]
        length = (tx * tx + ty * ty) ** 0.5
        nx, ny = -ty / length, tx / length
        if nx * (o[0] - a[0]) + ny * (o[1] - a[1]) < 0:
            nx, ny = -nx, -ny
        dl = (-ny, nx)
        v = (magnetization[0] * dl[0] + magnetization[1] * dl[1]) * length / 2
        out.append((a, v))
        out.append((b, v))
    return out


def c_matrix_boundary_conditions(problem, positions, node_to_index):
    """finite_elements/analysis.py:267-279."""
    n = len(problem.nodes) * problem.dimension
    row_ind, col_ind, data = [], [], []
    for i, (node, _, dim) in enumerate(collapsed_boundary_conditions(problem)):
        data.extend((1, 1))
        pos = positions[(node_to_index[node], dim)]
        row_ind.extend((n + i, pos))
        col_ind.extend((pos, n + i))
    return data, row_ind, col_ind


def c_matrix_continuity_conditions(problem, node_to_index):
    """finite_elements/analysis.py:175-199 with conditions.py:223-231: per condition the entries
    (1, 1, -value, -value) at (R, i1), (i1, R), (R, i2), (i2, R).  The reference's row
    R = len(nodes) + len(node_loads) + i lands on Dirichlet rows; here R = n + n_bc + i, after the
    Dirichlet rows, which is the layout of its source vector length n + n_continuity + n_bc
    (analysis.py:737-739).  Scalar problems only (the reference uses node indices as DOF rows)."""
    if problem.continuity and problem.dimension != 1:
        raise ValueError('continuity conditions are defined for scalar problems')
    base = len(problem.nodes) + len(collapsed_boundary_conditions(problem))
    row_ind, col_ind, data = [], [], []
    for i, (node1, node2, value) in enumerate(problem.continuity):
        data.extend((1, 1, -value, -value))
        i1, i2 = node_to_index[node1], node_to_index[node2]
        row_ind.extend((base + i, i1, base + i, i2))
        col_ind.extend((i1, base + i, i2, base + i))
    return data, row_ind, col_ind


def create_matrix(problem):
    """COO lists of K, the continuity entries and the Dirichlet entries -> scipy CSR
    (finite_elements/analysis.py:617-663)."""
    positions = global_matrix_positions(problem.dimension, len(problem.nodes))
    node_to_index = problem.node_to_index()
    data, row, col = matrix_data(problem, 'k')
    for d2, r2, c2 in (c_matrix_continuity_conditions(problem, node_to_index),
                       c_matrix_boundary_conditions(problem, positions, node_to_index)):
        data.extend(d2)
        row.extend(r2)
        col.extend(c2)
    return sparse.csr_matrix((data, (row, col)))


def create_source_matrix(problem):
    """finite_elements/analysis.py:665-708 (loads, then magnet loads (none), then BC values)."""
    positions = global_matrix_positions(problem.dimension, len(problem.nodes))
    node_to_index = problem.node_to_index()
    bcs = collapsed_boundary_conditions(problem)
    n = len(problem.nodes) * problem.dimension
    matrix = np.zeros((n + len(bcs) + len(problem.continuity), 1))
    for node, value, dim in collapsed_node_loads(problem):
        matrix[positions[(node_to_index[node], dim)]][0] += value
    for i, (_, value, _) in enumerate(bcs):
        matrix[n + i][0] += value
    return matrix


def solve(problem):
    """spsolve(K, F, permc_spec='NATURAL') on the augmented system (finite_elements/analysis.py:816-830)."""
    k_sparse = create_matrix(problem)
    force = create_source_matrix(problem)
    x = scipy.sparse.linalg.spsolve(k_sparse, force, permc_spec='NATURAL', use_umfpack=True)
    return list(x)


def modal_largest(problem, k):
    """eigsh(K, M, which='LM', k) on the unconstrained K, M (finite_elements/analysis.py:773-786)."""
    m_matrix = matrix_sparse(problem, 'm')
    k_matrix = matrix_sparse(problem, 'k')
    eigvals, eigvecs = scipy.sparse.linalg.eigsh(A=k_matrix, M=m_matrix, which='LM', k=k)
    return eigvals, eigvecs.T


def problem_from_arrays(kind, coords, conn, materials, plane_strain=None, plane_stress=None,
                        fixed_dofs=(), fixed_values=None, f=None, continuity=()):
    """Oracle problem from arrays: node tuples in index order, per-element materials, DOF-level
    loads (as node loads), fixed DOFs (as node BCs, in the given order) and continuity conditions
    as (node1 index, node2 index, value)."""
    nodes = [tuple(float(v) for v in c) for c in np.asarray(coords)]
    d = {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}[kind]
    elements = []
    mats = {k: np.broadcast_to(np.asarray(v, dtype=np.float64), (len(conn),)) for k, v in materials.items()}
    for e, row in enumerate(np.asarray(conn)):
        elements.append((tuple(nodes[i] for i in row), {k: float(v[e]) for k, v in mats.items()}))
    node_loads = []
    if f is not None:
        for dof in np.nonzero(np.asarray(f))[0].tolist():
            node_loads.append((nodes[dof // d], float(f[dof]), dof % d + 1))
    node_bcs = []
    vals = np.zeros(len(fixed_dofs)) if fixed_values is None else np.asarray(fixed_values)
    for dof, v in zip(np.asarray(fixed_dofs).tolist(), vals.tolist()):
        node_bcs.append((nodes[dof // d], float(v), dof % d + 1))
    cont = [(nodes[int(a)], nodes[int(b)], float(v)) for a, b, v in continuity]
    return Problem(kind, nodes, elements, node_loads=node_loads, node_bcs=node_bcs,
                   plane_strain=plane_strain, plane_stress=plane_stress, continuity=cont)
