"""Element formulas of the reference vectorised over elements with numpy (test infrastructure only).

Same formulas as oracle.reference (finite_elements/elements.py:395-536, 719-857, 93-118), evaluated
for all elements at once, and the reference's COO -> CSR (finite_elements/analysis.py:714-735, 661).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sparse
import scipy.sparse.linalg


def _tri_pts(coords, conn):
    p = np.asarray(coords, dtype=np.float64)[np.asarray(conn)]
    return p[:, :, 0], p[:, :, 1]


def tri_area(coords, conn):
    x, y = _tri_pts(coords, conn)
    return np.abs((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])) / 2


def tri_b_matrix(coords, conn):
    """(E, 3, 6) B of finite_elements/elements.py:395-416."""
    x, y = _tri_pts(coords, conn)
    yc = [y[:, i] - y[:, j] for (i, j) in [(1, 2), (2, 0), (0, 1)]]
    xc = [x[:, i] - x[:, j] for (i, j) in [(2, 1), (0, 2), (1, 0)]]
    det = (x[:, 0] - x[:, 2]) * (y[:, 1] - y[:, 2]) - (y[:, 0] - y[:, 2]) * (x[:, 1] - x[:, 2])
    B = np.zeros((x.shape[0], 3, 6))
    for a in range(3):
        B[:, 0, 2 * a] = yc[a]
        B[:, 1, 2 * a + 1] = xc[a]
        B[:, 2, 2 * a] = xc[a]
        B[:, 2, 2 * a + 1] = yc[a]
    return (1 / det)[:, None, None] * B


def tri_d(E, nu, plane_strain):
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    if plane_strain:
        a = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    else:
        a = (E * nu) / (1 - nu ** 2)
    b = E / (2 * (1 + nu))
    c = a + 2 * b
    D = np.zeros(E.shape + (3, 3))
    D[..., 0, 0] = D[..., 1, 1] = c
    D[..., 0, 1] = D[..., 1, 0] = a
    D[..., 2, 2] = b
    return D


def tri_elastic_ke(coords, conn, E, nu, thickness, plane_strain):
    """(E, 36) t*A*B^T D B, row-major."""
    B = tri_b_matrix(coords, conn)
    D = tri_d(E, nu, plane_strain)
    A = tri_area(coords, conn)
    K = np.einsum('eki,ekl,elj->eij', B, D, B)
    return ((np.asarray(thickness) * A)[:, None, None] * K).reshape(-1, 36)


def tri_mass(coords, conn, rho, thickness):
    from .reference import TRI_MASS_PATTERN
    f = (np.asarray(rho) * tri_area(coords, conn) * np.asarray(thickness)) / 12
    return (f[:, None, None] * TRI_MASS_PATTERN[None]).reshape(-1, 36)


def tri_grads(coords, conn):
    """(E, 3, 2) gradients (b_i, c_i) of the linear shape functions (inverse Vandermonde columns)."""
    p = np.asarray(coords, dtype=np.float64)[np.asarray(conn)]
    V = np.concatenate([np.ones(p.shape[:2] + (1,)), p], axis=2)
    inv = np.linalg.inv(V)  # (E, 3, 3); column i = (a_i, b_i, c_i)
    return np.transpose(inv[:, 1:3, :], (0, 2, 1))


def magnetic_ke(coords, conn, mu):
    g = tri_grads(coords, conn)
    A = tri_area(coords, conn)
    K = np.einsum('eid,ejd->eij', g, g)
    return ((1 / np.asarray(mu))[:, None, None] * K * A[:, None, None]).reshape(-1, 9)


def tet_volume(coords, conn):
    p = np.asarray(coords, dtype=np.float64)[np.asarray(conn)]
    return np.abs(np.einsum('ei,ei->e', np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])) / 6


def tet_grads(coords, conn):
    """(E, 4, 3) true gradients of the linear shape functions (cofactor form / 6V)."""
    p = np.asarray(coords, dtype=np.float64)[np.asarray(conn)]
    V = np.concatenate([np.ones(p.shape[:2] + (1,)), p], axis=2)
    inv = np.linalg.inv(V)
    return np.transpose(inv[:, 1:4, :], (0, 2, 1))


def tet_b_matrix(coords, conn):
    g = tet_grads(coords, conn)
    B = np.zeros((g.shape[0], 6, 12))
    for i in range(4):
        a, b, c = g[:, i, 0], g[:, i, 1], g[:, i, 2]
        B[:, 0, 3 * i] = a
        B[:, 1, 3 * i + 1] = b
        B[:, 2, 3 * i + 2] = c
        B[:, 3, 3 * i], B[:, 3, 3 * i + 1] = b, a
        B[:, 4, 3 * i + 1], B[:, 4, 3 * i + 2] = c, b
        B[:, 5, 3 * i], B[:, 5, 3 * i + 2] = c, a
    return B


def tet_d(E, nu):
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    coeff = E / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros(E.shape + (6, 6))
    for i in range(3):
        for j in range(3):
            D[..., i, j] = coeff * (1 - nu) if i == j else coeff * nu
    for i in range(3, 6):
        D[..., i, i] = coeff * ((1 - 2 * nu) / 2)
    return D


def tet_elastic_ke(coords, conn, E, nu, chunk=200_000):
    """(E, 144) V*B^T D B, row-major, evaluated in chunks."""
    conn = np.asarray(conn)
    E = np.broadcast_to(np.asarray(E, dtype=np.float64), (conn.shape[0],))
    nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), (conn.shape[0],))
    out = np.empty((conn.shape[0], 144))
    for s in range(0, conn.shape[0], chunk):
        c = conn[s:s + chunk]
        B = tet_b_matrix(coords, c)
        D = tet_d(E[s:s + chunk], nu[s:s + chunk])
        K = np.einsum('eki,ekl,elj->eij', B, D, B)
        out[s:s + chunk] = (tet_volume(coords, c)[:, None, None] * K).reshape(-1, 144)
    return out


def tet_mass(coords, conn, rho):
    from .reference import TET_MASS_PATTERN
    f = (np.asarray(rho) * tet_volume(coords, conn)) / 20
    return (f[:, None, None] * TET_MASS_PATTERN[None]).reshape(-1, 144)


def element_matrices(kind, coords, conn, materials, plane_mode=0, what='stiffness'):
    mats = {k: np.broadcast_to(np.asarray(v, dtype=np.float64), (len(conn),)) for k, v in materials.items()}
    if what == 'mass':
        if kind == 'tri_elastic':
            return tri_mass(coords, conn, mats['rho'], mats['thickness'])
        if kind == 'tet_elastic':
            return tet_mass(coords, conn, mats['rho'])
        raise AttributeError('magnetic elements have no mass matrix')
    if kind == 'tri_elastic':
        return tri_elastic_ke(coords, conn, mats['E'], mats['nu'], mats['thickness'], plane_mode == 1)
    if kind == 'tet_elastic':
        return tet_elastic_ke(coords, conn, mats['E'], mats['nu'])
    return magnetic_ke(coords, conn, mats['mu'])


def coo_indices(conn, dof):
    """Row/col index arrays in the order of get_row_col_indices (finite_elements/analysis.py:714-735)."""
    conn = np.asarray(conn, dtype=np.int64)
    g = (conn[:, :, None] * dof + np.arange(dof)[None, None, :]).reshape(conn.shape[0], -1)
    m = g.shape[1]
    row = np.repeat(g, m, axis=1).reshape(-1)
    col = np.tile(g, (1, m)).reshape(-1)
    return row, col


def assemble_csr(kind, coords, conn, materials, plane_mode=0, what='stiffness'):
    """scipy csr_matrix((data, (row, col))) of the element matrices -- the reference's call."""
    dof = {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}[kind]
    ke = element_matrices(kind, coords, conn, materials, plane_mode, what)
    row, col = coo_indices(conn, dof)
    n = np.asarray(coords).shape[0] * dof
    return sparse.csr_matrix((ke.reshape(-1), (row, col)), shape=(n, n))


def reduced_solve(K, f, fixed_dofs, fixed_values=None):
    """Direct FP64 solve of the eliminated system K_ff u_f = f_f - K_fc g (SuperLU, default ordering)."""
    n = K.shape[0]
    fixed = np.zeros(n, dtype=bool)
    fixed[np.asarray(fixed_dofs, dtype=np.int64)] = True
    g = np.zeros(n)
    if fixed_values is not None:
        g[np.asarray(fixed_dofs, dtype=np.int64)] = fixed_values
    free = ~fixed
    K = K.tocsr()
    rhs = np.asarray(f, dtype=np.float64)[free] - K[free][:, fixed] @ g[fixed]
    u = g.copy()
    u[free] = scipy.sparse.linalg.spsolve(K[free][:, free].tocsc(), rhs)
    return u


def kkt_solve(K, f, fixed_dofs, fixed_values=None, continuity=()):
    """The reference's Lagrange formulation solved directly (finite_elements/analysis.py:175-199,
    241-279, 617-663, 798-830) for large scalar systems: [[K, C^T], [C, 0]] [u; lambda] = [f; g; 0]
    with one row per fixed DOF (u_p = g_p) and, after them, one per continuity condition
    (u_node1 - value u_node2 = 0).  SuperLU with its default ordering (the reference's NATURAL
    ordering does not finish at this size).  Returns (u, lambda of the fixed DOFs, lambda of the
    conditions)."""
    n = K.shape[0]
    fixed_dofs = np.asarray(fixed_dofs, dtype=np.int64)
    nb = fixed_dofs.size
    cont = np.asarray(continuity, dtype=np.float64).reshape(-1, 3)
    nc = cont.shape[0]
    s, m, v = cont[:, 0].astype(np.int64), cont[:, 1].astype(np.int64), cont[:, 2]
    rows = np.concatenate([np.arange(nb), nb + np.arange(nc), nb + np.arange(nc)])
    cols = np.concatenate([fixed_dofs, s, m])
    vals = np.concatenate([np.ones(nb), np.ones(nc), -v])
    C = sparse.csr_matrix((vals, (rows, cols)), shape=(nb + nc, n))
    A = sparse.bmat([[K, C.T], [C, None]], format='csc')
    g = np.zeros(nb) if fixed_values is None else np.asarray(fixed_values, dtype=np.float64)
    rhs = np.concatenate([np.asarray(f, dtype=np.float64), g, np.zeros(nc)])
    x = scipy.sparse.linalg.spsolve(A, rhs)
    return x[:n], x[n:n + nb], x[n + nb:]


def dirichlet_matrix(K, fixed_dofs):
    """K with fixed rows/columns replaced by identity (the eliminated system the solver sees)."""
    n = K.shape[0]
    keep = np.ones(n)
    keep[np.asarray(fixed_dofs, dtype=np.int64)] = 0.0
    Dk = sparse.diags(keep)
    return (Dk @ K @ Dk + sparse.diags(1.0 - keep)).tocsr()


def element_fields(kind, coords, conn, materials, u, plane_mode=0):
    """Per-element fields of a solution u (interleaved DOFs), as the reference's results.py computes
    them: strain = B u_e, stress = D B u_e (results.py:810-830), energy = 1/2 u_e^T Ke u_e
    (elements.py:275-292, results.py:738-770) with u_e the element's DOFs in vertex-major order
    (results.py:689-715); magnetic: B = (sum c_i A_i, -sum b_i A_i) (results.py:121-152).
    Returns (strain (E, ns), stress (E, ns) or None, energy (E,) or None)."""
    conn = np.asarray(conn, dtype=np.int64)
    u = np.asarray(u, dtype=np.float64)
    mats = {k: np.broadcast_to(np.asarray(v, dtype=np.float64), (len(conn),)) for k, v in materials.items()}
    if kind == 'tri_magnetic':
        g = tri_grads(coords, conn)  # (E, 3, 2): (b_i, c_i)
        a = u[conn]
        B = np.stack([np.einsum('ei,ei->e', g[:, :, 1], a), -np.einsum('ei,ei->e', g[:, :, 0], a)], 1)
        return B, None, None
    dof = 2 if kind == 'tri_elastic' else 3
    ue = u[(conn[:, :, None] * dof + np.arange(dof)).reshape(len(conn), -1)]
    if kind == 'tri_elastic':
        Bm = tri_b_matrix(coords, conn)
        D = tri_d(mats['E'], mats['nu'], plane_mode == 1)
    else:
        Bm = tet_b_matrix(coords, conn)
        D = tet_d(mats['E'], mats['nu'])
    strain = np.einsum('eij,ej->ei', Bm, ue)
    stress = np.einsum('eij,ej->ei', D, strain)
    m = ue.shape[1]
    ke = element_matrices(kind, coords, conn, materials, plane_mode).reshape(-1, m, m)
    energy = 0.5 * np.einsum('ei,eij,ej->e', ue, ke, ue)
    return strain, stress, energy
