"""Structured meshes and the benchmark problem definitions (inputs only; no compute on the path).

* ``structured_tri``: nx x ny quads, each split along the (i,j)-(i+1,j+1) diagonal into the CCW
  triangles [(i,j),(i+1,j),(i+1,j+1)] and [(i,j),(i+1,j+1),(i,j+1)]; node (i,j) -> j*(nx+1)+i.
* ``kuhn_tets``: n^3 unit-cube cells, each split into the 6 tets sharing the cell diagonal
  (0,0,0)-(1,1,1); node (i,j,k) -> i + (n+1) j + (n+1)^2 k.
* ``cantilever_2d`` / ``magnetostatic_2d`` / ``cube_elasticity`` / ``cube_modal`` / ``plates``:
  the plane-stress cantilever, the iron-annulus magnetostatics, the lognormal-field tet elasticity
  cube, the perturbed-stiffness modal cube and the batched uneven plates.  Each returns a ``Problem``
  with arrays: coords, conn, materials, fixed DOFs and values, and the load vector.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

MU0 = 4 * math.pi * 1e-7


@dataclass
class Problem:
    name: str
    kind: str
    coords: np.ndarray
    conn: np.ndarray
    materials: dict
    plane_mode: int = 0
    fixed_dofs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    fixed_values: Optional[np.ndarray] = None
    f: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def dof(self):
        return {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}[self.kind]

    @property
    def n(self):
        return self.coords.shape[0] * self.dof


def structured_tri(nx, ny, lx=1.0, ly=1.0, x0=0.0, y0=0.0):
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing='xy')
    coords = np.stack([x0 + lx * i.ravel() / nx, y0 + ly * j.ravel() / ny], axis=1)
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing='xy')
    ci, cj = ci.ravel(), cj.ravel()
    n00 = cj * (nx + 1) + ci
    n10 = n00 + 1
    n11 = n00 + nx + 2
    n01 = n00 + nx + 1
    t1 = np.stack([n00, n10, n11], axis=1)
    t2 = np.stack([n00, n11, n01], axis=1)
    conn = np.stack([t1, t2], axis=1).reshape(-1, 3).astype(np.int32)
    return coords, conn


def kuhn_tets(n, length=1.0):
    m = n + 1
    g = np.arange(m) * (length / n)
    x, y, z = np.meshgrid(g, g, g, indexing='ij')
    # node (i,j,k) -> i + m j + m^2 k
    coords = np.stack([x.transpose(2, 1, 0).ravel(), y.transpose(2, 1, 0).ravel(),
                       z.transpose(2, 1, 0).ravel()], axis=1)
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
    ci, cj, ck = (a.transpose(2, 1, 0).ravel() for a in (ci, cj, ck))  # cell order: i fastest

    def node(di, dj, dk):
        return (ci + di) + m * (cj + dj) + m * m * (ck + dk)

    units = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    tets = []
    for a, b in itertools.permutations(range(3), 2):
        ea, eb = units[a], units[b]
        p1 = ea
        p2 = tuple(ea[k] + eb[k] for k in range(3))
        tets.append(np.stack([node(0, 0, 0), node(*p1), node(*p2), node(1, 1, 1)], axis=1))
    conn = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return coords, conn


def last_write_wins(nodes, dims, values):
    """Collapse (node, dim) keys the way the reference's load/BC dicts do (analysis.py:28-110).

    Returns keys in first-insertion order with the last written value.
    """
    nodes = np.asarray(nodes, dtype=np.int64)
    dims = np.asarray(dims, dtype=np.int64)
    values = np.asarray(values, dtype=np.float64)
    key = nodes * 8 + dims
    _, first = np.unique(key, return_index=True)
    _, last_rev = np.unique(key[::-1], return_index=True)
    last = len(key) - 1 - last_rev
    order = np.argsort(first, kind='stable')
    return nodes[first[order]], dims[first[order]], values[last[order]]


def gaussian_random_field(n, corr_len, seed, embed=2):
    """Unit-variance Gaussian field on n^3 cell centres (spacing 1/n) with covariance
    exp(-r^2 / (2 corr_len^2)), by spectral filtering of white noise on a periodic (embed*n)^3 grid."""
    N = embed * n
    h = 1.0 / n
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((N, N, N))
    k = 2 * np.pi * np.fft.fftfreq(N, d=h)
    kx, ky, kz = np.meshgrid(k, k, k, indexing='ij')
    spec = np.exp(-0.5 * corr_len ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    field_ = np.real(np.fft.ifftn(np.sqrt(spec) * np.fft.fftn(white)))[:n, :n, :n]
    field_ = (field_ - field_.mean()) / field_.std()
    return field_  # indexed [i, j, k]


# ------------------------------------------------------------------------------------------------
# benchmark problems
# ------------------------------------------------------------------------------------------------
def cantilever_2d(nx=40, ny=10, lx=1.0, ly=0.25, thickness=0.01, E=210e9, nu=0.3, load=-1e4):
    """Plane-stress cantilever: clamp x=0 (both components), point load on the right-mid node."""
    coords, conn = structured_tri(nx, ny, lx, ly)
    ne = conn.shape[0]
    mats = {'E': np.full(ne, E), 'nu': np.full(ne, nu), 'thickness': np.full(ne, thickness),
            'rho': np.full(ne, 7850.0)}
    left = np.nonzero(coords[:, 0] == 0.0)[0]
    fixed = np.sort(np.concatenate([2 * left, 2 * left + 1]))
    f = np.zeros(coords.shape[0] * 2)
    tip = (ny // 2) * (nx + 1) + nx
    f[2 * tip + 1] = load
    return Problem('cantilever_2d', 'tri_elastic', coords, conn, mats, plane_mode=2,
                   fixed_dofs=fixed, fixed_values=np.zeros(fixed.size), f=f,
                   meta={'tip_node': int(tip), 'nx': nx, 'ny': ny})


def magnetostatic_2d(n=450, mu_iron=1000.0, r_in=0.30, r_out=0.45, current=1e6,
                     consistent_loads=False):
    """Iron annulus + two rectangular coils on [-1,1]^2, A = 0 on the outer boundary.

    Loads follow the reference's ``ElementsLoad`` lowering (value*A_e/sum A per element, factors
    (0, A, A), last write wins per node) unless ``consistent_loads`` (J*A/3 per vertex, summed).
    """
    coords, conn = structured_tri(n, n, 2.0, 2.0, -1.0, -1.0)
    cen = coords[conn].mean(axis=1)
    r = np.hypot(cen[:, 0], cen[:, 1])
    mur = np.where((r >= r_in) & (r <= r_out), mu_iron, 1.0)
    mats = {'mu': MU0 * mur}
    p = coords[conn]
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                        (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    coil_pos = (cen[:, 0] >= -0.25) & (cen[:, 0] <= -0.15) & (np.abs(cen[:, 1]) <= 0.1)
    coil_neg = (cen[:, 0] >= 0.15) & (cen[:, 0] <= 0.25) & (np.abs(cen[:, 1]) <= 0.1)
    N = coords.shape[0]
    f = np.zeros(N)
    if consistent_loads:
        for sel, J in ((coil_pos, current), (coil_neg, -current)):
            idx = np.nonzero(sel)[0]
            np.add.at(f, conn[idx].ravel(), np.repeat(J * area[idx] / 3.0, 3))
    else:
        nodes_l, vals_l = [], []
        for sel, J in ((coil_pos, current), (coil_neg, -current)):
            idx = np.nonzero(sel)[0]
            per = J * area[idx] / area[idx].sum()
            # factors (0, A, A): vertices 2 and 3 of each element, in element order
            for e_i, e in enumerate(idx.tolist()):
                for a, fac in ((0, 0.0), (1, area[e]), (2, area[e])):
                    nodes_l.append(conn[e, a])
                    vals_l.append(per[e_i] * fac)
        kn, _, kv = last_write_wins(nodes_l, np.ones(len(nodes_l)), vals_l)
        f[kn] += kv
    x, y = coords[:, 0], coords[:, 1]
    bnd = np.nonzero((x == -1.0) | (x == 1.0) | (y == -1.0) | (y == 1.0))[0]
    return Problem('magnetostatic_2d', 'tri_magnetic', coords, conn, mats, fixed_dofs=bnd,
                   fixed_values=np.zeros(bnd.size), f=f,
                   meta={'n': n, 'consistent_loads': consistent_loads})


def top_face_loads(coords, conn, traction, z_top=1.0):
    """Consistent nodal loads of a uniform traction on the tet faces lying on z = z_top (A/3 each)."""
    N = coords.shape[0]
    on_top = np.isclose(coords[:, 2], z_top)
    f = np.zeros(N * 3)
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    for fa in faces:
        tri = conn[:, fa]
        sel = on_top[tri].all(axis=1)
        t = tri[sel]
        p = coords[t]
        area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        for d in range(3):
            if traction[d] != 0.0:
                np.add.at(f, 3 * t.ravel() + d, np.repeat(traction[d] * area / 3.0, 3))
    return f


def cube_elasticity(n=64, sigma=1.0, corr_len=0.1, E_mean=1e11, nu=0.3, traction=-1e6, seed=4):
    """Kuhn-tet cube, lognormal random-field Young's modulus, z=0 clamped, traction on z=1."""
    coords, conn = kuhn_tets(n)
    g = gaussian_random_field(n, corr_len, seed)
    E_cell = E_mean * np.exp(sigma * g - sigma ** 2 / 2)  # [i, j, k]
    E_cell_flat = E_cell.transpose(2, 1, 0).ravel()  # cell order i fastest
    E = np.repeat(E_cell_flat, 6)
    ne = conn.shape[0]
    mats = {'E': E, 'nu': np.full(ne, nu), 'rho': np.full(ne, 7850.0)}
    bottom = np.nonzero(coords[:, 2] == 0.0)[0]
    fixed = np.sort(np.concatenate([3 * bottom, 3 * bottom + 1, 3 * bottom + 2]))
    f = top_face_loads(coords, conn, (0.0, 0.0, traction))
    return Problem('cube_elasticity', 'tet_elastic', coords, conn, mats, fixed_dofs=fixed,
                   fixed_values=np.zeros(fixed.size), f=f, meta={'n': n, 'sigma': sigma, 'seed': seed})


def cube_modal(n=48, seed=5000, E0=210e9, nu=0.3, rho=7850.0, spread=1.2):
    """Kuhn-tet cube with E_e = E0 * exp(ln(spread) Z_e), consistent mass, z=0 clamped."""
    coords, conn = kuhn_tets(n)
    ne = conn.shape[0]
    rng = np.random.default_rng(seed)
    E = E0 * np.exp(math.log(spread) * rng.standard_normal(ne))
    mats = {'E': E, 'nu': np.full(ne, nu), 'rho': np.full(ne, rho)}
    bottom = np.nonzero(coords[:, 2] == 0.0)[0]
    fixed = np.sort(np.concatenate([3 * bottom, 3 * bottom + 1, 3 * bottom + 2]))
    return Problem('cube_modal', 'tet_elastic', coords, conn, mats, fixed_dofs=fixed,
                   fixed_values=np.zeros(fixed.size), f=None, meta={'n': n, 'seed': seed})


def plates(count=4096, seed=3, h=0.01, thickness=0.01):
    """Uneven plane-stress plates: nx ~ U{20..200}, ny ~ U{5..50}, lognormal E, nu ~ U[0.25, 0.35],
    x = 0 clamped, unit tip load at node (nx, ny//2) in a random direction.  Returns a list of
    Problems (one per plate)."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(count):
        nx = int(rng.integers(20, 201))
        ny = int(rng.integers(5, 51))
        coords, conn = structured_tri(nx, ny, nx * h, ny * h)
        ne = conn.shape[0]
        E = 1.4e11 * np.exp(0.5 * rng.standard_normal(ne))
        nu = rng.uniform(0.25, 0.35, ne)
        theta = rng.uniform(0, 2 * np.pi)
        mats = {'E': E, 'nu': nu, 'thickness': np.full(ne, thickness), 'rho': np.full(ne, 7850.0)}
        left = np.nonzero(coords[:, 0] == 0.0)[0]
        fixed = np.sort(np.concatenate([2 * left, 2 * left + 1]))
        f = np.zeros(coords.shape[0] * 2)
        tip = (ny // 2) * (nx + 1) + nx
        f[2 * tip] = math.cos(theta)
        f[2 * tip + 1] = math.sin(theta)
        out.append(Problem(f'plate{p}', 'tri_elastic', coords, conn, mats, plane_mode=2,
                           fixed_dofs=fixed, fixed_values=np.zeros(fixed.size), f=f,
                           meta={'nx': nx, 'ny': ny}))
    return out
