"""Pins of the CPU oracle against what mechanics fixes independently of the reference's formulas.

For a linear element every nodal displacement vector is the trace of a linear field, so the element
stiffness is fully determined by (i) zero energy for rigid motions and (ii) the energy of uniform
strain states: u^T K v = (t*A or V) * eps(u)^T D eps(v).  D itself is checked against the textbook
Lame / plane-stress forms, and mass matrices against exact integrals of shape-function products
computed with an independent quadrature rule.  A dropped term, wrong sign, transposed B or wrong
normalisation in the oracle fails one of these.
"""

import itertools
from fractions import Fraction

import numpy as np
import pytest

from oracle import reference as ref
from oracle import vectorized as vec

# meshes of the reference's example scripts
TRI_EX1 = [((3, 0), (3, 2), (0, 0)), ((0, 2), (0, 0), (3, 2))]  # scripts/Elasticity/beam2d_example_1.py:20-21
TET_EX1 = ((2, 3, 4), (6, 3, 2), (2, 5, 1), (4, 3, 6))  # scripts/Elasticity/beam3d_example_1.py:22


def _tri_ex2():
    """18 unit right triangles of scripts/Elasticity/beam2d_example_2.py:37-70."""
    out = []
    for i in range(9):
        out.append(((i, 0), (i + 1, 0), (i, 1)))
        out.append(((i + 1, 1), (i + 1, 0), (i, 1)))
    return out


def cross2(a, b):
    return a[0] * b[1] - a[1] * b[0]


def lame(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return lam, mu


def linear_fields_2d(pts):
    """Nodal vectors of the 6 linear fields: 2 translations, 1 rotation, 3 unit strain states."""
    P = np.array(pts, dtype=np.float64)
    x, y = P[:, 0], P[:, 1]
    rigid = [np.stack([np.ones(3), np.zeros(3)], 1), np.stack([np.zeros(3), np.ones(3)], 1),
             np.stack([-y, x], 1)]
    # strain states (voigt, engineering shear): exx, eyy, gxy
    strains = {(1, 0, 0): np.stack([x, 0 * x], 1), (0, 1, 0): np.stack([0 * x, y], 1),
               (0, 0, 1): np.stack([0.5 * y, 0.5 * x], 1)}
    return [r.reshape(-1) for r in rigid], {k: v.reshape(-1) for k, v in strains.items()}


def linear_fields_3d(pts):
    P = np.array(pts, dtype=np.float64)
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    o, n = np.ones(4), np.zeros(4)
    rigid = [np.stack([o, n, n], 1), np.stack([n, o, n], 1), np.stack([n, n, o], 1),
             np.stack([-y, x, n], 1), np.stack([n, -z, y], 1), np.stack([z, n, -x], 1)]
    strains = {
        (1, 0, 0, 0, 0, 0): np.stack([x, n, n], 1),
        (0, 1, 0, 0, 0, 0): np.stack([n, y, n], 1),
        (0, 0, 1, 0, 0, 0): np.stack([n, n, z], 1),
        (0, 0, 0, 1, 0, 0): np.stack([0.5 * y, 0.5 * x, n], 1),   # gamma_xy
        (0, 0, 0, 0, 1, 0): np.stack([n, 0.5 * z, 0.5 * y], 1),   # gamma_yz
        (0, 0, 0, 0, 0, 1): np.stack([0.5 * z, n, 0.5 * x], 1),   # gamma_zx
    }
    return [r.reshape(-1) for r in rigid], {k: v.reshape(-1) for k, v in strains.items()}


def textbook_d_2d(E, nu, plane_strain):
    if plane_strain:
        lam, mu = lame(E, nu)
        return np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]])
    c = E / (1 - nu ** 2)
    return np.array([[c, nu * c, 0], [nu * c, c, 0], [0, 0, E / (2 * (1 + nu))]])


def textbook_d_3d(E, nu):
    lam, mu = lame(E, nu)
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    return D


@pytest.mark.parametrize('plane_strain', [True, False])
def test_d_matrices_are_textbook(plane_strain):
    E, nu = 210e9, 0.3
    got = ref.tri_d_plane_strain(E, nu) if plane_strain else ref.tri_d_plane_stress(E, nu)
    np.testing.assert_allclose(got, textbook_d_2d(E, nu, plane_strain), rtol=1e-14, atol=0)
    np.testing.assert_allclose(ref.tet_d(E, nu), textbook_d_3d(E, nu), rtol=1e-14, atol=0)


def test_plane_stress_uniaxial():
    # sigma_yy = 0 under uniaxial stress: eps_yy = -nu eps_xx and sigma_xx = E eps_xx
    E, nu = 30e6, 0.25
    D = ref.tri_d_plane_stress(E, nu)
    s = D @ np.array([1.0, -nu, 0.0])
    assert abs(s[1]) < 1e-9 * E and abs(s[0] - E) < 1e-9 * E


@pytest.mark.parametrize('pts,E,nu,t,ps', [
    (TRI_EX1[0], 30e6, 0.25, 0.5, False), (TRI_EX1[1], 30e6, 0.25, 0.5, False),
    (_tri_ex2()[3], 210e9, 0.25, 1.0, True), (((0.1, -0.3), (2.2, 0.4), (-0.7, 1.9)), 1.0, 0.33, 2.0, True),
    (((0.0, 0.0), (0.0, 1.0), (1.0, 0.0)), 5.0, 0.2, 1.0, False),  # clockwise
])
def test_tri_ke_energy_characterisation(pts, E, nu, t, ps):
    K = ref.tri_elastic_ke(pts, E, nu, t, ps, not ps).reshape(6, 6)
    A = 0.5 * abs(cross2(np.subtract(pts[1], pts[0]), np.subtract(pts[2], pts[0])))
    D = textbook_d_2d(E, nu, ps)
    rigid, strains = linear_fields_2d(pts)
    scale = np.abs(K).max()
    for r in rigid:
        assert np.abs(K @ r).max() <= 1e-12 * scale * max(1, np.abs(r).max())
    for (e1, u), (e2, v) in itertools.product(strains.items(), repeat=2):
        want = t * A * np.array(e1) @ D @ np.array(e2)
        assert abs(u @ K @ v - want) <= 1e-12 * abs(t * A * D).max() * max(1.0, np.abs(np.array(pts)).max() ** 2)
    np.testing.assert_allclose(K, K.T, rtol=0, atol=1e-13 * scale)
    assert np.linalg.eigvalsh(K).min() > -1e-10 * scale


def test_tet_ke_energy_characterisation():
    for pts, E, nu in [(TET_EX1, 210e9, 0.25),
                       (((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)), 1.0, 0.3),
                       (((0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, 1)), 2.0, 0.1),  # negative orientation
                       (((0.3, -0.2, 0.1), (1.7, 0.2, -0.4), (0.1, 1.3, 0.5), (0.4, 0.3, 2.1)), 7.0, 0.45)]:
        K = ref.tet_elastic_ke(pts, E, nu).reshape(12, 12)
        P = np.array(pts, dtype=np.float64)
        V = abs(np.linalg.det(np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]]))) / 6
        D = textbook_d_3d(E, nu)
        rigid, strains = linear_fields_3d(pts)
        scale = np.abs(K).max()
        for r in rigid:
            assert np.abs(K @ r).max() <= 1e-11 * scale * max(1, np.abs(r).max())
        for (e1, u), (e2, v) in itertools.product(strains.items(), repeat=2):
            want = V * np.array(e1) @ D @ np.array(e2)
            assert abs(u @ K @ v - want) <= 1e-11 * abs(V * D).max() * max(1.0, np.abs(P).max() ** 2)
        np.testing.assert_allclose(K, K.T, rtol=0, atol=1e-12 * scale)


def test_tet_b_matrix_is_shape_function_gradient():
    # B u for a linear field u = eps*x gives exactly that strain (patch test at element level)
    B = ref.tet_b_matrix(TET_EX1)
    _, strains = linear_fields_3d(TET_EX1)
    for e, u in strains.items():
        np.testing.assert_allclose(B @ u, np.array(e, dtype=float), atol=1e-13)


def test_magnetic_closed_form_right_isosceles():
    for L in (1.0, 0.01, 7.5):
        pts = ((0.0, 0.0), (L, 0.0), (0.0, L))
        mu = 4e-7 * np.pi * 1000
        K = np.array(ref.magnetic_ke(pts, mu)).reshape(3, 3)
        want = (1 / mu) / 2 * np.array([[2, -1, -1], [-1, 1, 0], [-1, 0, 1]])
        np.testing.assert_allclose(K, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


def test_magnetic_energy_characterisation():
    pts = ((0.2, 0.1), (1.9, -0.3), (0.4, 1.7))
    mu = 3.0
    K = np.array(ref.magnetic_ke(pts, mu)).reshape(3, 3)
    P = np.array(pts)
    A = 0.5 * abs(cross2(P[1] - P[0], P[2] - P[0]))
    fields = {(1.0, 0.0): P[:, 0], (0.0, 1.0): P[:, 1]}
    for (g1, u), (g2, v) in itertools.product(fields.items(), repeat=2):
        assert abs(u @ K @ v - (1 / mu) * A * np.dot(g1, g2)) < 1e-13
    assert np.abs(K @ np.ones(3)).max() < 1e-14


def _quad_tri(pts, f):
    # edge-midpoint rule: exact for quadratics
    P = np.array(pts, dtype=np.float64)
    A = 0.5 * abs(cross2(P[1] - P[0], P[2] - P[0]))
    mids = [(P[0] + P[1]) / 2, (P[1] + P[2]) / 2, (P[2] + P[0]) / 2]
    return A / 3 * sum(f(m) for m in mids)


def test_tri_mass_exact_integrals():
    pts = ((0.2, 0.1), (1.9, -0.3), (0.4, 1.7))
    rho, t = 7850.0, 0.01
    M = ref.tri_mass(pts, rho, t).reshape(6, 6)
    ff = ref.tri_form_functions(pts)
    N = [lambda p, c=c: c[0] + c[1] * p[0] + c[2] * p[1] for c in ff]
    for a in range(3):
        for b in range(3):
            want = rho * t * _quad_tri(pts, lambda p: N[a](p) * N[b](p))
            for d in range(2):
                assert abs(M[2 * a + d, 2 * b + d] - want) < 1e-12 * rho * t
                assert M[2 * a + d, 2 * b + 1 - d] == 0
    assert abs(M[0::2, 0::2].sum() - rho * t * 0.5 * abs(cross2(np.subtract(pts[1], pts[0]), np.subtract(pts[2], pts[0])))) < 1e-9


def test_tet_mass_exact_integrals():
    P = np.array(TET_EX1, dtype=np.float64)
    rho = 7860.0
    M = ref.tet_mass(TET_EX1, rho).reshape(12, 12)
    V = ref.tet_volume(TET_EX1)
    # exact: int N_a N_b = V (1 + delta_ab) / 20 (barycentric moment formula a!b!c!d! 3! V/(3+sum)!)
    for a in range(4):
        for b in range(4):
            want = rho * V * (2 if a == b else 1) * 6 / 120
            assert abs(M[3 * a, 3 * b] - want) < 1e-10 * rho * V
    assert abs(M[0::3, 0::3].sum() - rho * V) < 1e-9 * rho * V


def test_element_to_node_factors_identity():
    for pts in TRI_EX1 + _tri_ex2()[:4] + [((0.3, 0.1), (2.0, 0.5), (0.7, 1.9))]:
        A = 0.5 * abs(cross2(np.subtract(pts[1], pts[0]), np.subtract(pts[2], pts[0])))
        f = ref.element_to_node_factors(pts)
        assert abs(f[0]) < 1e-13 * max(1, A) and abs(f[1] - A) < 1e-13 * max(1, A) and abs(f[2] - A) < 1e-13 * max(1, A)


def test_exact_rational_tri_ex1():
    """Exact rational B^T D B on the integer triangle of beam2d_example_1 equals the float oracle."""
    pts = TRI_EX1[0]
    E, nu, t = Fraction(30_000_000), Fraction(1, 4), Fraction(1, 2)
    (x1, y1), (x2, y2), (x3, y3) = [(Fraction(a), Fraction(b)) for a, b in pts]
    det = (x1 - x3) * (y2 - y3) - (y1 - y3) * (x2 - x3)
    gx = [(y2 - y3) / det, (y3 - y1) / det, (y1 - y2) / det]
    gy = [(x3 - x2) / det, (x1 - x3) / det, (x2 - x1) / det]
    c = E / (1 - nu * nu)
    lam, mu = nu * c, E / (2 * (1 + nu))
    A = abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)) / 2
    K = [[Fraction(0)] * 6 for _ in range(6)]
    for a in range(3):
        for b in range(3):
            ga, gb = (gx[a], gy[a]), (gx[b], gy[b])
            for di in range(2):
                for dj in range(2):
                    if di == dj:
                        v = c * ga[di] * gb[dj] + mu * ga[1 - di] * gb[1 - dj]
                    else:
                        v = lam * ga[di] * gb[dj] + mu * ga[dj] * gb[di]
                    K[2 * a + di][2 * b + dj] = t * A * v
    Kf = ref.tri_elastic_ke(pts, 30e6, 0.25, 0.5, False, True).reshape(6, 6)
    Ke = np.array([[float(v) for v in row] for row in K])
    np.testing.assert_allclose(Kf, Ke, rtol=1e-14, atol=1e-14 * np.abs(Ke).max())


def test_vectorized_matches_per_element():
    rng = np.random.default_rng(0)
    coords = rng.uniform(-1, 1, (30, 3))
    conn = np.array([rng.choice(30, 4, replace=False) for _ in range(20)], dtype=np.int32)
    E = rng.uniform(1, 2, 20)
    nu = rng.uniform(0.2, 0.4, 20)
    kv = vec.tet_elastic_ke(coords, conn, E, nu)
    for e in range(20):
        ke = ref.tet_elastic_ke([tuple(coords[i]) for i in conn[e]], E[e], nu[e])
        np.testing.assert_allclose(kv[e], ke, rtol=1e-10, atol=1e-10 * np.abs(ke).max())
    c2 = coords[:, :2]
    conn2 = conn[:, :3]
    for ps in (True, False):
        kt = vec.tri_elastic_ke(c2, conn2, E, nu, 0.3, ps)
        for e in range(20):
            ke = ref.tri_elastic_ke([tuple(c2[i]) for i in conn2[e]], E[e], nu[e], 0.3, ps, not ps)
            np.testing.assert_allclose(kt[e], ke, rtol=1e-10, atol=1e-10 * np.abs(ke).max())
    km = vec.magnetic_ke(c2, conn2, E)
    for e in range(20):
        ke = ref.magnetic_ke([tuple(c2[i]) for i in conn2[e]], E[e])
        np.testing.assert_allclose(km[e], ke, rtol=1e-10, atol=1e-10 * np.abs(ke).max())


def test_flags_raise_like_reference():
    with pytest.raises(ValueError):
        ref.tri_elastic_ke(TRI_EX1[0], 1.0, 0.3, 1.0, True, True)
    with pytest.raises(ValueError):
        ref.tet_elastic_ke(TET_EX1, 1.0, 0.3, False, False)


def test_spsolve_augmented_equals_eliminated():
    """The reference's Lagrange system and the eliminated SPD system give the same displacements, and
    the multipliers equal f_c - (K u)_c."""
    from finite_elements_amd import meshgen
    pr = meshgen.cantilever_2d(8, 2)
    prob = ref.problem_from_arrays(pr.kind, pr.coords, pr.conn, pr.materials, plane_strain=False,
                                   plane_stress=True, fixed_dofs=pr.fixed_dofs, f=pr.f)
    x = np.array(ref.solve(prob))
    n = pr.n
    K = vec.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, 2)
    u = vec.reduced_solve(K, pr.f, pr.fixed_dofs)
    np.testing.assert_allclose(x[:n], u, rtol=1e-10, atol=1e-12 * np.abs(u).max())
    lam = pr.f[pr.fixed_dofs] - (K @ u)[pr.fixed_dofs]
    np.testing.assert_allclose(x[n:], lam, rtol=1e-8, atol=1e-8 * np.abs(lam).max())
