"""Finite-element classes with the reference's names and methods (finite_elements/elements.py).

* ``MagneticElement2D(triangular_element, mu_total)`` -- P1 magnetostatics, 1 DOF per node.
* ``ElasticityTriangularElement2D(mesh_element, E, nu, rho, thickness=1.0)`` -- P1 plane elasticity.
* ``ElasticityTetrahedralElement3D(mesh_element, E, nu, rho)`` -- P1 3D elasticity.

``elementary_matrix`` and ``elementary_mass_matrix`` run the HIP element kernels of libfe_hip.so
(one-element launches; bulk work goes through ``FiniteElementAnalysis``).  ``b_matrix`` and the
``d_matrix_*`` attributes are small host arrays with the reference's formulas, kept for the result
post-processing that reads them.
"""

from __future__ import annotations

import numpy as np

from .mesh import TetrahedralElement, TriangularElement2D


def _plane_mode(plane_strain, plane_stress):
    """The reference's flag check (finite_elements/elements.py:265-273)."""
    if plane_strain and plane_stress:
        raise ValueError('just one of plane_strain or plane_stress can be True')
    if not plane_strain and not plane_stress:
        raise ValueError('one of plane_strain or plane_stress must be True')
    return 1 if plane_strain else 2


def _single_element(kind, points, materials, plane_mode=0, what='stiffness'):
    from . import device
    ctx = device.Context.get()
    coords = np.array([p._c for p in points], dtype=np.float64)
    conn = np.arange(len(points), dtype=np.int32)[None, :]
    dm = device.DeviceMesh(ctx, kind, coords, conn, materials, plane_mode)
    ke = device.element_matrices(dm, what=what, sym=False)
    return ke[:, 0].cpu().numpy()


class Element2D(TriangularElement2D):
    """Triangle with the reference's element-to-node load factors."""

    def element_to_node_factors(self):
        """|detJ| * N_i(midpoint of edge 2-3), as finite_elements/elements.py:18-53 computes it.

        In exact arithmetic this is (0, A, A).
        """
        pts = self._triangle_points()
        (x_1, y_1), (x_2, y_2), (x_3, y_3) = (p._c for p in pts)
        det_jacobien = abs((x_2 - x_1) * (y_3 - y_1) - (x_3 - x_1) * (y_2 - y_1))
        ff = self._triangle_form_functions()
        out = []
        for a, b, c in ff:
            out.append(det_jacobien * (a + 0.5 * b * x_2 + 0.5 * c * y_2 + 0.5 * b * x_3 + 0.5 * c * y_3))
        return tuple(out)

    def _triangle_points(self):
        return self.points

    def _triangle_form_functions(self):
        return self.form_functions


class MagneticElement2D(Element2D):
    """P1 magnetostatic triangle (finite_elements/elements.py:56-191)."""

    def __init__(self, triangular_element, mu_total: float, name: str = ''):
        self.triangular_element = triangular_element
        TriangularElement2D.__init__(self, points=triangular_element.points)
        self.mu_total = mu_total

    @property
    def dimension(self):
        return 1

    def _triangle_points(self):
        return self.triangular_element.points

    def _triangle_form_functions(self):
        return self.triangular_element.form_functions

    def elementary_matrix(self):
        """9-tuple (1/mu)*A*(b_i b_j + c_i c_j), row-major (finite_elements/elements.py:93-118)."""
        ke = _single_element('tri_magnetic', self.triangular_element.points, {'mu': self.mu_total})
        return tuple(float(v) for v in ke)


class ElasticityElement:
    """Common part of the elasticity elements (finite_elements/elements.py:194-330)."""

    def __init__(self, mesh_element, elasticity_modulus, poisson_ratio, mass_density,
                 displacements=None, stress=None, strain=None, name: str = ''):
        self.mesh_element = mesh_element
        self.elasticity_modulus = elasticity_modulus
        self.poisson_ratio = poisson_ratio
        self.mass_density = mass_density
        self.displacements = displacements
        self.stress = stress
        self.strain = strain
        self.name = name
        self._b = None

    @property
    def b_matrix(self):
        if self._b is None:
            self._b = self._b_matrix()
        return self._b

    @property
    def d_matrix_plane_strain(self):
        return self._d_matrix_plane_strain()

    @property
    def d_matrix_plane_stress(self):
        return self._d_matrix_plane_stress()

    def d_matrix(self, plane_strain: bool, plane_stress: bool):
        if _plane_mode(plane_strain, plane_stress) == 1:
            return self.d_matrix_plane_strain
        return self.d_matrix_plane_stress

    def energy(self, plane_strain: bool, plane_stress: bool):
        """0.5 * u^T Ke u for the element's ``displacements`` (finite_elements/elements.py:275-292)."""
        shape = self.dimension * len(self.mesh_element.points)
        ke = np.asarray(self.elementary_matrix(plane_strain, plane_stress)).reshape(shape, shape)
        u = np.asarray(self.displacements, dtype=np.float64)
        return 0.5 * float(u @ ke @ u)

    @classmethod
    def with_material_object(cls, mesh_element, material, displacements=None, stress=None,
                             strain=None, name: str = ''):
        return cls(mesh_element=mesh_element, elasticity_modulus=material.elasticity_modulus,
                   poisson_ratio=material.poisson_ratio, mass_density=material.mass_density,
                   displacements=displacements, stress=stress, strain=strain, name=name)


class ElasticityTriangularElement2D(ElasticityElement, Element2D):
    """P1 plane-elasticity triangle (finite_elements/elements.py:333-660)."""

    def __init__(self, mesh_element, elasticity_modulus, poisson_ratio, mass_density,
                 thickness: float = 1.0, displacements=None, stress=None, strain=None,
                 name: str = ''):
        self.thickness = thickness
        ElasticityElement.__init__(self, mesh_element, elasticity_modulus, poisson_ratio,
                                   mass_density, displacements, stress, strain, name)
        TriangularElement2D.__init__(self, points=mesh_element.points, name=name)

    @property
    def dimension(self):
        return 2

    def _b_matrix(self):
        """B (3x6) of finite_elements/elements.py:395-416."""
        p = self.points
        y = [(p[i].y - p[j].y) for (i, j) in [(1, 2), (2, 0), (0, 1)]]
        x = [(p[i].x - p[j].x) for (i, j) in [(2, 1), (0, 2), (1, 0)]]
        det = (p[0].x - p[2].x) * (p[1].y - p[2].y) - (p[0].y - p[2].y) * (p[1].x - p[2].x)
        data = [y[0], 0, y[1], 0, y[2], 0,
                0, x[0], 0, x[1], 0, x[2],
                x[0], y[0], x[1], y[1], x[2], y[2]]
        return (1 / det) * np.array(data, dtype=np.float64).reshape(3, 6)

    def _d_matrix_plane_strain(self):
        E, nu = self.elasticity_modulus, self.poisson_ratio
        a = (E * nu) / ((1 + nu) * (1 - 2 * nu))
        b = E / (2 * (1 + nu))
        c = a + 2 * b
        return np.array([c, a, 0, a, c, 0, 0, 0, b], dtype=np.float64).reshape(3, 3)

    def _d_matrix_plane_stress(self):
        E, nu = self.elasticity_modulus, self.poisson_ratio
        a = (E * nu) / (1 - nu ** 2)
        b = E / (2 * (1 + nu))
        c = a + 2 * b
        return np.array([c, a, 0, a, c, 0, 0, 0, b], dtype=np.float64).reshape(3, 3)

    def elementary_matrix(self, plane_strain: bool, plane_stress: bool):
        """Flat (36,) t*A*B^T D B, row-major (finite_elements/elements.py:466-511); HIP kernel."""
        mode = _plane_mode(plane_strain, plane_stress)
        return _single_element('tri_elastic', self.points,
                               {'E': self.elasticity_modulus, 'nu': self.poisson_ratio,
                                'thickness': self.thickness}, plane_mode=mode)

    def elementary_mass_matrix(self):
        """Flat (36,) (rho*A*t/12)*P6 (finite_elements/elements.py:513-536); HIP kernel."""
        return _single_element('tri_elastic', self.points,
                               {'rho': self.mass_density, 'thickness': self.thickness,
                                'E': 0.0, 'nu': 0.0}, what='mass')

    @classmethod
    def from_element(cls, mesh_element, elasticity_element):
        return cls(mesh_element, elasticity_element.elasticity_modulus,
                   elasticity_element.poisson_ratio, elasticity_element.mass_density,
                   elasticity_element.thickness)


class ElasticityTetrahedralElement3D(ElasticityElement, TetrahedralElement):
    """P1 tetrahedron (finite_elements/elements.py:663-876)."""

    def __init__(self, mesh_element, elasticity_modulus, poisson_ratio, mass_density,
                 displacements=None, stress=None, strain=None, name: str = ''):
        ElasticityElement.__init__(self, mesh_element, elasticity_modulus, poisson_ratio,
                                   mass_density, displacements, stress, strain, name)
        TetrahedralElement.__init__(self, points=mesh_element.points, name=name)

    @property
    def dimension(self):
        return 3

    def _b_matrix(self):
        """B (6x12) = (1/(6V)) [cofactor rows] (finite_elements/elements.py:719-751)."""
        ff = self.mesh_element.form_functions
        rows = np.zeros((6, 12))
        for i, f in enumerate(ff):
            a, b, c = f[1], f[2], f[3]
            rows[0, 3 * i] = a
            rows[1, 3 * i + 1] = b
            rows[2, 3 * i + 2] = c
            rows[3, 3 * i], rows[3, 3 * i + 1] = b, a
            rows[4, 3 * i + 1], rows[4, 3 * i + 2] = c, b
            rows[5, 3 * i], rows[5, 3 * i + 2] = c, a
        return (1 / (6 * self.mesh_element.volume)) * rows

    def _d_matrix_(self):
        E, nu = self.elasticity_modulus, self.poisson_ratio
        ca, cb = 1 - nu, (1 - 2 * nu) / 2
        data = np.zeros((6, 6))
        data[:3, :3] = nu
        np.fill_diagonal(data[:3, :3], ca)
        data[3, 3] = data[4, 4] = data[5, 5] = cb
        return (E / ((1 + nu) * (1 - 2 * nu))) * data

    def _d_matrix_plane_strain(self):
        return self._d_matrix_()

    def _d_matrix_plane_stress(self):
        return self._d_matrix_()

    def elementary_matrix(self, plane_strain: bool, plane_stress: bool):
        """Flat (144,) V*B^T D B, row-major (finite_elements/elements.py:809-828); HIP kernel."""
        _plane_mode(plane_strain, plane_stress)
        return _single_element('tet_elastic', self.points,
                               {'E': self.elasticity_modulus, 'nu': self.poisson_ratio})

    def elementary_mass_matrix(self):
        """Flat (144,) (rho*V/20)*P12 (finite_elements/elements.py:830-857); HIP kernel."""
        return _single_element('tet_elastic', self.points,
                               {'rho': self.mass_density, 'E': 0.0, 'nu': 0.0}, what='mass')

    @classmethod
    def from_element(cls, mesh_element, elasticity_element):
        return cls(mesh_element, elasticity_element.elasticity_modulus,
                   elasticity_element.poisson_ratio, elasticity_element.mass_density)
