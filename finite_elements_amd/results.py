"""Result containers (finite_elements/results.py:23-54, 57-152, 490-830).

``Result(mesh, result_vector)`` is what ``solve()`` returns.  The elasticity and magnetic result
classes read the first n entries of ``result_vector`` through the interleaved DOF map, as the
reference's do.  Per-element strain, stress, energy and the magnetic field are computed on the GPU
by ``fe_element_fields`` (one thread per element); the dictionaries keyed by element objects that
the reference returns are built from those arrays.  ``generate_vtk_file`` writes the displacement
field in the reference's legacy-VTK layout (finite_elements_amd.vtk).  Plotting is not part of this
package.
"""

from __future__ import annotations

from typing import List

import numpy as np

from .mesh import ArrayMesh


def _device_fields(mesh, kind, u, plane_mode):
    """(fields (ns, E), stress or None, energy or None) of u on the GPU, as numpy arrays."""
    from . import device
    from .analysis import mesh_arrays
    _, coords, conn, mats = mesh_arrays(mesh)
    ctx = device.Context.get(0)
    dm = device.DeviceMesh(ctx, kind, coords, conn, mats, plane_mode)
    out = device.element_fields(dm, ctx.tensor(np.ascontiguousarray(u, dtype=np.float64)))
    return tuple(None if t is None else t.cpu().numpy() for t in out)


def _elements(mesh):
    return [el for group in mesh.elements_groups for el in group.elements]


class Result:
    def __init__(self, mesh, result_vector: List[float]):
        self.mesh = mesh
        self.result_vector = result_vector

    @property
    def dimension(self):
        if isinstance(self.mesh, ArrayMesh):
            return {'tri_elastic': 2, 'tet_elastic': 3, 'tri_magnetic': 1}[self.mesh.element_kind]
        return self.mesh.elements_groups[0].elements[0].dimension


class ElasticityResults(Result):
    def __init__(self, mesh, result_vector, plane_strain=None, plane_stress=None):
        super().__init__(mesh, result_vector)
        self.plane_strain = plane_strain
        self.plane_stress = plane_stress
        self._displacement_vectors_per_node = None
        self._fields = None
        self._strain = self._stress = self._energy_per_element = None
        self._displacements_per_element = None

    def displacements_array(self):
        """(N, dim) displacements: node i's entries are result_vector[i*dim : (i+1)*dim]."""
        d = self.dimension
        n = len(self.mesh.nodes)
        return np.asarray(self.result_vector[:n * d], dtype=np.float64).reshape(n, d)

    @property
    def displacement_vectors_per_node(self):
        """{node: displacement tuple} (finite_elements/results.py:651-674)."""
        if self._displacement_vectors_per_node is None:
            arr = self.displacements_array()
            self._displacement_vectors_per_node = {node: tuple(arr[i]) for i, node in
                                                   enumerate(self.mesh.nodes)}
        return self._displacement_vectors_per_node


    def generate_vtk_file(self, file_name_output: str) -> str:
        """Mesh + displacement field as a legacy-ASCII VTK file (finite_elements/results.py:832-871):
        '.vtk' is appended when missing; POINT_DATA holds Displacement_Magnitude and
        Displacement_Vectors (2D vectors padded with 0), one entry per mesh node in node order."""
        from . import vtk
        from .analysis import mesh_arrays
        if file_name_output[-3:] != 'vtk':
            file_name_output += '.vtk'
        _, coords, conn, _ = mesh_arrays(self.mesh)
        name = getattr(self.mesh, 'name', '') or 'mesh'
        return vtk.write_vtk(file_name_output, coords, conn, self.displacements_array(), title=name)

    # ---- per-element fields (finite_elements/results.py:689-830) ---------------------------
    def element_fields(self):
        """(strain (E, ns), stress (E, ns), energy (E,)) arrays, Voigt order (tri: exx, eyy, gxy;
        tet: exx, eyy, ezz, gxy, gyz, gzx), elements group by group."""
        if self._fields is None:
            from .elements import _plane_mode
            kind = {2: 'tri_elastic', 3: 'tet_elastic'}[self.dimension]
            mode = _plane_mode(self.plane_strain, self.plane_stress) if kind == 'tri_elastic' else 0
            n = len(self.mesh.nodes) * self.dimension
            strain, stress, energy = _device_fields(self.mesh, kind, self.result_vector[:n], mode)
            self._fields = (strain.T.copy(), stress.T.copy(), energy)
        return self._fields

    @property
    def displacements_per_element(self):
        """{element: [u of vertex 0 (dim values), vertex 1, ...]} (results.py:689-715)."""
        if self._displacements_per_element is None:
            idx = self.mesh.node_to_index
            d = self.dimension
            u = self.result_vector
            self._displacements_per_element = {
                el: [float(u[idx[p] * d + c]) for p in el.points for c in range(d)]
                for el in _elements(self.mesh)}
        return self._displacements_per_element

    @property
    def strain(self):
        """{element: strain vector B u_e} (results.py:784-830)."""
        if self._strain is None:
            self._strain = dict(zip(_elements(self.mesh), self.element_fields()[0]))
        return self._strain

    @property
    def stress(self):
        """{element: stress vector D B u_e} (results.py:784-830)."""
        if self._stress is None:
            self._stress = dict(zip(_elements(self.mesh), self.element_fields()[1]))
        return self._stress

    @property
    def energy_per_element(self):
        """{element: 1/2 u_e^T Ke u_e} (results.py:753-770)."""
        if self._energy_per_element is None:
            self._energy_per_element = {el: float(w) for el, w in
                                        zip(_elements(self.mesh), self.element_fields()[2])}
        return self._energy_per_element

    @property
    def energy(self):
        """Total strain energy, the sum of the per-element energies (results.py:716-751)."""
        return float(np.sum(self.element_fields()[2]))


    def _component(self, which: str, k: int):
        """{element: component k of the element's strain or stress vector}."""
        arr = self.element_fields()[0 if which == 'strain' else 1][:, k]
        return {el: float(v) for el, v in zip(_elements(self.mesh), arr)}

    def displacement_per_node_x(self):
        """[u_x of every node] in node order (results.py:618-626)."""
        return self.displacements_array()[:, 0].tolist()

    def displacement_per_node_y(self):
        """[u_y of every node] in node order (results.py:628-636)."""
        return self.displacements_array()[:, 1].tolist()


class ElasticityResults2D(ElasticityResults):
    """Plane strain / stress components per element (finite_elements/results.py:930-1466): the
    strain and stress vectors are (xx, yy, xy)."""

    def axial_strain_x(self):
        return self._component('strain', 0)

    def axial_strain_y(self):
        return self._component('strain', 1)

    def shear_strain_xy(self):
        return self._component('strain', 2)

    def axial_stress_x(self):
        return self._component('stress', 0)

    def axial_stress_y(self):
        return self._component('stress', 1)

    def shear_stress_xy(self):
        return self._component('stress', 2)


class ElasticityResults3D(ElasticityResults):
    """Solid components per element (finite_elements/results.py:1468-1740): the strain and stress
    vectors are (xx, yy, zz, xy, yz, zx)."""

    def axial_strain_x(self):
        return self._component('strain', 0)

    def axial_strain_y(self):
        return self._component('strain', 1)

    def axial_strain_z(self):
        return self._component('strain', 2)

    def shear_strain_xy(self):
        return self._component('strain', 3)

    def shear_strain_yz(self):
        return self._component('strain', 4)

    def shear_strain_zx(self):
        return self._component('strain', 5)

    def axial_stress_x(self):
        return self._component('stress', 0)

    def axial_stress_y(self):
        return self._component('stress', 1)

    def axial_stress_z(self):
        return self._component('stress', 2)

    def shear_stress_xy(self):
        return self._component('stress', 3)

    def shear_stress_yz(self):
        return self._component('stress', 4)

    def shear_stress_zx(self):
        return self._component('stress', 5)

    def displacement_per_node_z(self):
        """[u_z of every node] in node order (results.py:1585-1593)."""
        return self.displacements_array()[:, 2].tolist()


class MagneticResults(Result):
    """B = (sum c_i A_i, -sum b_i A_i) per element (finite_elements/results.py:121-152), computed on
    the GPU by ``fe_element_fields``."""

    def __init__(self, mesh, result_vector):
        super().__init__(mesh, result_vector)
        self._field = None
        self.magnetic_field_per_element = self._magnetic_field_per_element()

    def magnetic_field_array(self):
        """(E, 2) field B per element, elements group by group."""
        if self._field is None:
            n = len(self.mesh.nodes)
            b, _, _ = _device_fields(self.mesh, 'tri_magnetic', self.result_vector[:n], 0)
            self._field = b.T.copy()
        return self._field

    def _magnetic_field_per_element(self):
        return {el: (float(b[0]), float(b[1])) for el, b in
                zip(_elements(self.mesh), self.magnetic_field_array())}

    def magnetic_field_norm(self):
        """|B| per element (results.py:108-119)."""
        return list(np.hypot(*self.magnetic_field_array().T))
