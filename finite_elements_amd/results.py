"""Result containers (finite_elements/results.py:23-54, 57-152, 490-714).

``Result(mesh, result_vector)`` is what ``solve()`` returns.  The elasticity and magnetic result
classes read the first n entries of ``result_vector`` through the interleaved DOF map, as the
reference's do; plotting and VTK export are not part of this package.
"""

from __future__ import annotations

from typing import List

import numpy as np

from .mesh import ArrayMesh


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


class ElasticityResults2D(ElasticityResults):
    pass


class ElasticityResults3D(ElasticityResults):
    pass


class MagneticResults(Result):
    """B = (sum c_i A_i, -sum b_i A_i) per element (finite_elements/results.py:121-152)."""

    def __init__(self, mesh, result_vector):
        super().__init__(mesh, result_vector)
        self.magnetic_field_per_element = self._magnetic_field_per_element()

    def _magnetic_field_per_element(self):
        out = {}
        idx = self.mesh.node_to_index
        a = self.result_vector
        for group in self.mesh.elements_groups:
            for element in group.elements:
                ff = element.form_functions
                ids = [idx[p] for p in element.points]
                bx = sum(ff[k][2] * a[ids[k]] for k in range(3))
                by = -sum(ff[k][1] * a[ids[k]] for k in range(3))
                out[element] = (float(bx), float(by))
        return out
