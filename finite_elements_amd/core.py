"""Core utilities: vacuum permeability, the DOF map and the material record.

Follows finite_elements/core.py:31 (``MU``), 89-108 (``global_matrix_positions``) and 139-171
(``Material``).
"""

import math

import numpy as np

MU = 4 * math.pi * 1e-7


def global_matrix_positions(dimension, nodes_number):
    """Interleaved DOF map ``{(node_index, d): node_index*dimension + d - 1}`` with 1-based ``d``.

    Same mapping as finite_elements/core.py:89-108.
    """
    positions = {}
    count = 0
    for i in range(nodes_number):
        for j in range(dimension):
            positions[(i, j + 1)] = count
            count += 1
    return positions


def dof_index(node_index, dimension_1based, dof_per_node):
    """Vectorised form of the DOF map: ``node_index*dof_per_node + (d - 1)``."""
    return np.asarray(node_index, dtype=np.int64) * dof_per_node + (np.asarray(dimension_1based) - 1)


class Material:
    """Material record (finite_elements/core.py:139-171)."""

    def __init__(self, elasticity_modulus, poisson_ratio, mass_density, name: str = ''):
        self.elasticity_modulus = elasticity_modulus
        self.poisson_ratio = poisson_ratio
        self.mass_density = mass_density
        self.name = name
