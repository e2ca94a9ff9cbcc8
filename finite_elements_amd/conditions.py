"""Boundary-condition value objects (finite_elements/conditions.py).

Each Dirichlet condition contributes a unit constraint coefficient (``c_matrix() == (1, 1)``) and its
value to the right-hand side (finite_elements/conditions.py:40-58, 89-107).
"""


class BoundaryCondition:
    def __init__(self, application, value: float, dimension, name: str = ''):
        self.application = application
        self.value = value
        self.dimension = dimension
        self.name = name

    def c_matrix(self):
        return (1, 1)

    def source_c_matrix(self):
        return self.value


class NodeBoundaryCondition(BoundaryCondition):
    """Dirichlet value on one node and 1-based ``dimension``."""


class EdgeBoundaryCondition(BoundaryCondition):
    """Dirichlet value on an edge; lowered to its endpoints with ``value * 0.5``."""

    def to_node_boundary_condition(self):
        return [NodeBoundaryCondition(point, self.value, self.dimension)
                for point in [self.application.start, self.application.end]]


class ElementBoundaryCondition(BoundaryCondition):
    """Dirichlet value on an element; lowered with ``value * element_to_node_factors()[i]``."""

    def to_node_boundary_condition(self):
        return [NodeBoundaryCondition(point, self.value, self.dimension)
                for point in self.application.points]


class ContinuityCondition:
    """Periodic / anti-periodic link between two nodes, A(node1) = value * A(node2)
    (finite_elements/conditions.py:196-241): value 1 periodic, -1 anti-periodic.

    The analysis eliminates node1 in favour of node2 (``finite_elements_amd.continuity``); its
    constraint rows follow the Dirichlet rows, where the reference's index
    len(nodes) + len(node_loads) + i (finite_elements/analysis.py:190-197) lands on them.
    """

    def __init__(self, node1, node2, value):
        self.node1 = node1
        self.node2 = node2
        self.value = value

    def c_matrix(self):
        return (1, 1, -self.value, -self.value)

    def source_c_matrix(self):
        return ()
