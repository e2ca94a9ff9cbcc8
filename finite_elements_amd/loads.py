"""Load value objects (finite_elements/loads.py).

``NodeLoad`` and ``ElementsLoad`` are lowered to the right-hand side by the analysis with the
reference's semantics; ``EdgeLoad`` endpoints take ``value * 0.5`` each.  ``MagnetLoad`` is accepted
for signature compatibility but its contour lowering is not supported by this package.
"""

from typing import List


class ElementsLoad:
    """Load spread over elements in proportion to their area (finite_elements/loads.py:24-35)."""

    def __init__(self, elements: List, value: float, dimension: int):
        self.elements = elements
        self.value = value
        self.dimension = dimension
        total_area = sum(elem.area for elem in self.elements)
        self.value_per_element = [value * element.area / total_area for element in self.elements]


class ElementLoad:
    """Single-element load record (finite_elements/loads.py:38-56; unused by the analysis)."""

    def __init__(self, element, value: float, dimension: int):
        self.element = element
        self.value = value
        self.dimension = dimension


class EdgeLoad:
    """Load on an edge with ``start``/``end`` points (finite_elements/loads.py:59-69)."""

    def __init__(self, edge, value, dimension):
        self.edge = edge
        self.value = value
        self.dimension = dimension


class NodeLoad:
    """Nodal load; ``dimension`` is 1-based (finite_elements/loads.py:72-102)."""

    def __init__(self, node, value: float, dimension):
        self.node = node
        self.value = value
        self.dimension = dimension

    def c_matrix(self):
        return ()

    def source_c_matrix(self):
        return self.value


class MagnetLoad:
    """Magnet load record (finite_elements/loads.py:105-147).  Lowering is not supported here."""

    def __init__(self, elements: List, non_contour_nodes: List, magnetization_vector):
        self.elements = elements
        self.non_contour_nodes = non_contour_nodes
        self.magnetization_vector = magnetization_vector
