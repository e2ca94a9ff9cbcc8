"""MI355X-native element-matrix, assembly and linear/eigen-solve path with the API of
Dessia-tech/finite_elements (``FiniteElementAnalysis``, element classes, loads, conditions, results).

Compute runs in hand-written HIP kernels for gfx950 (libfe_hip.so, C ABI in include/fe_hip.h) called
through ctypes; PyTorch-ROCm only provides device memory and the stream.
"""

__version__ = '0.1.0'

from . import core, mesh, loads, conditions, elements, results, analysis  # noqa: F401
