"""CPU oracle for the element-matrix / assembly / solve path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import this
package, and only as the checker or the timed CPU baseline -- never as part of the product path.

* ``oracle.reference``  -- per-element restatement of Dessia-tech/finite_elements v0.2.0
  (finite_elements/elements.py, analysis.py, core.py, loads.py, conditions.py) in plain Python/numpy,
  with the reference's own scipy calls (COO->CSR at analysis.py:661, spsolve NATURAL at 820-822,
  eigsh at 779-782).  Sized for small meshes; it is also the CPU baseline.
* ``oracle.vectorized`` -- the same element formulas vectorised over elements with numpy, and the same
  COO->CSR, for the larger parity cases.

Pinning (see tests/test_oracle_pins.py): every element formula is checked against what the mechanics
fixes independently of the formula -- the energy of constant-strain states (u^T K v = V eps(u)^T D
eps(v) for all linear displacement fields, which span every element's DOF space), rigid-body null
spaces, textbook Lame/plane-stress constitutive matrices, exact integrals of shape-function products
by quadrature, closed forms for right isosceles triangles, total mass; the COO->CSR pattern is the
reference's scipy call itself; loads/conditions lowering against hand-worked cases.

Not pinnable here (the reference's geometry dependency volmdlr is absent): volmdlr's tet
``form_functions`` normalisation and sign and its area/volume sign conventions.  The restatement uses
the cofactor form (6V * grad N, unsigned V) and unsigned areas, the only choices for which the
reference's formulas produce a positive semi-definite stiffness; they are pinned by the mechanics
checks above, not by volmdlr itself.
"""
