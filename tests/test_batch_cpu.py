"""Host side of the batched path: concatenation of independent meshes into one block-diagonal system
(node offsets, connectivity shift, DOF shift of conditions and loads)."""

import numpy as np
import scipy.sparse as sp

from finite_elements_amd import meshgen
from finite_elements_amd.batch import concatenate
from oracle import vectorized as vz


def test_concatenate_is_block_diagonal():
    plates = meshgen.plates(count=5, seed=1)
    kind, pm, coords, conn, mats, fixed, vals, f, offs = concatenate(plates)
    assert kind == 'tri_elastic' and pm == 2
    assert offs[0] == 0 and offs[-1] == coords.shape[0]
    assert np.array_equal(np.diff(offs), [p.coords.shape[0] for p in plates])
    # the combined stiffness is the block diagonal of the per-plate stiffness matrices
    K = vz.assemble_csr(kind, coords, conn, mats, plane_mode=pm)
    blocks = [vz.assemble_csr(p.kind, p.coords, p.conn, p.materials, plane_mode=p.plane_mode)
              for p in plates]
    ref = sp.block_diag(blocks, format='csr')
    assert (K - ref).count_nonzero() == 0 or abs(K - ref).max() == 0.0
    # conditions and loads land on the shifted DOFs
    for g, p in enumerate(plates):
        d0 = 2 * offs[g]
        assert np.array_equal(f[d0:d0 + 2 * p.coords.shape[0]], p.f)
    ref_fixed = np.concatenate([p.fixed_dofs + 2 * o for p, o in zip(plates, offs[:-1])])
    assert np.array_equal(fixed, ref_fixed)
    assert vals.size == fixed.size


def test_aspect_ratio_classes():
    """The precision classes of SplitBatch come from each mesh's bounding box: a plate of nx x ny
    square cells has aspect ratio max(nx, ny) / min(nx, ny) (the 4096 plates: 2,063 below 4, 2,033
    at or above)."""
    from finite_elements_amd.batch import aspect_ratio
    plates = meshgen.plates(count=64, seed=3)
    for pr in plates:
        nx, ny = pr.meta['nx'], pr.meta['ny']
        assert np.isclose(aspect_ratio(pr), max(nx, ny) / min(nx, ny))
    cube = meshgen.cube_elasticity(n=4)
    assert aspect_ratio(cube) == 1.0
