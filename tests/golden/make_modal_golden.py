"""Golden eigenvalues of the 48^3 modal cube (variant seeds given on the command line) from the oracle:
the 20 lowest eigenvalues of the clamped pencil by scipy eigsh in shift-invert mode (oracle/eigen.py).

    python tests/golden/make_modal_golden.py 5000 [5001 ...]

Writes tests/golden/modal48_eigs.npz (seeds, eigenvalues).  Takes minutes and several GB per variant
(sparse LU of the 346k-DOF free stiffness).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from finite_elements_amd import meshgen  # noqa: E402
from oracle import eigen, vectorized as vz  # noqa: E402


def main():
    seeds = [int(s) for s in sys.argv[1:]] or [5000]
    n = int(os.environ.get('FE_MODAL_N', '48'))
    out = []
    for sd in seeds:
        t0 = time.time()
        pr = meshgen.cube_modal(n=n, seed=sd)
        K = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, what='stiffness')
        M = vz.assemble_csr(pr.kind, pr.coords, pr.conn, pr.materials, what='mass')
        vals, _ = eigen.smallest_clamped(K, M, pr.fixed_dofs, 20)
        out.append(vals)
        print(f'seed {sd}: {time.time() - t0:.0f} s, {vals[:4]}', flush=True)
        # written after every variant so that a long run keeps what it has finished
        path = os.path.join(ROOT, 'tests', 'golden', f'modal{n}_eigs.npz')
        np.savez(path, seeds=np.array(seeds[:len(out)]), eigenvalues=np.array(out), n=np.int64(n))
        print('wrote', path, flush=True)


if __name__ == '__main__':
    main()
