"""Train the learned block-Jacobi preconditioner on the GPU and save its weights.

    python -m finite_elements_amd.train_learned --n 32 --systems 60 --epochs 20 --out weights.npz

Training data: residual/error pairs of plain two-level PCG (block Jacobi + rigid-body coarse
correction; --block-jacobi-only: plain block Jacobi) on n^3 Kuhn-tet cubes with lognormal random-field
Young's modulus (sigma 1, correlation length 0.1; seeds 1000, 1001, ...).  The network learns the
block-Jacobi part (the smoother) with the coarse correction held fixed in the loss.  The forward, loss
gradient, backward and Adam steps are the fe_gnn_* kernels of libfe_hip.so.  After training the
iteration counts with and without the network are compared on held-out seeds (2000+).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'weights', 'learned_block_jacobi_tet.npz')


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--n', type=int, default=32)
    ap.add_argument('--systems', type=int, default=60)
    ap.add_argument('--epochs', type=int, default=20)
    ap.add_argument('--lr', type=float, default=3e-3)
    ap.add_argument('--snapshots', default='3,10,30,90')
    ap.add_argument('--probes', type=int, default=2, help='random-error pairs per training system')
    ap.add_argument('--block-jacobi-only', action='store_true',
                    help='train the network as plain block Jacobi (no coarse correction)')
    ap.add_argument('--loss', default='damped', choices=('damped', 'undamped'))
    ap.add_argument('--eval-n', type=int, default=32)
    ap.add_argument('--eval-seeds', type=int, default=3)
    ap.add_argument('--out', default=DEFAULT_OUT)
    args = ap.parse_args(argv)

    from finite_elements_amd import device, learned, meshgen
    from finite_elements_amd.system import FESystem

    def log(m):
        print(f'[train] {m}', flush=True)

    t0 = time.perf_counter()
    snaps = tuple(int(s) for s in args.snapshots.split(','))
    two = not args.block_jacobi_only
    systems = learned.make_training_systems(n=args.n, seeds=range(1000, 1000 + args.systems),
                                            snapshots=snaps, log=log, two_level=two,
                                            probes=args.probes)
    log(f'{len(systems)} training systems in {time.perf_counter() - t0:.1f} s')
    model = learned.LearnedPreconditioner(3, seed=0)
    t1 = time.perf_counter()
    hist = model.train(systems, epochs=args.epochs, lr=args.lr, log=log, loss=args.loss)
    train_s = time.perf_counter() - t1
    log(f'trained in {train_s:.1f} s; loss {hist[0]:.5f} -> {hist[-1]:.5f}')
    evals = []
    sysm = None
    for seed in range(2000, 2000 + args.eval_seeds):
        pr = meshgen.cube_elasticity(n=args.eval_n, seed=seed)
        if sysm is None:
            sysm = FESystem(pr.kind, pr.coords, pr.conn, pr.materials)
        else:
            sysm.set_materials(pr.materials)
        K = sysm.stiffness_bsr()
        fixed, g = sysm.dirichlet(pr.fixed_dofs)
        KD, b = sysm.prepare(K, fixed, g, sysm.ctx.tensor(pr.f))
        coarse = sysm.coarse_space(pr.fixed_dofs) if two else None
        pc = 'two_level' if two else 'block_jacobi'
        _, ib = sysm.solver(KD, pc, 'fp64', coarse=coarse).solve(b, tol=1e-8)
        M = model.preconditioner(sysm, KD, fixed)
        _, il = sysm.solver(KD, pc, 'fp64', M64=M, coarse=coarse).solve(b, tol=1e-8)
        evals.append({'seed': seed, 'without_network_it': ib['iterations'], 'learned_it': il['iterations'],
                      'learned_relres': il['relres']})
        log(f'eval seed {seed}: {pc} {ib["iterations"]} it, learned {il["iterations"]} it')
    model.meta.update({'train_n': args.n, 'snapshots': list(snaps), 'train_seconds': train_s,
                       'two_level': two, 'eval': evals})
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    model.save(args.out)
    log(f'saved {args.out}')
    print(json.dumps({'loss_first': hist[0], 'loss_last': hist[-1], 'eval': evals}), flush=True)


if __name__ == '__main__':
    sys.exit(main())
