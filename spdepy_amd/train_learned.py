"""Trains the learned preconditioner on the GPU and writes its weights.

    python -m spdepy_amd.train_learned [--draws 750] [--grid 100 100 20] [--samples 16] [--loss kaporin]

Each draw is a seeded parameter setting of the B-spline advection /
anisotropic-diffusion model (``learned.draw_parameters``) with observations
at 2.5% of the lattice (tau = 100), and one Adam step lowers the loss of the
corrected IC(0) factor (``learned.train``; --loss: Kaporin's condition number
by default, from K Rademacher probes, or the divergences against K samples of
N(0, Q_c^-1) that the exact FP64 Cholesky path supplies).  Progress goes to stdout, the
weights to spdepy_amd/weights/learned_ic.npz (or --out)."""
import argparse
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--draws", type=int, default=750)
    ap.add_argument("--grid", type=int, nargs=3, default=[100, 100, 20])
    ap.add_argument("--nbs", type=int, default=8)
    ap.add_argument("--samples", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--loss", default="kaporin", choices=["kaporin", "kaporin+jeffreys", "jeffreys", "kl"])
    a = ap.parse_args(argv)
    import spdepy_amd as sp
    from spdepy_amd import spdes
    from spdepy_amd.learned import LearnedPreconditioner, train, default_weights_path
    M, N, T = a.grid
    ax, ay = (np.arange(M) + 0.5) / M, (np.arange(N) + 0.5) / N
    g = sp.grid(x=ax, y=ay, t=np.arange(float(T)), Nbs=a.nbs)
    g0 = sp.grid(x=ax, y=ay, Nbs=a.nbs)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=1)
    mod = spdes.VarAdvectionVarDiffusion2D(grid=g, mod0=mod0, bc=1)
    net = LearnedPreconditioner(seed=a.seed, lr=a.lr)
    t0 = time.time()
    losses = train(net, mod, a.draws, K=a.samples, seed=a.seed + 1, lr=a.lr, loss=a.loss,
                   log=lambda s: print("%s (%.0fs)" % (s, time.time() - t0), flush=True))
    out = a.out or default_weights_path()
    os.makedirs(os.path.dirname(out), exist_ok=True)
    net.save(out)
    print("trained %d draws on %dx%dx%d in %.0fs: loss/row first 20 %.6f, last 20 %.6f -> %s"
          % (a.draws, M, N, T, time.time() - t0, losses[:20].mean(), losses[-20:].mean(), out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
