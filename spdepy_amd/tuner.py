"""Autotuner of the supernode leaf size per lattice shape.

The nested-dissection plan stops bisecting at boxes of at most `leaf`
vertices; the leaf size sets the width of the leaf supernodes and with it
the tile shapes of the whole factorisation (the 64 x 64 diagonal blocks and
the GEMM tile lists of every level).  ``tune_leaf`` times one factorisation
and selected inverse of a representative precision (the stationary
anisotropic advection-diffusion model at unit parameters, observations on
2.5% of the lattice) for each candidate and keeps the fastest; choices are
cached per (M, N, T, bc) in a JSON file."""
import json
import os

import numpy as np

CANDIDATES = (32, 64, 128, 256)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("spdepy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _representative(grid, plan):
    """Q values of the stationary anisotropic model on `plan` and a diagonal
    observation term."""
    from .engine import Coefficients
    torch = _torch()
    c = Coefficients(grid, variable=False, kappa=1.0, H=np.array([[0.6, 0.1], [0.1, 0.4]]), G=(0.2, -0.1),
                     sigma=1.0, kappa0=1.0, H0=np.eye(2) * 0.5)
    q = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
    plan.assemble(c.c, q)
    n = plan.n
    d = np.zeros(n)
    d[np.random.default_rng(0).choice(n, max(1, n // 40), replace=False)] = 100.0
    return q, torch.as_tensor(d).cuda()


def time_leaf(M, N, T, bc, leaf, reps=2):
    """Milliseconds of one factorisation + selected inverse of Q_c with the
    given leaf size (best of `reps` after one warm-up)."""
    import spdepy_amd as sp
    from .engine import get_plan
    torch = _torch()
    grid = sp.grid(x=(np.arange(M) + 0.5) / M, y=(np.arange(N) + 0.5) / N, t=np.arange(float(T)))
    plan = get_plan(M, N, T, bc, leaf_max=leaf)
    q, d = _representative(grid, plan)
    sig = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
    best = np.inf
    for k in range(reps + 1):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        plan.factor_async(1, q, d)
        plan.selinv(1, sig)
        b.record()
        plan.factor_finish(1)
        torch.cuda.synchronize()
        if k > 0:
            best = min(best, a.elapsed_time(b))
    return best


def tune_leaf(M, N, T, bc, candidates=CANDIDATES, reps=2, cache=None):
    """(best leaf, {leaf: ms}); with `cache` (a JSON path) a stored choice is
    returned without timing, and new choices are written back."""
    key = "%dx%dx%d_bc%d" % (M, N, T, bc)
    table = {}
    if cache is not None and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if key in table:
            return int(table[key]["leaf"]), {int(k): v for k, v in table[key]["ms"].items()}
    ms = {int(leaf): float(time_leaf(M, N, T, bc, leaf, reps)) for leaf in candidates}
    best = min(ms, key=ms.get)
    if cache is not None:
        table[key] = {"leaf": best, "ms": {str(k): v for k, v in ms.items()}}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return best, ms
