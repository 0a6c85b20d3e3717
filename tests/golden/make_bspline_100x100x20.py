"""Writes tests/golden/bspline_100x100x20.npz: the oracle's exact log-likelihood and full
386-entry gradient at 100 x 100 x 20 (n = 200,000; B-spline advection and
anisotropic diffusion, Nbs = 8, bc = 1, 5,000 observations:
``oracle.fixtures.bspline_100x100x20()``), with the observation indices and values.

The exact trace at this size comes from the Woodbury form
(``oracle.models.loglike_woodbury``): Q^{-1} S^T through Q = F^T D F with dense
inverses of the two 100 x 100-cell spatial operators, the 5,000 x 5,000
observation-space matrix K, and <dQ_i, W K^{-1} W^T> for each of the 385
matrices dQ_i built by the oracle's block formula -- a route independent of
the GPU's Cholesky and selected inverse of Q_c.  About 15 minutes and 20 GB
on 8 cores.  Run in the CPU container:

    python tests/golden/make_bspline_100x100x20.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import fixtures as F, models as Mo  # noqa: E402


def main():
    t0 = time.time()
    c = F.bspline_100x100x20()
    g, model, par = c["grid"], c["model"], c["par"]
    M, N, T = g.shape
    Q, _ = model.makeQ(par, grad=False)
    A, q, sigma = model.spatial_operator(par)
    A0 = c["mod0"].A0(c["mod0"].par)[0]
    cols = g.n2e(c["idx"])
    print("fixture built %.0fs" % (time.time() - t0), flush=True)
    like, jac, mu = Mo.loglike_woodbury(Q, model.dQ_iter(par), A, q, A0, g.V, T, sigma, g.dt, M, N, cols,
                                        c["y"], par[-1])
    ld_q = Mo.logdet_closed_form(A, q, A0, g.V, T, sigma, g.dt)
    np.savez_compressed(os.path.join(HERE, "bspline_100x100x20.npz"), par=par, idx=c["idx"], y=c["y"],
                        like=np.array(like), jac=jac, logdet_q=np.array(ld_q),
                        mu_sum=np.array(mu.sum()), mu_sq=np.array((mu ** 2).sum()))
    print("like %.15e |jac| %.6e  (%.0fs)" % (like, np.linalg.norm(jac), time.time() - t0))


if __name__ == "__main__":
    main()
