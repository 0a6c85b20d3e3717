"""Writes tests/golden/stationary_30x30x10.npz: the oracle's log-likelihood and exact
gradient at the CPU-runnable configuration (stationary isotropic
advection-diffusion, 30 x 30 cells x 10 steps, n = 9,000, bc = 1,
par = [log 1, log 0.5, 0.3, -0.2, log 1, log 100], mod0 = isotropic
Whittle-Matern [0, log 0.5, 0], one seeded sample, 200 observations with
noise sd 0.1: ``oracle.fixtures.stationary()``), with the observation
indices and values, Q's pattern (indptr, indices of the sorted CSR) and
checksums of its values.  The exact trace uses dense inverses
of Q and Q_c (n = 9,000: about a minute on 8 cores).  Run in the CPU
container:

    python tests/golden/make_stationary_30x30x10.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import fixtures as F, models as Mo  # noqa: E402


def main():
    c = F.stationary()
    g = c["grid"]
    Q, dQ = c["model"].makeQ(c["par"], grad=True)
    Q = Q.tocsr()
    Q.sort_indices()
    S = g.getS(c["idx"])
    like, jac, _ = Mo.loglike(Q.tocsc(), dQ, S, c["y"], c["par"][-1], trace="exact")
    np.savez_compressed(os.path.join(HERE, "stationary_30x30x10.npz"), par=c["par"], idx=c["idx"], y=c["y"],
                        like=np.array(like), jac=jac, q_nnz=np.array(Q.nnz),
                        q_abs_sum=np.array(np.abs(Q.data).sum()), q_diag=Q.diagonal(),
                        q_indptr=Q.indptr.astype(np.int32), q_indices=Q.indices.astype(np.int32))
    print("like", like, "jac", jac)


if __name__ == "__main__":
    main()
