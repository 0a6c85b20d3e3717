"""Writes tests/golden/examples.npz from the reference's example artefacts
(berild/spdepy ``examples/*/fits/*.npy`` parameter vectors, ``examples/*/data``):
the fit vectors, the initial-state parameters, one realisation of two example
data sets, and the oracle's values at those fits.  Run in the CPU container:

    python tests/golden/make_examples.py

Grids as the examples define them (their grid.npz is not in the repository):
50 x 50 cells on [0, 15]^2 extended by 5 cells, 12 time steps; t = 0..11 in
advection_example/master.ipynb, t = linspace(0, 2, 12) in
advection_diffusion_example/master.py.  Observed cells: all but the block
19 < i < 38, 12 < j < 31 (the examples' idxD)."""
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import grid as G, models as Mo  # noqa: E402

EX = "/root/reference/examples"
NH1, SEED = 20, 7


def idx_observed(M=50, N=50, T=12):
    hole = [t * M * N + j * M + i for t in range(T) for i in range(M) for j in range(N)
            if 19 < i < 38 and 12 < j < 31]
    return np.delete(np.arange(M * N * T), hole)


def probes(n):
    np.random.seed(SEED)
    return (2 * np.random.randint(1, 3, n * NH1) - 3).reshape(n, NH1)


def main():
    out = {}
    for f in sorted(glob.glob(EX + "/*/fits/*.npy") + glob.glob(EX + "/fits/*.npy")):
        rel = os.path.relpath(f, EX).replace("/fits/", "/").replace(".npy", "")
        if rel.startswith("fits/"):
            rel = "sinmod/" + rel[5:]
        out["fit/" + rel] = np.load(f, allow_pickle=False)
    for ex in ("advection_example", "diffusion_example", "advection_diffusion_example"):
        out["mod0/" + ex] = np.load(EX + "/%s/data/mod0pars.npy" % ex, allow_pickle=False)
    idx = idx_observed()
    for ex in ("advection_example", "advection_diffusion_example"):
        assert np.array_equal(idx, np.load(EX + "/%s/data/idxD.npy" % ex, allow_pickle=False))
        out["data/" + ex] = np.load(EX + "/%s/data/data.npy" % ex, allow_pickle=False)[:, 0].copy()
    x = np.linspace(0, 15, 50)

    # A: advection_example, var-advection-diffusion (isotropic, bc 3), mod0 anisotropic WM
    g = G.STGrid(x, x, np.arange(12.0), extend=5, Nbs=3)
    mod0 = Mo.WhittleMatern(G.SpatialGrid(x, x, extend=5, Nbs=3), "ani", par=out["mod0/advection_example"], bc=3)
    m = Mo.FieldAdvDiff(g, mod0, "iso", dvar=False, gvar=True, bc=3)
    par = out["fit/advection_example/var_advection_diffusion_bc3"]
    Q, dQ = m.makeQ(par, grad=True)
    like, jac, _ = Mo.loglike(Q, dQ, g.getS(idx), out["data/advection_example"], par[-1], trace="hutchinson",
                              probes=probes(g.n))
    out["expect/A_like"], out["expect/A_jac_hutchinson"] = np.array(like), jac

    # B: advection_diffusion_example, var-advection-var-diffusion (anisotropic, bc 1), mod0 isotropic WM
    g = G.STGrid(x, x, np.linspace(0, 2, 12), extend=5, Nbs=3)
    mod0 = Mo.WhittleMatern(G.SpatialGrid(x, x, extend=5, Nbs=3), "iso", par=out["mod0/advection_diffusion_example"],
                            bc=1)
    m = Mo.VarAdvVarDiff(g, mod0, bc=1)
    par = out["fit/advection_diffusion_example/var_advection_var_diffusion_ani_bc1"]
    Q, _ = m.makeQ(par, grad=False)
    like, _, _ = Mo.loglike(Q, [], g.getS(idx), out["data/advection_diffusion_example"], par[-1], trace="hutchinson",
                            probes=probes(g.n))
    out["expect/B_like"] = np.array(like)

    # C: advection_diffusion_example, separable (anisotropic, bc 1)
    m = Mo.Separable(g, "ani", bc=1)
    par = out["fit/advection_diffusion_example/seperable_spatial_temporal_ani_bc1"]
    Q, _ = m.makeQ(par, grad=False)
    like, _, _ = Mo.loglike(Q, [], g.getS(idx), out["data/advection_diffusion_example"], par[-1], trace="hutchinson",
                            probes=probes(g.n))
    out["expect/C_like"] = np.array(like)
    np.savez_compressed(os.path.join(HERE, "examples.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
