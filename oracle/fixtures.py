"""Seeded workloads of the hot path (test infrastructure).

stationary_30x30x10: stationary isotropic advection-diffusion, 30x30 cells x 10 steps
         (n = 9,000), bc=1, par = [log 1, log 0.5, 0.3, -0.2, log 1, log 100],
         mod0 = isotropic Whittle-Matern [0, log 0.5, 0]; one seeded sample,
         200 observations at uniform random (cell, time) indices + N(0, 0.1^2).
bspline_100x100x20: B-spline (Nbs=8) advection and anisotropic diffusion on 100x100 x 20
         steps (n = 200,000), 386 parameters, 5,000 observations.
small(): the 100x100x20 B-spline model class on reduced grids, for dense checks.
field_model(): the families with some coefficients as B-spline fields and
         the rest constant, on small grids (dense sampling).
"""
import numpy as np
from . import grid as G
from . import models as Mo


def axis(n):
    return (np.arange(n) + 0.5) / n


def mod0_iso(M, N, Nbs=3, bc=1, ne=None):
    g0 = G.SpatialGrid(axis(M), axis(N), extend=ne, Nbs=Nbs)
    return Mo.WhittleMatern(g0, "iso", par=np.array([0.0, np.log(0.5), 0.0]), bc=bc)


def observations(n_interior, m, x_full, grid, rng_idx, rng_noise, sd=0.1):
    idx = np.sort(rng_idx.choice(n_interior, m, replace=False))
    cols = grid.n2e(idx)
    y = x_full[cols, 0] + sd * rng_noise.normal(size=m)
    return idx, y


def stationary(M=30, N=30, T=10, bc=1, kind="iso", m=200, par=None, seeds=(0, 1, 2)):
    g = G.STGrid(axis(M), axis(N), np.arange(float(T)), Nbs=3)
    mod0 = mod0_iso(M, N, bc=bc)
    model = Mo.StationaryAdvDiff(g, mod0, kind=kind, bc=bc)
    if par is None:
        if kind == "iso":
            par = np.array([0.0, np.log(0.5), 0.3, -0.2, 0.0, np.log(100.0)])
        else:
            par = np.array([0.0, np.log(0.5), 0.4, -0.3, 0.3, -0.2, 0.0, np.log(100.0)])
    A, q, sigma = Mo.stationary_spatial_operator(model, par)
    A0, _, _, _ = mod0.A0(mod0.par)
    z = np.random.default_rng(seeds[0]).normal(size=g.n).reshape(g.n, 1)
    x = Mo.sample_fform(A, q, A0, g.V, T, sigma, g.dt, z)
    idx, y = observations(M * N * T, m, x, g, np.random.default_rng(seeds[1]),
                          np.random.default_rng(seeds[2]))
    return dict(grid=g, mod0=mod0, model=model, par=par, idx=idx, y=y, x=x)


def var_par(Np, rng, sd=0.3):
    return np.hstack([np.zeros(Np), np.full(Np, np.log(0.5)),
                      sd * rng.normal(size=4 * Np), [0.0, np.log(100.0)]])


def var_model(M=100, N=100, T=20, Nbs=8, bc=1, m=5000, seeds=(3, 4, 6, 7), ne=None):
    g = G.STGrid(axis(M), axis(N), np.arange(float(T)), Nbs=Nbs, extend=ne)
    mod0 = mod0_iso(M, N, Nbs=Nbs, bc=bc, ne=ne)
    model = Mo.VarAdvVarDiff(g, mod0, bc=bc)
    par = var_par(g.Nbs2, np.random.default_rng(seeds[0]))
    A, q, sigma = model.spatial_operator(par)
    A0, _, _, _ = mod0.A0(mod0.par)
    z = np.random.default_rng(seeds[2]).normal(size=g.n).reshape(g.n, 1)
    x = Mo.sample_fform(A, q, A0, g.V, T, sigma, g.dt, z)
    idx, y = observations(M * N * T, m, x, g, np.random.default_rng(seeds[1]),
                          np.random.default_rng(seeds[3]))
    return dict(grid=g, mod0=mod0, model=model, par=par, idx=idx, y=y, x=x)


def stationary_30x30x10():
    return stationary()


def bspline_100x100x20():
    return var_model()


def small(M=8, N=7, T=4, Nbs=3, bc=1, m=40, ne=None):
    return var_model(M, N, T, Nbs=Nbs, bc=bc, m=m, ne=ne)


def currents(g):
    """Smooth synthetic cell-centred currents (we, wn) on the interior cells
    (a rotating cell pattern like the reference's advection example)."""
    X, Y = np.meshgrid(g.x, g.y)
    we = np.sin(3 * X) * np.cos(3 * Y) + 0.2
    wn = -np.cos(3 * X) * np.sin(3 * Y) - 0.1
    return we.ravel(), wn.ravel()


def field_par(model, rng, sd=0.3):
    """kappa = 1, gamma = 0.5 plus field noise; random v and w."""
    nd, nh, nw = model.sizes()
    blocks = [np.zeros(nd) + (sd * rng.normal(size=nd) if nd > 1 else 0.0),
              np.full(nd, np.log(0.5)) + (sd * rng.normal(size=nd) if nd > 1 else 0.0)]
    blocks += [sd * rng.normal(size=nd) + 0.2 for _ in range(nh - 1)]
    blocks.append(sd * rng.normal(size=nw))
    return np.hstack(blocks + [[0.0, np.log(100.0)]])


def field_model(M=8, N=7, T=4, kind="iso", dvar=True, gvar=False, bc=3, m=40, seeds=(8, 9, 10, 11), cov=None):
    """cov=lambda: the cov-advection family with advection lambda * ww,
    ww = assimilate_adv(currents)."""
    g = G.STGrid(axis(M), axis(N), np.arange(float(T)), Nbs=3)
    mod0 = mod0_iso(M, N, bc=bc)
    ww = None if cov is None else g.assimilate_adv(*currents(g))
    model = Mo.FieldAdvDiff(g, mod0, kind=kind, dvar=dvar, gvar=gvar, bc=bc, ww=ww)
    par = field_par(model, np.random.default_rng(seeds[0]))
    if cov is not None:
        par[-3] = cov
    Q, _ = model.makeQ(par, grad=False)
    L = np.linalg.cholesky(Q.toarray())
    z = np.random.default_rng(seeds[2]).normal(size=(g.n, 1))
    x = np.linalg.solve(L.T, z)
    idx, y = observations(M * N * T, m, x, g, np.random.default_rng(seeds[1]),
                          np.random.default_rng(seeds[3]))
    return dict(grid=g, mod0=mod0, model=model, par=par, idx=idx, y=y, x=x, ww=ww)
