"""Batched evaluation of models of uneven size (spde_loglik_batch through the
C ABI): every model's log-likelihood and gradient equal its own logLike, two
models on one lattice run in separate waves, conditional samples come from
each model's factor of Q_c, and the leaf-size autotuner picks and caches a
candidate."""
import os

import numpy as np
import pytest
from scipy import sparse

from oracle import krylov as K
from tests.gpu_util import step

pytestmark = pytest.mark.gpu


def stationary_model(M, N, T, m_obs, seed):
    import spdepy_amd as sp
    from spdepy_amd import spdes
    rng = np.random.default_rng(seed)
    x, y = (np.arange(M) + 0.5) / M, (np.arange(N) + 0.5) / N
    g = sp.grid(x=x, y=y, t=np.arange(float(T)))
    g0 = sp.grid(x=x, y=y)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=1)
    mod = spdes.AdvectionDiffusion2D(grid=g, mod0=mod0, bc=1)
    idx = np.sort(rng.choice(M * N * T, m_obs, replace=False))
    mod.initFit(rng.normal(size=m_obs), idx=idx, fitQ0=False)
    par = np.array([rng.uniform(np.log(0.5), np.log(2.0)), rng.uniform(np.log(0.2), np.log(1.0)),
                    rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5),
                    0.0, np.log(50.0)])
    return mod, par


def test_batch_equals_single_evaluations():
    from spdepy_amd.batch import BatchEvaluator
    shapes = [(12, 10, 4, 60), (16, 14, 5, 120), (12, 10, 4, 80), (20, 18, 6, 150), (10, 12, 3, 40), (14, 14, 4, 90)]
    models, pars = zip(*[stationary_model(M, N, T, m, 100 + i) for i, (M, N, T, m) in enumerate(shapes)])
    step("batch of %d models on 4 streams" % len(models))
    ev = BatchEvaluator(models, nstreams=4)
    out = ev.loglik(pars)
    for i, (m, par) in enumerate(zip(models, pars)):
        step("single evaluation %d" % i)
        like, jac = m.logLike(par)
        lb, jb = out[i]
        assert abs(lb - like) <= 1e-12 * abs(like), (i, lb, like)
        assert np.allclose(jb, jac, rtol=1e-10, atol=1e-11 * np.abs(jac).max()), i
    # a second call reuses every plan's buffers
    step("batch again")
    out2 = ev.loglik(pars)
    assert all(abs(a[0] - b[0]) <= 1e-13 * abs(b[0]) for a, b in zip(out2, out))


def test_batch_conditional_samples():
    import torch
    from spdepy_amd import _native as nat
    from spdepy_amd.batch import BatchEvaluator
    mod, par = stationary_model(10, 9, 4, 50, 7)
    ev = BatchEvaluator([mod], nstreams=2)
    ev.loglik([par], grad=False)
    step("256 conditional samples")
    X = ev.conditional_samples(0, 256, seed=3)
    assert X.shape == (mod.grid.n, 256)
    Q = mod.makeQ(par, grad=False)[0]
    cols = mod._obs_cols()
    tau = np.exp(par[-1])
    counts = np.bincount(cols, minlength=mod.grid.n).astype(float)
    Qc = (Q + sparse.diags(tau * counts)).toarray()
    b = np.zeros(mod.grid.n)
    np.add.at(b, cols, tau * np.asarray(mod.data))
    mu = np.linalg.solve(Qc, b)
    plan = mod._engine().plan
    Z = torch.as_tensor(K.gaussian(mod.grid.n, 256, 3)).cuda()
    dev = (X - plan.solve(1, Z, mode=nat.SOLVE_LT)).cpu().numpy()
    assert np.allclose(dev, mu[:, None], rtol=1e-9, atol=1e-10 * np.abs(mu).max())
    # x - mu = P^T L^-T z has covariance Q_c^-1: (x - mu)^T Q_c (x - mu) = z^T z per sample
    D = (X.cpu().numpy() - mu[:, None])
    zz = (Z.cpu().numpy() ** 2).sum(0)
    assert np.allclose(np.einsum("is,ij,js->s", D, Qc, D), zz, rtol=1e-8)


def test_leaf_tuner_picks_and_caches(tmp_path):
    from spdepy_amd.tuner import tune_leaf
    cache = str(tmp_path / "leaf.json")
    step("leaf tuner 16x14x5")
    best, ms = tune_leaf(16, 14, 5, 1, candidates=(32, 64), reps=1, cache=cache)
    assert best in (32, 64) and set(ms) == {32, 64} and all(v > 0 for v in ms.values())
    assert os.path.exists(cache)
    again, ms2 = tune_leaf(16, 14, 5, 1, candidates=(32, 64), reps=1, cache=cache)
    assert again == best and ms2 == ms
