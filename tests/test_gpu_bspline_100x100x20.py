"""Parity at the headline configuration (B-spline advection and anisotropic
diffusion, 100 x 100 x 20, n = 200,000, Nbs = 8, 386 parameters, 5,000
observations, bc = 1): the drop-in model's exact log-likelihood and all 386
gradient entries against the oracle's values in tests/golden/bspline_100x100x20.npz,
which tests/golden/make_bspline_100x100x20.py computes by the Woodbury form (Q^-1 S^T
through Q = F^T D F, no Cholesky of Q_c) -- a route independent of the GPU's
factorisation and selected inverse."""
import os

import numpy as np
import pytest

from tests.gpu_util import step

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bspline_100x100x20.npz"))


def _model():
    import spdepy_amd as sp
    from spdepy_amd import spdes
    ax = (np.arange(100) + 0.5) / 100
    g = sp.grid(x=ax, y=ax, t=np.arange(20.0), Nbs=8)
    g0 = sp.grid(x=ax, y=ax, Nbs=8)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=1)
    mod = spdes.VarAdvectionVarDiffusion2D(grid=g, mod0=mod0, bc=1)
    mod.initFit(GOLD["y"], idx=GOLD["idx"], fitQ0=False)
    return mod


def test_loglike_and_full_gradient_100x100x20():
    mod = _model()
    par = GOLD["par"]
    step("100x100x20 exact logLike + 386-entry gradient vs Woodbury oracle")
    like, jac = mod.logLike(par)
    like_o, jac_o = float(GOLD["like"]), GOLD["jac"]
    scale = np.abs(jac_o).max()
    err = np.abs(jac - jac_o)
    print("  like %.15e oracle %.15e rel %.2e; grad max err / scale %.2e (entry %d)"
          % (like, like_o, abs(like - like_o) / abs(like_o), err.max() / scale, int(err.argmax())))
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    assert jac.shape == jac_o.shape == (386,)
    # every field coefficient and tau within 1e-9 of the gradient's scale;
    # d/d log sigma (entry 384) contains tr(Sigma_c,00 Q0) over the first
    # slice, a sum that cancels from ~|Q0| |Sigma_c| to ~Ns at condition
    # ~1e8, so the Cholesky / selected-inverse route carries ~7e-9 of the
    # scale there (the Woodbury oracle forms no such sum)
    sig = 6 * 64
    others = np.delete(np.arange(386), sig)
    print("  largest error off the sigma entry: %.2e of scale" % (err[others].max() / scale))
    assert np.allclose(jac[others], jac_o[others], rtol=1e-9, atol=1e-9 * scale)
    assert abs(jac[sig] - jac_o[sig]) <= 1e-8 * scale
    # the factorised route to log|Q| agrees too
    step("100x100x20 logLike, log|Q| factorised")
    like_f = mod.logLike(par, grad=False, logdet="factor")
    assert abs(like_f - like_o) <= 1e-9 * abs(like_o)
