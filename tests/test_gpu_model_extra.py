"""Model-level GPU checks beyond the space-time families: the spatial
Whittle-Matern model behind Model.setModel / sample / update / qinv and its
factor's log-determinant (whittle_matern2D.py:59-85, model.py:73-153)."""
import numpy as np
import pytest

from oracle import fixtures as F, grid as OG, models as Mo
from tests.gpu_util import step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,bc", [("iso", 1), ("ani", 3)])
def test_whittle_matern_model_sample_update(kind, bc):
    import spdepy_amd as sp
    from spdepy_amd import spdes
    M_, N_ = 9, 8
    x, y = F.axis(M_), F.axis(N_)
    par = np.array([0.2, np.log(0.5), 0.0]) if kind == "iso" else np.array([0.2, np.log(0.5), 0.3, -0.4, 0.0])
    g = sp.grid(x=x, y=y, Nbs=3)
    cls = spdes.WhittleMatern2D if kind == "iso" else spdes.WhittleMaternAnisotropic2D
    wm = cls(grid=g, par=par, bc=bc)
    og = OG.SpatialGrid(x, y, Nbs=3)
    Qo, _ = Mo.WhittleMatern(og, kind, par=par, bc=bc).makeQ(par, grad=False)
    Qd = Qo.toarray()
    step("WhittleMatern makeQ + factor (%s, bc=%d)" % (kind, bc))
    Q, fac, _ = wm.makeQ(par, grad=False)
    assert fac is not None
    assert np.allclose(Q.toarray(), Qd, rtol=1e-12, atol=1e-12 * np.abs(Qd).max())
    assert abs(fac.logdet() - np.linalg.slogdet(Qd)[1]) <= 1e-11 * abs(np.linalg.slogdet(Qd)[1])
    wm.setQ(par=par)
    mod = sp.Model.__new__(sp.Model)
    mod.grid, mod.mod, mod.Q, mod.mvar, mod.useCov = g, wm, None, None, False
    mod.mu = np.zeros(g.n)
    mod.setModel()
    step("WhittleMatern Model.sample")
    xs = mod.sample(n=2, simple=True, seed=3)
    back = mod.Q_fac.solve_Lt(mod.Q_fac.solve_L(Qd @ xs))
    assert np.allclose(np.asarray(back), xs, rtol=1e-8, atol=1e-10)
    step("WhittleMatern Model.update")
    idx = np.array([3, 10, 40, 41, 60])
    yv = np.array([0.5, -0.2, 1.0, 0.3, -0.7])
    mod.update(yv, idx, tau=4.0)
    S = g.getS(idx).toarray()
    Qc = Qd + 4.0 * S.T @ S
    assert np.allclose(mod.mu, np.linalg.solve(Qc, 4.0 * S.T @ yv), rtol=1e-9, atol=1e-12)
    step("WhittleMatern Model.qinv")
    mvar = mod.qinv()
    assert np.allclose(mvar, np.diag(np.linalg.inv(Qc)), rtol=1e-10)
