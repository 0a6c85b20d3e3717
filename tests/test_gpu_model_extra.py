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
    back = mod.Q_fac.pt_solve_Lt(mod.Q_fac.solve_L_p(Qd @ xs))
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


@pytest.mark.parametrize("two", [True, False])
def test_model_with_covariates(two):
    """Model.setModel(useCov=True) (model.py:129-153): intercept + covariate
    (two sigmas) or intercept only (one sigma) border columns beside the
    lattice; sample, update and qinv (model.py:73-127) against dense algebra
    on the augmented precision, the sample through a dense Cholesky factor
    in this library's ordering extended by the border."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    M_, N_ = 9, 8
    x, y = F.axis(M_), F.axis(N_)
    par = np.array([0.2, np.log(0.5), 0.3, -0.4, 0.0])
    g = sp.grid(x=x, y=y, Nbs=3)
    wm = spdes.WhittleMaternAnisotropic2D(grid=g, par=par, bc=3)
    wm.setQ(par=par)
    Ql = wm.Q.toarray()
    n = g.n
    mod = sp.Model.__new__(sp.Model)
    mod.grid, mod.mod, mod.Q, mod.mvar, mod.useCov = g, wm, None, None, False
    mod.mu = np.zeros(n)
    cov = np.linspace(0.5, 2.0, n) + 0.1 * np.sin(np.arange(n))
    if two:
        mod.setModel(mu=cov, sigmas=np.log([0.5, 3.0]), useCov=True, scale=True)
        C = np.diag([0.5, 3.0])
        X = np.stack([np.ones(n), cov / cov.max()], 1)
        mu0 = np.r_[np.zeros(n), 0.0, cov.max()]
    else:
        mod.setModel(mu=cov, sigmas=np.log(2.0), useCov=True)
        C = np.array([[2.0]])
        X = np.ones((n, 1))
        mu0 = np.r_[cov, 0.0]
    k = C.shape[0]
    Qa = np.block([[Ql, np.zeros((n, k))], [np.zeros((k, n)), C]])
    assert np.allclose(mod.Q.toarray(), Qa, rtol=1e-13, atol=0)
    assert np.allclose(mod.mu, mu0, rtol=1e-13, atol=0)
    S = np.hstack([np.eye(n), X])
    assert np.allclose(g.getS().toarray(), S)

    def check_sample(Qdense):
        z = np.random.default_rng(5).normal(size=(n + k) * 2).reshape(n + k, 2)
        step("covariate sample (k=%d)" % k)
        xs = mod._sample_aug(z)
        perm = np.r_[mod._engine().plan.export()["perm"], np.arange(n, n + k)]
        Lt = np.linalg.cholesky(Qdense[np.ix_(perm, perm)]).T
        want = np.empty_like(z)
        want[perm] = np.linalg.solve(Lt, z)
        assert np.allclose(xs, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())
        data = mod.sample(n=2, simple=True, seed=5)
        assert np.allclose(data, S @ (want + mod.mu[:, None]), rtol=1e-9, atol=1e-11 * np.abs(data).max())

    check_sample(Qa)
    step("covariate qinv (k=%d)" % k)
    mvar = mod.qinv()
    want = np.r_[np.diag(np.linalg.inv(S @ Qa @ S.T)), np.zeros(k)]
    assert mvar.shape == (n + k,)
    assert np.allclose(mvar, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    idx = np.array([3, 10, 40, 41, 60, 10])
    yv = np.array([0.5, -0.2, 1.0, 0.3, -0.7, 0.1])
    step("covariate update (k=%d)" % k)
    mod.update(yv, idx, tau=4.0)
    Sd = g.getS(idx).toarray()
    Qc = Qa + 4.0 * Sd.T @ Sd
    mu1 = mu0 + 4.0 * np.linalg.solve(Qc, Sd.T @ (yv - Sd @ mu0))
    assert np.allclose(mod.Q.toarray(), Qc, rtol=1e-13, atol=1e-13 * np.abs(Qc).max())
    assert np.allclose(mod.mu, mu1, rtol=1e-9, atol=1e-11 * np.abs(mu1).max())
    check_sample(Qc)
    mvar = mod.qinv()
    want = np.r_[np.diag(np.linalg.inv(S @ Qc @ S.T)), np.zeros(k)]
    assert np.allclose(mvar, want, rtol=1e-8, atol=1e-12 * np.abs(want).max())


def test_kriging_driver_flow_on_an_extended_grid():
    """The reference's kriging drivers (examples/server/get_results.py:6-14):
    setModel(mu=cov, sigmas=2 values, useCov=True) on a space-time model over
    an extended grid, the observed interior cells by getIdx(extend=False),
    update(y, idx), and the prediction getS() @ mu -- against dense algebra
    on the covariate-augmented precision."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    M_, N_, T_ = 8, 7, 3
    x, y = F.axis(M_), F.axis(N_)
    g = sp.grid(x=x, y=y, t=np.arange(float(T_)), extend=2)
    g0 = sp.grid(x=x, y=y, extend=2)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=3)
    st = spdes.AdvectionDiffusion2D(grid=g, mod0=mod0, bc=3)
    par = np.array([0.1, np.log(0.6), 0.2, -0.1, 0.3, -0.2, 0.0, np.log(20.0)])
    st.setQ(par=par)
    mod = sp.Model.__new__(sp.Model)
    mod.grid, mod.mod, mod.Q, mod.mvar, mod.useCov = g, st, None, None, False
    n_int = M_ * N_ * T_
    mod.mu = np.zeros(n_int)
    cov = 1.0 + 0.5 * np.cos(np.arange(n_int) * 0.37)
    step("kriging flow: setModel(useCov) on an extended grid")
    mod.setModel(mu=cov, sigmas=np.log([0.05, 2.0]), useCov=True)
    Qa = mod.Q.toarray()
    n = g.n
    assert Qa.shape == (n + 2, n + 2)
    mu0 = mod.mu.copy()
    uidx = np.array([g.getIdx([i, j, 1], extend=False) for i in range(M_) for j in range(N_)])
    yobs = np.sin(np.arange(uidx.size) * 0.3)
    step("kriging flow: update + prediction")
    mod.update(y=yobs, idx=uidx)
    tau = np.exp(par[-1])
    Sd = g.getS(uidx).toarray()
    Qc = Qa + tau * Sd.T @ Sd
    mu1 = mu0 + tau * np.linalg.solve(Qc, Sd.T @ (yobs - Sd @ mu0))
    assert np.allclose(mod.mu, mu1, rtol=1e-9, atol=1e-11 * np.abs(mu1).max())
    pred = g.getS() @ mod.mu
    want = g.getS().toarray() @ mu1
    assert np.allclose(pred, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())
