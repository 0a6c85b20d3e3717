"""Model wrapper of the drop-in API (berild/spdepy ``src/spdepy/model.py``):
holds Q, mu and tau; unconditional and conditional sampling; Bayesian
update with new observations.  Factorisations and solves run on the GPU."""
import numpy as np

from .engine import Engine
from .optim import Optimize
from .spdes import Factor, spde_init


class Model:
    def __init__(self, spde=None, grid=None, parameters=None, ani=True, ha=True, bc=3, mod0=None):
        self.grid = grid
        if spde is not None:
            self.model(spde=spde, grid=grid, parameters=parameters, ani=ani, ha=ha, bc=bc, mod0=mod0)
        else:
            self.mod = None
        self.Q = None
        self.mvar = None
        self.mu = np.zeros(self.grid.shape)
        self.useCov = False
        self.sigmas = np.log(np.array([0.01, 140]))

    def setQ(self, par=None):
        self.mod.setQ(par=par)

    def model(self, spde=None, grid=None, parameters=None, ani=True, ha=True, bc=3, mod0=None):
        """model.py:25-46 -- a default Whittle-Matern initial-state model is
        created for space-time grids when none is given."""
        assert spde is not None and grid is not None
        if grid.type == "gridST" and mod0 is None:
            mod0 = spde_init(model="whittle-matern", grid=grid, parameters=parameters, ani=ani, ha=ha, bc=bc)
        self.mod = spde_init(model=spde, grid=grid, parameters=parameters, ani=ani, ha=ha, bc=bc, mod0=mod0)
        self.spde_type = self.mod.type
        self.optim = Optimize(self.mod.logLike)

    def fit(self, data, **kwargs):
        """model.py:49-71: initFit, then the optimiser over logLike (one GPU
        log-likelihood + gradient evaluation per step), then setQ at the
        Polyak-averaged parameters.  kwargs as Optimize.fit (x0, stepType,
        lr, max_steps, pol, fix, verbose, end) plus initFit's (idx, fitQ0)."""
        assert self.mod is not None
        x0 = self.mod.initFit(data, **kwargs)
        if kwargs.get("x0") is None:
            kwargs["x0"] = x0
        if kwargs.get("verbose") is not None:
            kwargs["print"] = self.mod.print
        res = self.optim.fit(**kwargs)
        self.setQ(par=res["x"])
        return res

    def _engine(self):
        return self.mod._engine()

    def setModel(self, mu=None, sigmas=None, useCov=None, scale=True):
        """model.py:129-153 without covariates: Q from the model, mu, tau."""
        if useCov:
            raise NotImplementedError("covariate-augmented models are outside the hot path")
        self.useCov = False
        self.mu = np.zeros(self.mod.Q.shape[0]) if mu is None else self.grid.getS().T @ mu
        self.Q = self.mod.Q.copy().tocsc()
        self._qval = self.mod._last_qval
        self._diag = None
        self.tau = np.exp(self.mod.tau)

    def _factor(self):
        import torch
        eng = self._engine()
        diag = None if self._diag is None else torch.as_tensor(self._diag).cuda()
        return Factor(eng, 0, self._qval, diag)

    def sample(self, n=1, simple=False, seed=None):
        """x = P^T L^-T z + mu with z ~ N(0, I) from default_rng(seed) (model.py:73-87);
        P is this library's nested-dissection ordering."""
        if seed is None:
            seed = np.random.randint(100)
        z = np.random.default_rng(seed).normal(size=self.Q.shape[0] * n).reshape(self.Q.shape[0], n)
        self.Q_fac = self._factor()
        x = self.Q_fac.apply_Pt(self.Q_fac.solve_Lt(z, use_LDLt_decomposition=False))
        data = self.grid.getS() @ (x + self.mu[:, np.newaxis])
        if not simple:
            data += self.grid.getS()[:, :np.prod(self.grid.shape)] @ z[:np.prod(self.grid.shape)] / np.sqrt(self.tau)
        return data

    def qinv(self, simple=False):
        """Marginal variances (model.py:89-118).  simple=True: variance of
        1000 unconditional samples, as the reference.  Otherwise Q^-1 on the
        pattern of Q from the Takahashi selected inverse of the GPU factor
        (in place of R-INLA's qinv, kept in FP64 instead of the reference's
        float32 cast) in ``self.pQinv`` (CSC); returns its diagonal."""
        if simple:
            z = self.sample(n=1000, simple=True)
            return z.var(axis=1)
        from scipy import sparse
        if self.Q is None:
            self.setModel()
        if self.useCov:
            raise NotImplementedError("covariate-augmented models are outside the hot path")
        eng = self._engine()
        fac = self._factor()
        sig = fac.selinv().cpu().numpy()
        indptr, indices = eng.plan.structure()
        n = eng.n
        full = sparse.csr_matrix((sig, indices, indptr), shape=(n, n))
        pattern = self.Q.tocsc()
        mask = sparse.csc_matrix((np.ones(pattern.nnz), pattern.indices, pattern.indptr), shape=pattern.shape)
        self.pQinv = full.tocsc().multiply(mask).tocsc()
        self.mvar = self.pQinv.diagonal()
        return self.mvar

    def update(self, y, idx, tau=None):
        """Q <- Q + tau S^T S, mu <- mu + tau Q^-1 S^T (y - S mu) (model.py:120-127)."""
        if tau is None:
            tau = self.tau
        S = self.grid.getS(idx)
        add = np.asarray((S.T @ S).diagonal(), dtype=np.float64) * tau
        self._diag = add if self._diag is None else self._diag + add
        self.Q = (self.Q + S.transpose() @ S * tau).tocsc()
        self.Q_fac = self._factor()
        self.mu = self.mu + self.Q_fac.solve_A(S.T @ (y - S @ self.mu)) * tau

    def getPars(self, onlySelf=True):
        return self.mod.getPars(onlySelf=onlySelf)
