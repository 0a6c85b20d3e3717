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
        """model.py:129-153.  Without covariates: Q from the model, mu, tau.
        With useCov: Q = block_diag(Q_model, diag(exp(sigmas))) over the
        lattice and k = 2 (intercept + covariate mu, scaled by its maximum
        when scale) or k = 1 (intercept) border variables, whose columns
        grid.getS appends; the border stays dense (B: n x k, C: k x k) beside
        the GPU factor of the lattice block."""
        from scipy import sparse
        self.useCov = useCov if useCov is not None else self.useCov
        self._qval = self.mod._last_qval
        self._diag = None
        self.tau = np.exp(self.mod.tau)
        if not self.useCov:
            self.mu = np.zeros(self.mod.Q.shape[0]) if mu is None else self.grid.getS().T @ mu
            self.Q = self.mod.Q.copy().tocsc()
            self._border = None
            return
        self.sigmas = sigmas if sigmas is not None else self.sigmas
        n = self.mod.Q.shape[0]
        if hasattr(self.sigmas, "__len__"):
            self.grid.addCov(mu, scale=scale)
            C = np.diag(np.exp(np.asarray(self.sigmas, dtype=np.float64)))
            self.Q = sparse.block_diag([self.mod.Q.copy(), C]).tocsc()
            if scale:
                self.mu = np.zeros(self.Q.shape[0])
                self.mu[-2:] = [0, np.max(mu)]
            else:
                self.mu = np.zeros(self.Q.shape[0]) if mu is None else self.grid.getS().T @ mu
        else:
            self.grid.addInt()
            C = np.diag([np.exp(float(self.sigmas))])
            self.Q = sparse.block_diag([self.mod.Q.copy(), C]).tocsc()
            S = self.grid.getS()
            self.mu = np.zeros(S.shape[1])
            self.mu[:-1] = S[:, :-1].T @ mu
        self._border = (np.zeros((n, C.shape[0])), C.copy())

    def _factor(self):
        import torch
        eng = self._engine()
        diag = None if self._diag is None else torch.as_tensor(self._diag).cuda()
        return Factor(eng, 0, self._qval, diag)

    def _n_lattice(self):
        return self._engine().n

    def _solve_aug(self, b):
        """Q^-1 b for the covariate-augmented precision [[Q_l, B], [B^T, C]]
        by block elimination: the Schur complement C - B^T Q_l^-1 B (k x k)."""
        fac = self._factor()
        B, C = self._border
        n = self._n_lattice()
        b = np.asarray(b, dtype=np.float64)
        b1, b2 = b[:n], b[n:]
        U = fac.solve_A(B)
        w = fac.solve_A(b1)
        Sc = C - B.T @ U
        x2 = np.linalg.solve(Sc, b2 - B.T @ w)
        return np.concatenate([w - U @ x2, x2])

    def _sample_aug(self, z):
        """x = P~^T L~^-T z for the augmented precision: with P Q_l P^T = L L^T,
        Bt = L^-1 P B and L2 L2^T = C - Bt^T Bt, L~ = [[L, 0], [Bt^T, L2]]."""
        fac = self._factor()
        self.Q_fac = fac
        B, C = self._border
        n = self._n_lattice()
        Bt = np.asarray(fac.solve_L_p(B)).reshape(n, -1)
        L2 = np.linalg.cholesky(C - Bt.T @ Bt)
        z1, z2 = z[:n], z[n:]
        v2 = np.linalg.solve(L2.T, z2)
        v1 = fac.pt_solve_Lt(z1 - Bt @ v2)
        return np.vstack([np.asarray(v1).reshape(n, -1), v2])

    def sample(self, n=1, simple=False, seed=None):
        """x = P^T L^-T z + mu with z ~ N(0, I) from default_rng(seed) (model.py:73-87);
        P is this library's nested-dissection ordering."""
        if seed is None:
            seed = np.random.randint(100)
        z = np.random.default_rng(seed).normal(size=self.Q.shape[0] * n).reshape(self.Q.shape[0], n)
        if self.useCov:
            x = self._sample_aug(z)
        else:
            self.Q_fac = self._factor()
            x = self.Q_fac.pt_solve_Lt(z)
        data = self.grid.getS() @ (x + self.mu[:, np.newaxis])
        if not simple:
            data += self.grid.getS()[:, :np.prod(self.grid.shape)] @ z[:np.prod(self.grid.shape)] / np.sqrt(self.tau)
        return data

    def qinv(self, simple=False):
        """Marginal variances (model.py:89-118).  simple=True: variance of
        1000 unconditional samples, as the reference.  Otherwise Q^-1 on the
        pattern of Q from the Takahashi selected inverse of the GPU factor
        (in place of R-INLA's qinv, kept in FP64 instead of the reference's
        float32 cast) in ``self.pQinv`` (CSC); returns its diagonal.  With
        covariates the reference inverts S Q S^T (model.py:108-117): for a
        lattice without border extension (S's lattice part the identity)
        that is Q_l + W Z W^T with W = [B, X] and Z = [[0, I], [I, C]], whose
        inverse's diagonal follows from the selected inverse of Q_l and a
        2k-column Woodbury correction (the diagonal only: pQinv is None;
        the k border entries are zero, as in the reference's n + k diagonal)."""
        if simple:
            z = self.sample(n=1000, simple=True)
            return z.var(axis=1)
        from scipy import sparse
        if self.Q is None:
            self.setModel()
        eng = self._engine()
        fac = self._factor()
        sig = fac.selinv().cpu().numpy()
        indptr, indices = eng.plan.structure()
        n = eng.n
        full = sparse.csr_matrix((sig, indices, indptr), shape=(n, n))
        if self.useCov:
            if self.grid.isExtended:
                raise NotImplementedError("qinv with covariates on an extended grid: the reference inverts "
                                          "S Q S^T, whose lattice part is then not the identity")
            B, C = self._border
            X = self.grid.cov_columns()
            k = C.shape[0]
            W = np.hstack([B, X])
            Z = np.block([[np.zeros((k, k)), np.eye(k)], [np.eye(k), C]])
            QW = fac.solve_A(W)
            inner = np.linalg.inv(Z) + W.T @ QW
            corr = np.einsum("ij,jk,ik->i", QW, np.linalg.inv(inner), QW)
            self.pQinv = None
            # the reference's pQinv has the shape of Q (n + k): zeros past the lattice
            self.mvar = np.concatenate([full.diagonal() - corr, np.zeros(k)])
            return self.mvar
        pattern = self.Q.tocsc()
        mask = sparse.csc_matrix((np.ones(pattern.nnz), pattern.indices, pattern.indptr), shape=pattern.shape)
        self.pQinv = full.tocsc().multiply(mask).tocsc()
        self.mvar = self.pQinv.diagonal()
        return self.mvar

    def update(self, y, idx, tau=None):
        """Q <- Q + tau S^T S, mu <- mu + tau Q^-1 S^T (y - S mu) (model.py:120-127);
        with covariates S carries the border columns and the update adds
        tau S_l^T X to B and tau X^T X to C."""
        if tau is None:
            tau = self.tau
        S = self.grid.getS(idx)
        if self.useCov:
            n = self._n_lattice()
            Sl = S[:, :n]
            X = S[:, n:].toarray()
            add = np.asarray((Sl.T @ Sl).diagonal(), dtype=np.float64).ravel() * tau
            self._diag = add if self._diag is None else self._diag + add
            B, C = self._border
            self._border = (B + tau * np.asarray(Sl.T @ X), C + tau * X.T @ X)
            self.Q = (self.Q + S.transpose() @ S * tau).tocsc()
            self.mu = self.mu + self._solve_aug(S.T @ (y - S @ self.mu)) * tau
            return
        add = np.asarray((S.T @ S).diagonal(), dtype=np.float64) * tau
        self._diag = add if self._diag is None else self._diag + add
        self.Q = (self.Q + S.transpose() @ S * tau).tocsc()
        self.Q_fac = self._factor()
        self.mu = self.mu + self.Q_fac.solve_A(S.T @ (y - S @ self.mu)) * tau

    def getPars(self, onlySelf=True):
        return self.mod.getPars(onlySelf=onlySelf)
