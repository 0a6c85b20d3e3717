"""Precision matrices, parameter derivatives, log-likelihood and sampling.

Oracle restatement (test infrastructure) of
  * ``spdes/whittle_matern{,_anisotropic,_ha}2D.py``  makeQ  (initial-state Q0)
  * ``spdes/advection_{i,,ha_}diffusion2D.py``       makeQ / logLike (stationary)
  * ``spdes/var_advection_var_diffusion2D.py``         makeQ / logLike (B-spline fields)
  * ``model.py:73-153``                                sample / update
using scipy.sparse operations in the reference's order, so the sparsity
pattern (scipy drops exact zeros in products and sums, keeps them in COO
construction and ``bmat``) is the reference's.
"""
import numpy as np
from scipy import sparse
from scipy.sparse import linalg as spla

from . import stencils as st


# --------------------------------------------------------------------- H maps
def h_map(kind, gamma, vx=0.0, vy=0.0):
    """2x2 diffusion tensor and its derivatives w.r.t. (log gamma, vx, vy).

    iso: gamma I  (``advection_idiffusion2D.py:97``)
    ani: gamma I + v v^T  (``advection_diffusion2D.py:100``)
    ha : gamma (cosh|v| I + sinh|v|/|v| [[vx,vy],[vy,-vx]])  (``advection_ha_diffusion2D.py:109-115``)
    """
    I = np.eye(2)
    if kind == "iso":
        return gamma * I, [gamma * I]
    if kind == "ani":
        v = np.array([vx, vy])
        H = gamma * I + v[:, None] * v[None, :]
        d = []
        for e in (np.array([1.0, 0.0]), np.array([0.0, 1.0])):
            d.append(e[:, None] * v[None, :] + v[:, None] * e[None, :])
        return H, [gamma * I] + d
    aV = np.sqrt(vx ** 2 + vy ** 2)
    mV = np.array([[vx, vy], [vy, -vx]])
    ch = (np.exp(aV) + np.exp(-aV)) / 2
    sh = (np.exp(aV) - np.exp(-aV)) / 2
    H = gamma * (ch * I + sh / aV * mV)
    dvx = gamma / aV * (vx * sh * I + vx / aV * (ch - sh / aV) * mV + sh * np.array([[1, 0], [0, -1]]))
    dvy = gamma / aV * (vy * sh * I + vy / aV * (ch - sh / aV) * mV + sh * np.array([[0, 1], [1, 0]]))
    return H, [H, dvx, dvy]


def ah_const_csc(grid2, H, bc):
    M, N = grid2.shape[0], grid2.shape[1]
    return st.triplets_to_csc(*st.ah_const(M, N, H, grid2.hx, grid2.hy, bc), M * N)


def ah_var_csc(grid2, H, bc):
    M, N = grid2.shape[0], grid2.shape[1]
    return st.triplets_to_csc(*st.ah_var(M, N, H, grid2.hx, grid2.hy, bc), M * N)


def aw_const_csc(grid2, G, bc, diff=3):
    M, N = grid2.shape[0], grid2.shape[1]
    return st.triplets_to_csc(*st.aw_const(M, N, G, grid2.hx, grid2.hy, diff, bc), M * N)


def aw_var_csc(grid2, G, bc, diff=3, dG=None):
    M, N = grid2.shape[0], grid2.shape[1]
    return st.triplets_to_csc(*st.aw_var(M, N, G, grid2.hx, grid2.hy, diff, bc, dG), M * N,
                              nan_to_zero=True)


# ------------------------------------------------------------ Whittle-Matern
class WhittleMatern:
    """Spatial Matern precision Q0 = A0^T iDv A0, A0 = Dv kappa - A_H(H0)
    (``whittle_matern2D.py:59-85``, ``whittle_matern_anisotropic2D.py:61-104``,
    ``whittle_matern_ha2D.py:68-110``).  par = [log kappa, log gamma, (vx, vy), log tau]."""

    def __init__(self, grid, kind="iso", par=None, bc=3):
        self.grid, self.kind, self.bc = grid, kind, bc
        self.npar = 3 if kind == "iso" else 5
        self.par = np.asarray(par, dtype=np.float64) if par is not None else None
        self.Q = None
        if par is not None:
            self.Q, _ = self.makeQ(self.par, grad=False)

    def getPars(self):
        return self.par

    def A0(self, par):
        g = self.grid
        kappa, gamma = np.exp(par[0]), np.exp(par[1])
        vx, vy = (par[2], par[3]) if self.kind != "iso" else (0.0, 0.0)
        H, dH = h_map(self.kind, gamma, vx, vy)
        Dk = kappa * sparse.eye(g.Ns)
        return g.Dv @ Dk - ah_const_csc(g, H, self.bc), H, dH, Dk

    def makeQ(self, par, grad=True):
        g = self.grid
        A, H, dH, Dk = self.A0(par)
        Q = A.transpose() @ g.iDv @ A
        if not grad:
            return Q, None
        dQ = []
        dA = (g.Dv @ Dk).tocsc()
        dQ.append(dA.T @ g.iDv @ A + A.T @ g.iDv @ dA)
        for dh in dH:
            dA = -ah_const_csc(g, dh, self.bc)
            dQ.append(dA.T @ g.iDv @ A + A.T @ g.iDv @ dA)
        return Q, dQ


class VarWhittleMatern:
    """Variable-coefficient spatial Matern precision (``var_whittle_matern2D.py:
    61-66,131-152``, ``var_whittle_matern_anisotropic2D.py``,
    ``var_whittle_matern_ha2D.py:83-123``): Q = A^T iDv A, A = Dv Dk - A_H(H)
    with the variable diffusion stencil; kappa = exp(bs c_k) at the centres,
    gamma = exp(bsH c_g) on the faces and per face H = gamma I (iso),
    gamma I + v v^T (ani), gamma expm([[vx, vy], [vy, -vx]]) (ha, written as
    cosh/sinh).  par = [c_k, c_g, (c_vx, c_vy,) log tau] (Np each)."""

    def __init__(self, grid, kind="iso", bc=3):
        self.grid, self.kind, self.bc = grid, kind, bc
        self.Np = grid.Nbs2

    def split(self, par):
        Np = self.Np
        ck, cg = par[:Np], par[Np:2 * Np]
        cvx, cvy = (par[2 * Np:3 * Np], par[3 * Np:4 * Np]) if self.kind != "iso" else (None, None)
        return ck, cg, cvx, cvy

    def fields(self, par):
        g = self.grid
        ck, cg, cvx, cvy = self.split(par)
        kappa = np.exp(g.evalB(ck))
        gam = np.exp(g.evalBH(cg))
        Ns = g.Ns
        I = np.eye(2)
        H = gam[:, :, None, None] * I
        dgam = H.copy()                       # d H / d log gamma at each face
        dvx = dvy = None
        if self.kind == "ani":
            vx, vy = g.evalBH(cvx), g.evalBH(cvy)
            vv = np.stack([vx, vy], axis=2)
            H = H + vv[:, :, :, None] * vv[:, :, None, :]
            ex = np.zeros_like(vv); ex[:, :, 0] = 1.0
            ey = np.zeros_like(vv); ey[:, :, 1] = 1.0
            dvx = vv[:, :, :, None] * ex[:, :, None, :] + ex[:, :, :, None] * vv[:, :, None, :]
            dvy = vv[:, :, :, None] * ey[:, :, None, :] + ey[:, :, :, None] * vv[:, :, None, :]
        elif self.kind == "ha":
            vx, vy = g.evalBH(cvx), g.evalBH(cvy)
            a = np.sqrt(vx ** 2 + vy ** 2)
            mV = np.zeros((Ns, 4, 2, 2))
            mV[:, :, 0, 0], mV[:, :, 0, 1], mV[:, :, 1, 0], mV[:, :, 1, 1] = vx, vy, vy, -vx
            ch, sh = (np.exp(a) + np.exp(-a)) / 2, (np.exp(a) - np.exp(-a)) / 2
            H = (gam * ch)[:, :, None, None] * I + (gam * sh / a)[:, :, None, None] * mV
            dgam = H.copy()
            Dx = np.zeros_like(mV); Dx[:, :, 0, 0], Dx[:, :, 1, 1] = 1.0, -1.0
            Dy = np.zeros_like(mV); Dy[:, :, 0, 1], Dy[:, :, 1, 0] = 1.0, 1.0
            dvx = ((gam * sh * vx / a)[:, :, None, None] * I + (gam * (ch * vx - sh * vx / a) / a ** 2)[:, :, None, None] * mV
                   + (gam * sh / a)[:, :, None, None] * Dx)
            dvy = ((gam * sh * vy / a)[:, :, None, None] * I + (gam * (ch * vy - sh * vy / a) / a ** 2)[:, :, None, None] * mV
                   + (gam * sh / a)[:, :, None, None] * Dy)
        return kappa, H, dgam, dvx, dvy

    def A(self, par):
        g = self.grid
        kappa, H, dgam, dvx, dvy = self.fields(par)
        return g.Dv @ sparse.diags(kappa) - ah_var_csc(g, H, self.bc), kappa, H, dgam, dvx, dvy

    def makeQ(self, par, grad=True):
        g = self.grid
        A, kappa, H, dgam, dvx, dvy = self.A(par)
        Q = A.transpose() @ g.iDv @ A
        if not grad:
            return Q, None
        dQ = []

        def add(dA):
            dQ.append((dA.transpose() @ g.iDv @ A + A.transpose() @ g.iDv @ dA).tocsc())
        for i in range(self.Np):
            add(g.Dv @ sparse.diags(g.bs[:, i] * kappa))
        for d in [dgam] + ([dvx, dvy] if self.kind != "iso" else []):
            for i in range(self.Np):
                add(-ah_var_csc(g, g.bsH[:, :, i][:, :, None, None] * d, self.bc))
        return Q, dQ


# --------------------------------------------------------- space-time blocks
class Separable:
    """Separable space-time precision Q = Qt (x) Qs
    (``seperable_spatial_temporal2D.py:64-124``, ``_ha2D.py``,
    ``_idiffusion2D.py:62-98``).  Qs is the variable Whittle-Matern precision
    of the same kind (A = Dv Dk - A_H(H), Qs = A^T iDv A); Qt is tridiagonal:
      ani:      par[.] = log rho,  Qt = tridiag(-rho; 1, 1+rho^2, .., 1+rho^2, 1) / (1 - rho^2)
      iso, ha:  par[.] = a, log sigma,  Qt = sigma tridiag(-a; 1, 1+a^2, .., 1+a^2, 1)
    (``makeQt``, ``:186-211`` and ``_ha2D.py:202-225``; the latter's loop is
    written for T = 10, the form here holds for every T).  par = [spatial
    fields (2 or 4 Np), temporal (1 or 2), log tau]; dQ in the same order."""

    def __init__(self, grid, kind="ani", bc=3):
        from .grid import SpatialGrid
        self.grid, self.kind, self.bc = grid, kind, bc
        self.sgrid = SpatialGrid(grid.x, grid.y, extend=grid.Ne if grid.Ne else None, Nbs=grid.Nbs)
        self.wm = VarWhittleMatern(self.sgrid, kind, bc)
        self.Np = grid.Nbs2
        self.ns = self.Np * (2 if kind == "iso" else 4)

    def qt(self, tp, T, diff=False):
        """Qt, or its derivative in log rho (ani) / in a (iso, ha)."""
        d, o = np.zeros(T), np.zeros(T - 1)
        if self.kind == "ani":
            rho = np.exp(tp[0])
            den = 1.0 - rho ** 2
            if diff:
                d[:] = 4 * rho ** 2 / den ** 2
                d[0] = d[-1] = 2 * rho ** 2 / den ** 2
                o[:] = -rho * (1 + rho ** 2) / den ** 2
            else:
                d[:] = (1 + rho ** 2) / den
                d[0] = d[-1] = 1.0 / den
                o[:] = -rho / den
        else:
            a, sig = tp[0], np.exp(tp[1])
            if diff:
                d[:] = 2 * a * sig
                d[0] = d[-1] = 0.0
                o[:] = -sig
            else:
                d[:] = (1 + a ** 2) * sig
                d[0] = d[-1] = sig
                o[:] = -a * sig
        return sparse.diags([o, d, o], [-1, 0, 1]).tocsc()

    def makeQ(self, par, grad=True):
        par = np.asarray(par, dtype=np.float64)
        T = self.grid.T
        Qs, dQs = self.wm.makeQ(np.hstack([par[:self.ns], [0.0]]), grad)
        tp = par[self.ns:-1]
        Qt = self.qt(tp, T)
        Q = sparse.kron(Qt, Qs).tocsc()
        if not grad:
            return Q, None
        dQ = [sparse.kron(Qt, d).tocsc() for d in dQs]
        dQ.append(sparse.kron(self.qt(tp, T, diff=True), Qs).tocsc())
        if self.kind != "ani":
            dQ.append(Q.copy())
        return Q, dQ


def assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt, diag0=None):
    """Block-tridiagonal space-time precision (``advection_idiffusion2D.py:110-115``):
        Q00 = sigma dt Q0 + Qs,  Qtt = A^T iDv Qs iDv A + Qs,  Q_{T-1} = A^T iDv Qs iDv A,
        Q_{t,t+1} = -Qs iDv A,   Q_{t+1,t} = -A^T iDv Qs,  all scaled by 1/(dt sigma).
    One ``bmat`` call; the per-block patterns are those of the reference's
    incremental concatenation."""
    up = -Qs @ iDv @ A
    lo = -A.T @ iDv @ Qs
    mid = A.T @ iDv @ Qs @ iDv @ A + Qs
    last = A.T @ iDv @ Qs @ iDv @ A
    first = sigma * dt * Q0 + Qs if diag0 is None else diag0
    blocks = [[None] * T for _ in range(T)]
    for t in range(T):
        if t == 0:
            blocks[t][t] = first
        elif t == T - 1:
            blocks[t][t] = last
        else:
            blocks[t][t] = mid
        if t + 1 < T:
            blocks[t][t + 1] = up
            blocks[t + 1][t] = lo
    return sparse.bmat(blocks).tocsc()


def d_blocks(A, Qs, iDv, T, dA, dQs=None):
    """Derivative of the block formula for a perturbation (dA, dQs)
    (``advection_idiffusion2D.py:120-155``), unscaled."""
    Ns = A.shape[0]
    if dQs is None:
        first = sparse.csc_matrix((Ns, Ns))
        up = -Qs @ iDv @ dA
        lo = -dA.T @ iDv @ Qs
        mid = dA.T @ iDv @ Qs @ iDv @ A + A.T @ iDv @ Qs @ iDv @ dA
        last = mid
    else:
        first = dQs
        up = -dQs @ iDv @ A - Qs @ iDv @ dA
        lo = -dA.T @ iDv @ Qs - A.T @ iDv @ dQs
        last = dA.T @ iDv @ Qs @ iDv @ A + A.T @ iDv @ dQs @ iDv @ A + A.T @ iDv @ Qs @ iDv @ dA
        mid = last + dQs
    blocks = [[None] * T for _ in range(T)]
    for t in range(T):
        blocks[t][t] = first if t == 0 else (last if t == T - 1 else mid)
        if t + 1 < T:
            blocks[t][t + 1] = up
            blocks[t + 1][t] = lo
    return sparse.bmat(blocks).tocsc()


def pad_q0(dQ0, Ns, T):
    """mod0 derivative placed in block (0,0) (``advection_idiffusion2D.py:157-162``)."""
    blocks = [[None] * T for _ in range(T)]
    blocks[0][0] = dQ0
    for t in range(1, T):
        blocks[t][t] = sparse.csc_matrix((Ns, Ns))
    return sparse.bmat(blocks).tocsc()


class _STBase:
    def _q0(self, par_mod0, grad):
        if par_mod0 is not None:
            return self.mod0.makeQ(par_mod0, grad=grad)
        return self.mod0.Q, None


class StationaryAdvDiff(_STBase):
    """Stationary advection-diffusion, kind in {iso, ani, ha}.
    par (iso) = [log kappa, log gamma, wx, wy, log sigma, (mod0...), log tau]
    par (ani/ha) = [log kappa, log gamma, vx, vy, wx, wy, log sigma, (mod0...), log tau]."""

    def __init__(self, grid, mod0, kind="iso", bc=3):
        self.grid, self.mod0, self.kind, self.bc = grid, mod0, kind, bc
        self.nself = 6 if kind == "iso" else 8

    def split(self, par):
        par = np.asarray(par, dtype=np.float64)
        if self.kind == "iso":
            kap, gam, vx, vy, ws, sig = par[0], par[1], 0.0, 0.0, par[2:4], par[4]
            p0 = par[5:] if par.size > 6 else None
        else:
            kap, gam, vx, vy, ws, sig = par[0], par[1], par[2], par[3], par[4:6], par[6]
            p0 = par[7:] if par.size > 8 else None
        return kap, gam, vx, vy, ws, sig, p0

    def makeQ(self, par, grad=True):
        g = self.grid
        kap, gam, vx, vy, ws, sig = self.split(par)[:6]
        p0 = self.split(par)[6]
        dt, T, Ns, Dv, iDv = g.dt, g.T, g.Ns, g.Dv, g.iDv
        kappa, gamma, sigma = np.exp(kap), np.exp(gam), np.exp(sig)
        H, dH = h_map(self.kind, gamma, vx, vy)
        Dk = kappa * sparse.eye(Ns)
        A_H = ah_const_csc(g, H, self.bc)
        As = Dv @ Dk
        Qs = As.transpose() @ iDv @ As
        A = Dv + Dv @ Dk * dt - A_H * dt + aw_const_csc(g, ws, self.bc) * dt
        Q0, dQ0 = self._q0(p0, grad)
        Q = 1 / (dt * sigma) * assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt)
        if not grad:
            return Q, None
        c = 1 / (dt * sigma)
        dQ = []
        dA = (Dv @ Dk * dt).tocsc()
        dAs = Dv @ Dk
        dQs = dAs.T @ iDv @ As + As.T @ iDv @ dAs
        dQ.append(c * d_blocks(A, Qs, iDv, T, dA, dQs))
        for dh in dH:
            dQ.append(c * d_blocks(A, Qs, iDv, T, -ah_const_csc(g, dh, self.bc) * dt))
        dQ.append(c * d_blocks(A, Qs, iDv, T, aw_const_csc(g, ws, self.bc, diff=1) * dt))
        dQ.append(c * d_blocks(A, Qs, iDv, T, aw_const_csc(g, ws, self.bc, diff=2) * dt))
        full = assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt, diag0=Qs)
        dQ.append(-c * full)
        if dQ0 is not None:
            dQ += [pad_q0(d, Ns, T) for d in dQ0]
        return Q, dQ


class VarAdvVarDiff(_STBase):
    """Non-stationary model with B-spline fields (``var_advection_var_diffusion2D.py``).
    par = [c_kappa(Np), c_gamma(Np), c_vx(Np), c_vy(Np), c_wx(Np), c_wy(Np), log sigma,
           (mod0...), log tau]."""

    def __init__(self, grid, mod0, bc=3):
        self.grid, self.mod0, self.bc = grid, mod0, bc
        self.Np = grid.Nbs2

    def fields(self, par):
        g, Np = self.grid, self.Np
        kappa = np.exp(g.evalB(par[:Np]))
        gamma = np.exp(g.evalBH(par[Np:2 * Np]))
        vx = g.evalBH(par[2 * Np:3 * Np])
        vy = g.evalBH(par[3 * Np:4 * Np])
        ws = g.evalAdv(par[4 * Np:6 * Np])
        return kappa, gamma, vx, vy, ws

    @staticmethod
    def h_field(gamma, vx, vy):
        vv = np.stack([vx, vy], axis=2)
        return (np.eye(2) * np.stack([gamma, gamma], axis=2)[:, :, :, None]
                + vv[:, :, :, None] * vv[:, :, None, :])

    def makeQ(self, par, grad=True):
        Q = self._operators(par)[0]
        if not grad:
            return Q, None
        return Q, list(self.dQ_iter(par))

    def _operators(self, par):
        g, Np = self.grid, self.Np
        par = np.asarray(par, dtype=np.float64)
        dt, T, Ns, Dv, iDv = g.dt, g.T, g.Ns, g.Dv, g.iDv
        kappa, gamma, vx, vy, ws = self.fields(par)
        sigma = np.exp(par[Np * 6])
        Hs = self.h_field(gamma, vx, vy)
        Dk = sparse.diags(kappa).tocsc()
        As = Dv @ Dk
        Qs = As.transpose() @ iDv @ As
        A = Dv + (Dv @ Dk - ah_var_csc(g, Hs, self.bc) + aw_var_csc(g, ws, self.bc)) * dt
        p0 = par[Np * 6 + 1:] if par.size > Np * 6 + 2 else None
        Q0 = self._q0(p0, False)[0]
        Q = assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt).tocsc() / (dt * sigma)
        return Q, A, As, Qs, Q0, kappa, gamma, vx, vy, ws, sigma, p0

    def dQ_iter(self, par):
        """The per-parameter derivatives dQ_i one at a time, in the parameter
        order (``var_advection_var_diffusion2D.py:118-197``)."""
        g, Np = self.grid, self.Np
        par = np.asarray(par, dtype=np.float64)
        dt, T, Ns, Dv, iDv = g.dt, g.T, g.Ns, g.Dv, g.iDv
        Q, A, As, Qs, Q0, kappa, gamma, vx, vy, ws, sigma, p0 = self._operators(par)
        c = dt * sigma
        for i in range(Np):
            dk = sparse.diags(g.bs[:, i] * kappa).tocsc()
            dA = dt * Dv @ dk
            dAs = Dv @ dk
            dQs = dAs.T @ iDv @ As + As.T @ iDv @ dAs
            yield d_blocks(A, Qs, iDv, T, dA, dQs) / c
        for i in range(Np):
            dHs = np.eye(2) * (np.stack([g.bsH[:, :, i] * gamma] * 2, axis=2)[:, :, :, None])
            yield d_blocks(A, Qs, iDv, T, -ah_var_csc(g, dHs, self.bc) * dt) / c
        vv = np.stack([vx, vy], axis=2)
        for comp in (0, 1):
            for i in range(Np):
                e = np.zeros(Np)
                e[i] = 1.0
                b = g.evalBH(e)
                z = np.zeros_like(b)
                dv = np.stack([b, z] if comp == 0 else [z, b], axis=2)
                dHs = vv[:, :, :, None] * dv[:, :, None, :] + dv[:, :, :, None] * vv[:, :, None, :]
                yield d_blocks(A, Qs, iDv, T, -ah_var_csc(g, dHs, self.bc) * dt) / c
        for i in range(2 * Np):
            e = np.zeros(2 * Np)
            e[i] = 1.0
            dws = g.evalAdv(e)
            diff = 1 if i < Np else 2
            yield d_blocks(A, Qs, iDv, T, aw_var_csc(g, ws, self.bc, diff, dws) * dt) / c
        full = assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt, diag0=Qs)
        yield -full.tocsc() / c
        if p0 is not None:
            for d in self._q0(p0, True)[1]:
                yield pad_q0(d, Ns, T)

    def spatial_operator(self, par):
        """A (Ns x Ns) and q = diag(Qs) for the factorised form."""
        g, Np = self.grid, self.Np
        kappa, gamma, vx, vy, ws = self.fields(par)
        Dk = sparse.diags(kappa).tocsc()
        A = g.Dv + (g.Dv @ Dk - ah_var_csc(g, self.h_field(gamma, vx, vy), self.bc)
                    + aw_var_csc(g, ws, self.bc)) * g.dt
        return A.tocsc(), g.V * kappa ** 2, np.exp(par[Np * 6])



def _face_tensors(kind, gamma, vx, vy):
    """Per-face H and its derivatives in (log gamma, vx, vy) for B-spline
    diffusion fields (the H maps of the advection-var-* and var-advection-var-*
    files): gamma I, gamma I + v v^T, gamma (cosh a I + sinh a / a mV)."""
    I = np.eye(2)
    H = gamma[:, :, None, None] * I
    if kind == "iso":
        return H, [H.copy()]
    if kind == "ani":
        vv = np.stack([vx, vy], axis=2)
        d = [H.copy()]
        for a in range(2):
            e = np.zeros_like(vv)
            e[:, :, a] = 1.0
            d.append(vv[:, :, :, None] * e[:, :, None, :] + e[:, :, :, None] * vv[:, :, None, :])
        return H + vv[:, :, :, None] * vv[:, :, None, :], d
    a = np.sqrt(vx ** 2 + vy ** 2)
    mV = np.zeros(vx.shape + (2, 2))
    mV[..., 0, 0], mV[..., 0, 1], mV[..., 1, 0], mV[..., 1, 1] = vx, vy, vy, -vx
    ch, sh = (np.exp(a) + np.exp(-a)) / 2, (np.exp(a) - np.exp(-a)) / 2
    H = (gamma * ch)[..., None, None] * I + (gamma * sh / a)[..., None, None] * mV
    Dx = np.zeros_like(mV)
    Dx[..., 0, 0], Dx[..., 1, 1] = 1.0, -1.0
    Dy = np.zeros_like(mV)
    Dy[..., 0, 1], Dy[..., 1, 0] = 1.0, 1.0
    d = [H.copy()]
    for u, D in ((vx, Dx), (vy, Dy)):
        d.append((gamma * sh * u / a)[..., None, None] * I + (gamma * (ch * u - sh * u / a) / a ** 2)[..., None, None] * mV
                 + (gamma * sh / a)[..., None, None] * D)
    return H, d


class FieldAdvDiff(_STBase):
    """The families between the stationary and the fully variable model:
    advection_var_{i,,ha_}diffusion2D.py (B-spline kappa and diffusion,
    constant advection w) and var_advection_{i,,ha_}diffusion2D.py,
    var_advection_var_{i,ha_}diffusion2D.py (B-spline advection; diffusion
    constant or B-spline).  A = Dv + (Dv Dk - A_H + A_w) dt.  par = [kappa
    block, gamma block, (vx, vy blocks,) w block, log sigma, (mod0...),
    log tau]; a block is Np coefficients when its field is a B-spline field
    (2 Np for w), else one log-scalar (two scalars for w)."""

    def __init__(self, grid, mod0, kind="iso", dvar=True, gvar=True, bc=3, ww=None):
        """ww (Ns x 4 faces E,N,W,S): the cov-advection families
        (cov_advection_{i,,ha_}diffusion2D.py, cov_advection_var_*): the w
        block is one scalar lambda, the advection lambda * ww (variable
        stencil), and its dQ uses the stencil of ww itself, dA = Aw(ww) dt
        (cov_advection_idiffusion2D.py:59,153-159)."""
        self.grid, self.mod0, self.kind, self.bc = grid, mod0, kind, bc
        self.dvar, self.gvar = dvar, gvar and ww is None
        self.ww = None if ww is None else np.asarray(ww, dtype=np.float64)
        self.Np = grid.Nbs2

    def sizes(self):
        nd = self.Np if self.dvar else 1
        nh = 1 if self.kind == "iso" else 3
        if self.ww is not None:
            return nd, nh, 1
        return nd, nh, (2 * self.Np if self.gvar else 2)

    def split(self, par):
        par = np.asarray(par, dtype=np.float64)
        nd, nh, nw = self.sizes()
        k = 0
        blocks = []
        for size in [nd] * (1 + nh) + [nw]:
            blocks.append(par[k:k + size])
            k += size
        sig = par[k]
        p0 = par[k + 1:] if par.size > k + 2 else None
        return blocks, sig, p0

    def makeQ(self, par, grad=True):
        g = self.grid
        dt, T, Ns, Dv, iDv = g.dt, g.T, g.Ns, g.Dv, g.iDv
        blocks, sig, p0 = self.split(par)
        sigma = np.exp(sig)
        ck, cw = blocks[0], blocks[-1]
        if self.dvar:
            kappa = np.exp(g.evalB(ck))
            gam = np.exp(g.evalBH(blocks[1]))
            vx = g.evalBH(blocks[2]) if self.kind != "iso" else None
            vy = g.evalBH(blocks[3]) if self.kind != "iso" else None
            Hs, dH = _face_tensors(self.kind, gam, vx, vy)
            A_H = ah_var_csc(g, Hs, self.bc)
        else:
            kappa = np.exp(ck[0]) * np.ones(Ns)
            vx, vy = (blocks[2][0], blocks[3][0]) if self.kind != "iso" else (0.0, 0.0)
            Hc, dHc = h_map(self.kind, np.exp(blocks[1][0]), vx, vy)
            A_H = ah_const_csc(g, Hc, self.bc)
        if self.ww is not None:
            A_w = aw_var_csc(g, cw[0] * self.ww, self.bc)
        elif self.gvar:
            ws = g.evalAdv(cw)
            A_w = aw_var_csc(g, ws, self.bc)
        else:
            A_w = aw_const_csc(g, cw, self.bc)
        Dk = sparse.diags(kappa).tocsc()
        As = Dv @ Dk
        Qs = As.transpose() @ iDv @ As
        A = Dv + (Dv @ Dk - A_H + A_w) * dt
        Q0, dQ0 = self._q0(p0, grad)
        Q = assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt).tocsc() / (dt * sigma)
        if not grad:
            return Q, None
        c = dt * sigma
        dQ = []

        def dkappa(dk):
            dk = sparse.diags(dk).tocsc()
            dAs = Dv @ dk
            dQs = dAs.T @ iDv @ As + As.T @ iDv @ dAs
            dQ.append(d_blocks(A, Qs, iDv, T, dt * Dv @ dk, dQs) / c)
        if self.dvar:
            for i in range(self.Np):
                dkappa(g.bs[:, i] * kappa)
            for d in dH:
                for i in range(self.Np):
                    dQ.append(d_blocks(A, Qs, iDv, T, -ah_var_csc(g, g.bsH[:, :, i][:, :, None, None] * d, self.bc)
                                       * dt) / c)
        else:
            dkappa(kappa)
            for d in dHc:
                dQ.append(d_blocks(A, Qs, iDv, T, -ah_const_csc(g, d, self.bc) * dt) / c)
        if self.ww is not None:
            dQ.append(d_blocks(A, Qs, iDv, T, aw_var_csc(g, self.ww, self.bc) * dt) / c)
        elif self.gvar:
            for i in range(2 * self.Np):
                e = np.zeros(2 * self.Np)
                e[i] = 1.0
                dQ.append(d_blocks(A, Qs, iDv, T, aw_var_csc(g, ws, self.bc, 1 if i < self.Np else 2, g.evalAdv(e))
                                   * dt) / c)
        else:
            for diff in (1, 2):
                dQ.append(d_blocks(A, Qs, iDv, T, aw_const_csc(g, cw, self.bc, diff=diff) * dt) / c)
        full = assemble_blocks(A, Qs, Q0, iDv, T, sigma, dt, diag0=Qs)
        dQ.append(-full.tocsc() / c)
        if dQ0 is not None:
            dQ += [pad_q0(d, Ns, T) for d in dQ0]
        return Q, dQ

def stationary_spatial_operator(model, par):
    g = model.grid
    kap, gam, vx, vy, ws, sig, _ = model.split(par)
    kappa, gamma = np.exp(kap), np.exp(gam)
    H, _ = h_map(model.kind, gamma, vx, vy)
    Dk = kappa * sparse.eye(g.Ns)
    A = g.Dv + g.Dv @ Dk * g.dt - ah_const_csc(g, H, model.bc) * g.dt \
        + aw_const_csc(g, ws, model.bc) * g.dt
    return A.tocsc(), np.full(g.Ns, g.V * kappa ** 2), np.exp(sig)


# ------------------------------------------------------------ factor helpers
def logdet_spd(Q):
    """log|Q| of a sparse SPD matrix: dense Cholesky for small n, else sparse LU
    with symmetric pivoting (stand-in for CHOLMOD's ``Factor.logdet``)."""
    n = Q.shape[0]
    if n <= 4000:
        L = np.linalg.cholesky(Q.toarray())
        return 2.0 * np.log(np.diag(L)).sum()
    lu = spla.splu(Q.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                   options=dict(SymmetricMode=True))
    return np.log(np.abs(lu.U.diagonal())).sum()


def logdet_closed_form(A, q, A0, V, T, sigma, dt):
    """log|Q| for Q = F^T D F (F block-bidiagonal with iDv^1/2 A0 and iDv A on
    the diagonal, -I below; D = blockdiag(I, I_(T-1) x Qs/(dt sigma))):
        2 log|det A0| - Ns log V + 2(T-1)(log|det A| - Ns log V)
        + (T-1) sum_k log(q_k / (dt sigma))."""
    Ns = A.shape[0]

    def logabsdet(M):
        lu = spla.splu(M.tocsc(), permc_spec="COLAMD")
        return np.log(np.abs(lu.U.diagonal())).sum()
    return (2 * logabsdet(A0) - Ns * np.log(V) + 2 * (T - 1) * (logabsdet(A) - Ns * np.log(V))
            + (T - 1) * np.log(q / (dt * sigma)).sum())


def sample_fform(A, q, A0, V, T, sigma, dt, z):
    """Ordering-free exact sampler x = F^{-1} D^{-1/2} z:
       x_0 = sqrt(V) A0^{-1} z_0,  x_t = V A^{-1} (x_{t-1} + sqrt(dt sigma / q) z_t)."""
    Ns = A.shape[0]
    z = z.reshape(T, Ns, -1)
    luA0 = spla.splu(A0.tocsc())
    luA = spla.splu(A.tocsc())
    x = np.empty_like(z)
    x[0] = np.sqrt(V) * luA0.solve(z[0])
    s = np.sqrt(dt * sigma / q)[:, None]
    for t in range(1, T):
        x[t] = V * luA.solve(x[t - 1] + s * z[t])
    return x.reshape(T * Ns, -1)


# ----------------------------------------------------------- log-likelihood
def loglike(Q, dQ, S, data, log_tau, trace="exact", probes=None, dense_inverse=None):
    """Normalised log-likelihood and gradient (``advection_idiffusion2D.py:167-192``):
        like = 1/2 r log|Q| + 1/2 m r log tau - 1/2 r log|Q_c| - 1/2 sum mu_c o (Q mu_c)
               - tau/2 ||y - S mu_c||^2,     Q_c = Q + tau S^T S,  mu_c = Q_c^{-1} tau S^T y
        g_i  = 1/2 r tr((Q^{-1} - Q_c^{-1}) dQ_i) - 1/2 mu_c^T dQ_i mu_c
        g_tau = m r/2 - 1/2 r tr(Q_c^{-1} tau S^T S) - tau/2 ||y - S mu_c||^2
    returning (-like/(m r), -g/(m r)).  trace="exact" uses dense inverses
    (small n); trace="hutchinson" uses the probe matrix V (n x nh1) exactly as
    the reference's estimator (``:182-189``)."""
    tau = np.exp(log_tau)
    y = data.reshape(data.shape[0], -1)
    m, r = y.shape
    StS = (S.T @ S).tocsc()
    Qc = (Q + StS * tau).tocsc()
    n = Q.shape[0]
    if n <= 4000:
        Qd, Qcd = Q.toarray(), Qc.toarray()
        Lc = np.linalg.cholesky(Qcd)
        mu = np.linalg.solve(Lc.T, np.linalg.solve(Lc, (S.T @ y) * tau))
        ld, ldc = logdet_spd(Q), 2.0 * np.log(np.diag(Lc)).sum()
        solveQ = lambda b: np.linalg.solve(Qd, b)
        solveQc = lambda b: np.linalg.solve(Qcd, b)
    else:
        luQ = spla.splu(Q.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                        options=dict(SymmetricMode=True))
        luQc = spla.splu(Qc, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                         options=dict(SymmetricMode=True))
        mu = luQc.solve((S.T @ y) * tau)
        ld = np.log(np.abs(luQ.U.diagonal())).sum()
        ldc = np.log(np.abs(luQc.U.diagonal())).sum()
        solveQ, solveQc = luQ.solve, luQc.solve
    resid = y - S @ mu
    like = (0.5 * ld * r + m * r * np.log(tau) / 2 - 0.5 * ldc * r
            - 0.5 * (mu * (Q @ mu)).sum() - tau / 2 * (resid ** 2).sum())
    g = np.zeros(len(dQ) + 1)
    if trace == "exact":
        Sig = np.linalg.inv(Q.toarray()) if dense_inverse is None else dense_inverse[0]
        Sigc = np.linalg.inv(Qc.toarray()) if dense_inverse is None else dense_inverse[1]
        D = Sig - Sigc
        for i, d in enumerate(dQ):
            tr = d.multiply(D).sum()
            g[i] = 0.5 * tr * r - 0.5 * (mu * (d @ mu)).sum()
        trc = (StS * tau).multiply(Sigc).sum()
        g[-1] = m * r / 2 - 0.5 * trc * r - tau / 2 * (resid ** 2).sum()
    else:
        Vp = probes
        nh1 = Vp.shape[1]
        TrQ = solveQ(Vp)
        TrQc = solveQc(Vp)
        for i, d in enumerate(dQ):
            g[i] = 0.5 * ((TrQ - TrQc) * (d @ Vp)).sum() * r / nh1 - 0.5 * (mu * (d @ mu)).sum()
        g[-1] = (m * r / 2 - 0.5 * (TrQc * (StS * tau @ Vp)).sum() * r / nh1
                 - tau / 2 * (resid ** 2).sum())
    return -like / (m * r), -g / (m * r), mu


def qinv_columns(A, q, A0, V, T, sigma, dt, cols):
    """W = Q^{-1} S^T for the unit columns S^T e_k = e_{cols[k]}, through the
    factorised form Q = F^T D F (``sample_fform``'s F and D):
        F^T u = e:   u_t = F_tt^{-T} (e_t + u_{t+1})   (t = T-1 .. 0),
        F w = D^{-1} u:  w_0 = F_00^{-1} u_0,  w_t = F_tt^{-1} (D_t^{-1} u_t + w_{t-1}),
    with F_00 = V^{-1/2} A0, F_tt = A / V, D_0 = I, D_t = q / (dt sigma), using
    dense inverses of the two spatial operators (Ns x Ns).  Returns W (n x m)."""
    Ns = A.shape[0]
    n = Ns * T
    cols = np.asarray(cols)
    m = cols.size
    iA = np.linalg.inv(A.toarray())
    iA0 = np.linalg.inv(A0.toarray())
    iF = [np.sqrt(V) * iA0] + [V * iA] * (T - 1)
    W = np.zeros((n, m))
    cell, slc = cols % Ns, cols // Ns
    carry = np.zeros((Ns, m))
    for t in range(T - 1, -1, -1):
        e = carry
        sel = np.nonzero(slc == t)[0]
        e[cell[sel], sel] += 1.0
        carry = iF[t].T @ e
        W[t * Ns:(t + 1) * Ns] = carry
    dinv = np.concatenate([np.ones(Ns), np.tile(dt * sigma / q, T - 1)])
    prev = np.zeros((Ns, m))
    for t in range(T):
        blk = W[t * Ns:(t + 1) * Ns] * dinv[t * Ns:(t + 1) * Ns, None]
        prev = iF[t] @ (blk + (prev if t > 0 else 0.0))
        W[t * Ns:(t + 1) * Ns] = prev
    return W


def pattern_product(E, M, N, T, rows_block=2048):
    """G = E E^T on the 43-point space-time pattern of Q: a dict mapping each
    lattice offset (dx, dy, dt) (|dx|, |dy| <= 2 in the same slice, <= 1 in
    the neighbouring slices) to an array g[a] = (E E^T)[a, a + offset] (NaN
    where a + offset leaves the lattice).  Blocked dense products over
    consecutive lattice rows."""
    n, m = E.shape
    Ns = M * N
    offs = [(dx, dy, 0) for dy in range(-2, 3) for dx in range(-2, 3)]
    offs += [(dx, dy, d) for d in (-1, 1) for dy in range(-1, 2) for dx in range(-1, 2)]
    out = {o: np.full(n, np.nan) for o in offs}
    a_all = np.arange(n)
    ia, ja, ta = a_all % M, (a_all // M) % N, a_all // Ns
    reach = 2 * M + 2
    for a0 in range(0, n, rows_block):
        a1 = min(n, a0 + rows_block)
        Ea = E[a0:a1]
        for d in (-1, 0, 1):
            b0 = max(0, a0 + d * Ns - reach)
            b1 = min(n, a1 + d * Ns + reach)
            if b0 >= b1:
                continue
            P = Ea @ E[b0:b1].T
            for o in offs:
                if o[2] != d:
                    continue
                dx, dy, dtt = o
                a = a_all[a0:a1]
                ok = ((ia[a0:a1] + dx >= 0) & (ia[a0:a1] + dx < M) & (ja[a0:a1] + dy >= 0)
                      & (ja[a0:a1] + dy < N) & (ta[a0:a1] + dtt >= 0) & (ta[a0:a1] + dtt < T))
                b = a + dx + dy * M + dtt * Ns
                vals = np.full(a1 - a0, np.nan)
                vals[ok] = P[np.nonzero(ok)[0], b[ok] - b0]
                out[o][a0:a1] = vals
    return out


def pattern_inner(G, D, M, N):
    """<D, G> for a sparse D whose entries lie on the 43-point pattern, G from
    ``pattern_product``."""
    D = D.tocoo()
    Ns = M * N
    a, b = D.row.astype(np.int64), D.col.astype(np.int64)
    dx = (b % M) - (a % M)
    dy = ((b // M) % N) - ((a // M) % N)
    dtt = b // Ns - a // Ns
    tot = 0.0
    keys = dx + 5 * dy + 25 * dtt
    for key in np.unique(keys):
        sel = keys == key
        o = (int(dx[sel][0]), int(dy[sel][0]), int(dtt[sel][0]))
        g = G[o][a[sel]]
        if np.isnan(g).any():
            raise ValueError("entry outside the lattice pattern")
        tot += float(np.dot(D.data[sel], g))
    return tot


def loglike_woodbury(Q, dQ, A, q, A0, V, T, sigma, dt, M, N, cols, data, log_tau):
    """The exact log-likelihood and gradient of ``loglike`` (trace="exact") by
    the Woodbury identity instead of factorising Q_c (no reference
    counterpart; it reproduces ``var_advection_var_diffusion2D.py:201-237``
    with the exact trace):
        W = Q^{-1} S^T,  K = I_m / tau + S W,   Q_c^{-1} = Q^{-1} - W K^{-1} W^T,
        log|Q_c| = log|Q| + m log tau + log|K|,   mu_c = W K^{-1} y,
        tr((Q^{-1} - Q_c^{-1}) dQ_i) = <dQ_i, W K^{-1} W^T>,
        tr(Q_c^{-1} tau S^T S) = tau sum_k [(S W)_kk - (W K^{-1} W^T)_{c_k c_k}],
    log|Q| from the closed form.  dQ may be a generator.  Returns
    (-like/(m r), -g/(m r), mu)."""
    from scipy import linalg as sla
    tau = np.exp(log_tau)
    y = np.asarray(data, dtype=np.float64).reshape(len(cols), -1)
    m, r = y.shape
    cols = np.asarray(cols)
    ld_q = logdet_closed_form(A, q, A0, V, T, sigma, dt)
    W = qinv_columns(A, q, A0, V, T, sigma, dt, cols)
    SW = W[cols]
    K = 0.5 * (SW + SW.T) + np.eye(m) / tau
    LK = np.linalg.cholesky(K)
    ld_qc = ld_q + m * np.log(tau) + 2.0 * np.log(np.diag(LK)).sum()
    mu = W @ sla.cho_solve((LK, True), y)
    resid = y - mu[cols]
    like = (0.5 * ld_q * r + m * r * np.log(tau) / 2 - 0.5 * ld_qc * r
            - 0.5 * (mu * (Q @ mu)).sum() - tau / 2 * (resid ** 2).sum())
    # E = W L_K^{-T}: E E^T = W K^{-1} W^T
    iLK = sla.solve_triangular(LK, np.eye(m), lower=True)
    for a0 in range(0, W.shape[0], 8192):
        W[a0:a0 + 8192] = W[a0:a0 + 8192] @ iLK.T
    G = pattern_product(W, M, N, T)
    g = []
    for d in dQ:
        g.append(0.5 * r * pattern_inner(G, d, M, N) - 0.5 * (mu * (d @ mu)).sum())
    diag_g = G[(0, 0, 0)][cols]
    trc = tau * float((np.diag(SW) - diag_g).sum())
    g.append(m * r / 2 - 0.5 * trc * r - tau / 2 * (resid ** 2).sum())
    g = np.asarray(g)
    return -like / (m * r), -g / (m * r), mu


def reference_probes(n, nh1=100, seed=None):
    """The reference's Rademacher probes: ``2*np.random.randint(1,3,n*nh1)-3``
    from the global legacy RNG, reshaped (n, nh1) (``:182``)."""
    if seed is not None:
        np.random.seed(seed)
    return (2 * np.random.randint(1, 3, n * nh1) - 3).reshape(n, nh1).astype(np.float64)
