"""Regular grids of the drop-in API: ``grid(x, y, z=None, t=None, extend=None, Nbs=3)``.

Mirrors berild/spdepy ``src/spdepy/grids/__init__.py:3-49``,
``grids/spat2Dtemp_regular_mesh.py`` and ``grids/spatial2D_regular_mesh.py``:
cell spacing from the coordinates (``hx = (x.max-x.min)/(M-1)``), global index
``i + j*M' + t*M'N'`` (x fastest), optional border extension, quadratic
B-spline coefficient bases at cell centres (``bs``), diffusion faces W,E,S,N
(``bsH``) and advection faces E,N,W,S (``bsA``), and the observation
operator.  The bases are host arrays; the device copies live in the engine.
"""
import numpy as np
from scipy import sparse


def _quadratic_bsplines(t, lo, hi, nbs):
    """nbs quadratic B-splines on [lo, hi]: uniform knots with two extra
    intervals each side, the two end functions on each side summed
    (``spat2Dtemp_regular_mesh.py:157-187``).  Partition of unity on [lo, hi]."""
    span = hi - lo
    kn = np.linspace(lo - 2.0 * span / nbs, hi + 2.0 * span / nbs, nbs + 5)
    t = np.asarray(t, dtype=np.float64)[:, None]
    last = np.zeros(kn.size - 1, dtype=bool)
    last[-1] = True
    b = ((t >= kn[:-1]) & ((t < kn[1:]) | (last & (t <= kn[1:])))).astype(np.float64)
    for deg in (1, 2):
        left = (t - kn[:-deg - 1]) / (kn[deg:-1] - kn[:-deg - 1])
        right = (kn[deg + 1:] - t) / (kn[deg + 1:] - kn[1:-deg])
        b = left * b[:, :-1] + right * b[:, 1:]
    return np.concatenate([b[:, :2].sum(1, keepdims=True), b[:, 2:-2], b[:, -2:].sum(1, keepdims=True)], 1)


class _Lattice2D:
    sdim = 2

    def _space(self, x, y, Nbs):
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.M, self.N = self.x.size, self.y.size
        self.A = self.x.max() - self.x.min()
        self.B = self.y.max() - self.y.min()
        self.hx = self.A / (self.M - 1)
        self.hy = self.B / (self.N - 1)
        self.V = self.hx * self.hy
        gx, gy = np.meshgrid(self.x, self.y)
        self.sx, self.sy = gx.ravel(), gy.ravel()
        self.Nbs, self.Nbs2 = Nbs, Nbs * Nbs
        self.Ne = 0
        self.isExtended = False

    def _points(self):
        return (self.sxe, self.sye) if self.isExtended else (self.sx, self.sy)

    def basis(self, dx=0.0, dy=0.0):
        px, py = self._points()
        if dx != 0 or dy != 0:
            xr = (self.sx.min() - self.hx / 2, self.sx.max() + self.hx / 2)
            yr = (self.sy.min() - self.hy / 2, self.sy.max() + self.hy / 2)
        else:
            xr = (self.sx.min(), self.sx.max())
            yr = (self.sy.min(), self.sy.max())
        return (_quadratic_bsplines(px + dx, xr[0], xr[1], self.Nbs),
                _quadratic_bsplines(py + dy, yr[0], yr[1], self.Nbs))

    def _tensor(self, dx=0.0, dy=0.0):
        bx, by = self.basis(dx, dy)
        return np.einsum("pi,pj->pij", by, bx).reshape(bx.shape[0], -1)

    def basisN(self):
        self.bs = self._tensor()

    def basisH(self):
        hx, hy = self.hx / 2, self.hy / 2
        self.bsH = np.stack([self._tensor(-hx, 0), self._tensor(hx, 0),
                             self._tensor(0, -hy), self._tensor(0, hy)], axis=1)

    def basisA(self):
        hx, hy = self.hx / 2, self.hy / 2
        self.bsA = np.stack([self._tensor(hx, 0), self._tensor(0, hy),
                             self._tensor(-hx, 0), self._tensor(0, -hy)], axis=1)

    def evalB(self, par, bs=None, d=None):
        if d is not None:
            par = np.zeros(np.shape(par))
            par[d] = 1
        return (self.bs if bs is None else bs) @ par

    def evalBH(self, par, bs=None, d=None):
        if d is not None:
            par = np.zeros(np.shape(par))
            par[d] = 1
        return (self.bsH if bs is None else bs) @ par

    def _extend(self, ne):
        def ext(v, h):
            return np.hstack([np.linspace(v[0] - ne * h, v[0], ne + 1)[:-1], v,
                              np.linspace(v[-1], v[-1] + ne * h, ne + 1)[1:]])
        self.xe, self.ye = ext(self.x, self.hx), ext(self.y, self.hy)
        gx, gy = np.meshgrid(self.xe, self.ye)
        self.sxe, self.sye = gx.ravel(), gy.ravel()
        self.isExtended = True
        self.Ne = ne

    def setDv(self):
        self.Dv = self.V * sparse.eye(self.Ns)
        self.iDv = sparse.eye(self.Ns) / self.V

    # ------------------------------------------------ covariates in S
    def addCov(self, cov, inter=True, scale=False):
        """Dense covariate columns appended to S (``spat2Dtemp_regular_mesh.py:
        60-64,35-51``): with an intercept [1, cov] (or [1, cov / cov.max()]
        when scale), without one [cov].  cov holds one value per interior
        cell (in the interior index order of getS)."""
        self.inter = bool(inter)
        self.scale = bool(scale)
        self.cov = np.asarray(cov, dtype=np.float64).reshape(-1)
        self.S = None

    def addInt(self):
        """An intercept column appended to S (``:66-68``)."""
        self.inter = True
        self.S = None

    def cov_columns(self):
        """The dense columns getS appends (n_interior x k), or None."""
        cov = getattr(self, "cov", None)
        inter = getattr(self, "inter", False)
        if cov is not None:
            if inter:
                return np.stack([np.ones(cov.size), cov / cov.max() if getattr(self, "scale", False) else cov], 1)
            return cov.reshape(-1, 1)
        if inter:
            return np.ones((self._n_interior(), 1))
        return None

    def _with_cov(self, S, rows):
        X = self.cov_columns()
        if X is None:
            return S
        return sparse.hstack([S, sparse.csc_matrix(X[rows])]).tocsc()

    def extend(self, extend=1):
        """Border extension by `extend` cells on every side (``:102-117``):
        extended coordinates, lattice sizes and the centre and diffusion-face
        bases (the advection faces keep the interior support, as in the
        reference).  The volume matrices follow the new size."""
        self._extend(extend)
        self._set_sizes()
        self.basisN()
        self.basisH()
        self.Ae = None
        self.S = None
        self.setDv()


class SpatialGrid(_Lattice2D):
    """2-D grid (``spatial2D_regular_mesh.py``); used for the initial-state model."""
    type = "gridS"

    def __init__(self):
        self.S = None
        self.Ae = None
        self.cov, self.inter, self.scale = None, False, False

    def _n_interior(self):
        return self.M * self.N

    def _set_sizes(self):
        self.Ns = self.n = (self.M + 2 * self.Ne) * (self.N + 2 * self.Ne)

    def setGrid(self, x, y, extend=None, Nbs=3):
        self._space(x, y, Nbs)
        self.basisN()
        self.basisH()
        self.Ns = self.n = self.M * self.N
        if extend is not None:
            self._extend(extend)
            self.Ns = self.n = (self.M + 2 * extend) * (self.N + 2 * extend)
            self.basisN()
            self.basisH()
        self.setDv()

    def getIdx(self, pos, extend=True):
        """Lattice index of the cell (idx x, idx y) (``spatial2D_regular_mesh.py:64-76``):
        on the extended lattice when extend, else on the interior grid."""
        if extend:
            return (pos[0] + self.Ne) + (pos[1] + self.Ne) * (self.M + 2 * self.Ne)
        return pos[0] + pos[1] * self.M

    @property
    def shape(self):
        return [self.M + 2 * self.Ne, self.N + 2 * self.Ne]

    def n2e(self, idx):
        idx = np.asarray(idx)
        return (idx % self.M + self.Ne) + (idx // self.M + self.Ne) * (self.M + 2 * self.Ne)

    def getS(self, idxs=None):
        rows = np.arange(self.M * self.N) if idxs is None else np.asarray(idxs)
        S = sparse.csc_matrix((np.ones(rows.size), (np.arange(rows.size), self.n2e(rows))),
                              shape=(rows.size, int(np.prod(self.shape))))
        return self._with_cov(S, rows)


class Grid(_Lattice2D):
    """Space-time grid (``spat2Dtemp_regular_mesh.py``)."""
    type = "gridST"

    def __init__(self):
        self.S = None
        self.Ae = None
        self.cov, self.inter, self.scale = None, False, False

    def _n_interior(self):
        return self.M * self.N * self.T

    def _set_sizes(self):
        self.Ns = (self.M + 2 * self.Ne) * (self.N + 2 * self.Ne)
        self.n = self.Ns * self.T

    def setGrid(self, x, y, t, extend=None, Nbs=3):
        self._space(x, y, Nbs)
        self.t = np.asarray(t, dtype=np.float64)
        self.T = self.t.size
        self.Tdur = self.t.max() - self.t.min()
        self.dt = self.Tdur / (self.T - 1)
        # interior cell coordinates of index i + j*M + t*M*N (``:137-141``)
        iy, it, ix = np.meshgrid(np.arange(self.N), np.arange(self.T), np.arange(self.M))
        self.ix, self.iy, self.it = ix.ravel(), iy.ravel(), it.ravel()
        self.basisN()
        self.basisH()
        self.basisA()   # advection faces keep the interior support (not rebuilt on extension)
        self.Ns = self.M * self.N
        self.n = self.Ns * self.T
        if extend is not None:
            self._extend(extend)
            self.Ns = (self.M + 2 * extend) * (self.N + 2 * extend)
            self.n = self.Ns * self.T
            self.basisN()
            self.basisH()
        self.setDv()

    @property
    def shape(self):
        return [self.M + 2 * self.Ne, self.N + 2 * self.Ne, self.T]

    def n2e(self, idx):
        """Interior index i + j*M + t*M*N -> lattice index (``:86-95``)."""
        idx = np.asarray(idx)
        MN = self.M * self.N
        t, rem = idx // MN, idx % MN
        Me, Nee = self.M + 2 * self.Ne, self.N + 2 * self.Ne
        return (rem % self.M + self.Ne) + (rem // self.M + self.Ne) * Me + t * Me * Nee

    def getS(self, idxs=None):
        """Observation operator of the interior indices idxs (all when None),
        with the covariate columns of addCov / addInt appended
        (``spat2Dtemp_regular_mesh.py:35-58``)."""
        rows = np.arange(self.M * self.N * self.T) if idxs is None else np.asarray(idxs)
        S = sparse.csc_matrix((np.ones(rows.size), (np.arange(rows.size), self.n2e(rows))),
                              shape=(rows.size, self.n))
        return self._with_cov(S, rows)

    def getIdx(self, pos, extend=True):
        """Lattice index of the cell (idx x, idx y, idx t) (``:73-84``): on the
        extended lattice when extend, else the interior index."""
        if extend:
            return (pos[0] + self.Ne) + (pos[1] + self.Ne) * (self.M + 2 * self.Ne) + pos[2] * self.Ns
        return pos[0] + pos[1] * self.M + pos[2] * self.M * self.N

    def ext_map(self):
        """Border extrapolation of advection faces (``advBound :347-391``):
        weights decay linearly with the distance to the interior."""
        if self.Ae is not None:
            return self.Ae
        M, N, ne = self.M, self.N, self.Ne
        Me, Nee = M + 2 * ne, N + 2 * ne
        Ae = sparse.lil_matrix((self.Ns, M * N))
        for jj in range(Nee):
            for ii in range(Me):
                k = ii + jj * Me
                fx, fy = Me - 1 - ii, Nee - 1 - jj
                lx, hx_, ly, hy_ = ii < ne, ii >= M + ne, jj < ne, jj >= N + ne
                if lx and ly:
                    Ae[k, 0] = min(ii, jj) / ne
                elif hx_ and hy_:
                    Ae[k, M * N - 1] = (fy if N - jj <= M - ii else fx) / ne
                elif hx_ and ly:
                    Ae[k, M - 1] = (fx if fx <= jj else jj) / ne
                elif lx and hy_:
                    Ae[k, M * (N - 1)] = (ii if ii <= fy else fy) / ne
                else:
                    if lx:
                        Ae[k, (jj - ne) * M] = ii / ne
                    elif hx_:
                        Ae[k, M - 1 + (jj - ne) * M] = fx / ne
                    if ly:
                        Ae[k, ii - ne] = jj / ne
                    elif hy_:
                        Ae[k, ii - ne + (N - 1) * M] = fy / ne
                    if not (lx or hx_ or ly or hy_):
                        Ae[k, (ii - ne) + (jj - ne) * M] = 1.0
        self.Ae = Ae.tocsr()
        return self.Ae

    def assimilate_adv(self, we, wn):
        """Face velocities (faces E, N, W, S) from cell-centred currents we, wn
        on the interior M x N cells (``:277-344``): a face takes the mean of its
        cell and the neighbour across it, the cell alone at the lattice border;
        extended grids extrapolate to the border cells as ``advBound`` does."""
        M, N = self.M, self.N
        we = np.asarray(we, dtype=np.float64).reshape(M * N)
        wn = np.asarray(wn, dtype=np.float64).reshape(M * N)
        k = np.arange(M * N)
        i, j = k % M, k // M
        nb = [np.where(i + 1 == M, k, k + 1), np.where(j + 1 == N, k, k + M), np.where(i == 0, k, k - 1),
              np.where(j == 0, k, k - M)]
        ww = np.stack([(w + w[b]) / 2 for w, b in zip((we, wn, we, wn), nb)], axis=1)
        return self.ext_map() @ ww if self.isExtended else ww

    def evalAdv(self, par, bs=None, d=None):
        if d is not None:
            par = np.zeros(np.shape(par))
            par[d] = 1
        bs = self.bsA if bs is None else bs
        k = self.Nbs2
        res = np.stack([bs[:, 0, :] @ par[:k], bs[:, 1, :] @ par[k:], bs[:, 2, :] @ par[:k],
                        bs[:, 3, :] @ par[k:]], axis=1)
        if self.isExtended:
            res = self.ext_map() @ res
        return res


def grid(x, y, z=None, t=None, extend=None, Nbs=3):
    """Grid factory (``grids/__init__.py:3-49``): spatial 2-D or space-time 2-D."""
    if z is not None:
        raise NotImplementedError("3-D grids have no SPDE model in the reference")
    if t is None:
        g = SpatialGrid()
        g.setGrid(x=x, y=y, extend=extend, Nbs=Nbs)
        return g
    g = Grid()
    g.setGrid(x=x, y=y, t=t, extend=extend, Nbs=Nbs)
    return g
