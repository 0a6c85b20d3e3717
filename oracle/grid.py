"""Regular 2-D and space-time grids with quadratic B-spline coefficient bases.

Oracle restatement (test infrastructure) of
``src/spdepy/grids/spat2Dtemp_regular_mesh.py`` and
``src/spdepy/grids/spatial2D_regular_mesh.py``.
"""
import numpy as np
from scipy import sparse


def bspline_1d(t, lo, hi, nbs):
    """Quadratic B-splines on nbs+5 uniform knots over [lo-2L/nbs, hi+2L/nbs],
    the two outermost functions on each side merged (nbs functions left).

    Follows ``spat2Dtemp_regular_mesh.py:157-187`` (Cox-de Boor recursion,
    half-open degree-0 intervals, the last one closed)."""
    L = hi - lo
    knots = np.linspace(lo - 2.0 * L / nbs, hi + 2.0 * L / nbs, nbs + 5)
    nk = knots.size
    t = np.asarray(t, dtype=np.float64)
    # degree 0
    b = np.zeros((t.size, nk - 1))
    for i in range(nk - 1):
        inside = (t >= knots[i]) & (t < knots[i + 1])
        if i == nk - 2:
            inside |= (t >= knots[i]) & (t <= knots[i + 1])
        b[:, i] = inside * 1.0
    for r in (1, 2):
        nb = nk - r - 1
        nxt = np.zeros((t.size, nb))
        for i in range(nb):
            left = (t - knots[i]) / (knots[i + r] - knots[i]) * b[:, i]
            right = (knots[i + r + 1] - t) / (knots[i + r + 1] - knots[i + 1]) * b[:, i + 1]
            nxt[:, i] = left + right
        b = nxt
    return np.hstack([(b[:, 0] + b[:, 1])[:, None], b[:, 2:-2], (b[:, -2] + b[:, -1])[:, None]])


class _Base:
    """Common geometry of the regular 2-D lattice (cell centres x fastest)."""

    def _set_space(self, x, y, nbs):
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.M = self.x.size
        self.N = self.y.size
        self.A = self.x.max() - self.x.min()
        self.B = self.y.max() - self.y.min()
        self.hx = self.A / (self.M - 1)
        self.hy = self.B / (self.N - 1)
        self.V = self.hx * self.hy
        sx, sy = np.meshgrid(self.x, self.y)
        self.sx = sx.ravel()
        self.sy = sy.ravel()
        self.Nbs = nbs
        self.Nbs2 = nbs * nbs
        self.Ne = 0
        self.isExtended = False

    def _eval_points(self):
        return (self.sxe, self.sye) if self.isExtended else (self.sx, self.sy)

    def basis(self, dx=0.0, dy=0.0):
        """1-D factors (bx, by) of the tensor basis at centres shifted by (dx, dy).
        Face bases widen the support by half a cell (``:164-173``)."""
        px, py = self._eval_points()
        if dx != 0 or dy != 0:
            xlo, xhi = self.sx.min() - self.hx / 2, self.sx.max() + self.hx / 2
            ylo, yhi = self.sy.min() - self.hy / 2, self.sy.max() + self.hy / 2
        else:
            xlo, xhi = self.sx.min(), self.sx.max()
            ylo, yhi = self.sy.min(), self.sy.max()
        return (bspline_1d(px + dx, xlo, xhi, self.Nbs),
                bspline_1d(py + dy, ylo, yhi, self.Nbs))

    @staticmethod
    def _tensor(bx, by):
        # column index = iy*Nbs + ix  (``:195-197``)
        return (by[:, :, None] * bx[:, None, :]).reshape(bx.shape[0], -1)

    def basisN(self):
        bx, by = self.basis()
        self.bs = self._tensor(bx, by)

    def _faces(self, shifts):
        cols = [self._tensor(*self.basis(dx=dx, dy=dy)) for dx, dy in shifts]
        return np.stack(cols, axis=1)

    def basisH(self):
        """Diffusion faces in the order W, E, S, N (``:201-224``)."""
        h = (self.hx / 2, self.hy / 2)
        self.bsH = self._faces([(-h[0], 0), (h[0], 0), (0, -h[1]), (0, h[1])])

    def basisA(self):
        """Advection faces in the order E, N, W, S (``:226-249``)."""
        h = (self.hx / 2, self.hy / 2)
        self.bsA = self._faces([(h[0], 0), (0, h[1]), (-h[0], 0), (0, -h[1])])

    def evalB(self, par):
        return self.bs @ par

    def evalBH(self, par):
        return self.bsH @ par

    def _extend_xy(self, ne):
        def ext(v, h):
            lo = np.linspace(v[0] - ne * h, v[0], ne + 1)[:-1]
            hi = np.linspace(v[-1], v[-1] + ne * h, ne + 1)[1:]
            return np.hstack([lo, v, hi])
        self.xe = ext(self.x, self.hx)
        self.ye = ext(self.y, self.hy)
        sxe, sye = np.meshgrid(self.xe, self.ye)
        self.sxe = sxe.ravel()
        self.sye = sye.ravel()
        self.isExtended = True
        self.Ne = ne

    def setDv(self):
        self.Dv = self.V * sparse.eye(self.Ns)
        self.iDv = sparse.eye(self.Ns) / self.V


class SpatialGrid(_Base):
    """Spatial 2-D grid (``spatial2D_regular_mesh.py:109-135``)."""
    type = "gridS"
    sdim = 2

    def __init__(self, x, y, extend=None, Nbs=3):
        self._set_space(x, y, Nbs)
        self.basisN()
        self.basisH()
        self.Ns = self.M * self.N
        self.n = self.Ns
        if extend is not None:
            self._extend_xy(extend)
            self.Ns = (self.M + 2 * extend) * (self.N + 2 * extend)
            self.n = self.Ns
            self.basisN()
            self.basisH()
        self.setDv()
        self._S = None

    @property
    def shape(self):
        return [self.M + 2 * self.Ne, self.N + 2 * self.Ne]

    def n2e(self, idx):
        idx = np.asarray(idx)
        i = idx % self.M
        j = idx // self.M
        return (i + self.Ne) + (j + self.Ne) * (self.M + 2 * self.Ne)

    def getS(self, idxs=None):
        rows = np.arange(self.M * self.N) if idxs is None else np.asarray(idxs)
        cols = self.n2e(rows)
        return sparse.csc_matrix((np.ones(rows.size), (np.arange(rows.size), cols)),
                                 shape=(rows.size, int(np.prod(self.shape))))


class STGrid(_Base):
    """Space-time grid, global index i + j*M' + t*M'N' (``spat2Dtemp_regular_mesh.py``)."""
    type = "gridST"
    sdim = 2

    def __init__(self, x, y, t, extend=None, Nbs=3):
        self._set_space(x, y, Nbs)
        self.t = np.asarray(t, dtype=np.float64)
        self.T = self.t.size
        self.Tdur = self.t.max() - self.t.min()
        self.dt = self.Tdur / (self.T - 1)
        self.basisN()
        self.basisH()
        self.basisA()          # not rebuilt by extend (``:102-117``)
        self.Ns = self.M * self.N
        self.n = self.Ns * self.T
        self.Ae = None
        if extend is not None:
            self._extend_xy(extend)
            self.Ns = (self.M + 2 * extend) * (self.N + 2 * extend)
            self.n = self.Ns * self.T
            self.basisN()
            self.basisH()
        self.setDv()

    @property
    def shape(self):
        return [self.M + 2 * self.Ne, self.N + 2 * self.Ne, self.T]

    def assimilate_adv(self, we, wn):
        """Face velocities from cell-centred currents (``:277-344``): each face
        takes the mean of its cell and the neighbour across it (the cell
        itself at the lattice border); E, N, W, S from we, wn, we, wn; on
        extended grids the border extrapolation of ``advBound``."""
        M, N = self.M, self.N
        we, wn = np.asarray(we, dtype=np.float64), np.asarray(wn, dtype=np.float64)
        k = np.arange(M * N)
        i, j = k % M, k // M
        east = np.where(i + 1 == M, k, k + 1)
        west = np.where(i == 0, k, k - 1)
        north = np.where(j + 1 == N, k, k + M)
        south = np.where(j == 0, k, k - M)
        ww = np.stack([(we + we[east]) / 2, (wn + wn[north]) / 2, (we + we[west]) / 2, (wn + wn[south]) / 2], axis=1)
        return self._ext_map() @ ww if self.Ne else ww

    def n2e(self, idx):
        idx = np.asarray(idx)
        MN = self.M * self.N
        t = idx // MN
        r = idx % MN
        i = r % self.M
        j = r // self.M
        Me, Nee = self.M + 2 * self.Ne, self.N + 2 * self.Ne
        return (i + self.Ne) + (j + self.Ne) * Me + t * Me * Nee

    def getS(self, idxs=None):
        """Observation operator: row r picks interior cell idxs[r] (``:35-58``)."""
        rows = np.arange(self.M * self.N * self.T) if idxs is None else np.asarray(idxs)
        cols = self.n2e(rows)
        return sparse.csc_matrix((np.ones(rows.size), (np.arange(rows.size), cols)),
                                 shape=(rows.size, self.n))

    def _ext_map(self):
        """Linear extrapolation of advection faces into the border (``advBound :347-391``)."""
        if self.Ae is not None:
            return self.Ae
        M, N, Ne = self.M, self.N, self.Ne
        Me, Nee = M + 2 * Ne, N + 2 * Ne
        Ae = np.zeros((self.Ns, M * N))
        for j in range(Nee):
            for i in range(Me):
                k = i + j * Me
                inx, iny = Ne <= i < M + Ne, Ne <= j < N + Ne
                lo_x, hi_x = i < Ne, i >= M + Ne
                lo_y, hi_y = j < Ne, j >= N + Ne
                rx_hi = Me - 1 - i        # distance to the far x border
                ry_hi = Nee - 1 - j
                if lo_x and lo_y:
                    Ae[k, 0] = (j if i >= j else i) / Ne
                    continue
                if hi_x and hi_y:
                    Ae[k, M * N - 1] = (ry_hi if N - j <= M - i else rx_hi) / Ne
                    continue
                if hi_x and lo_y:
                    Ae[k, M - 1] = (rx_hi if rx_hi <= j else j) / Ne
                    continue
                if lo_x and hi_y:
                    Ae[k, M * (N - 1)] = (i if i <= ry_hi else ry_hi) / Ne
                    continue
                if lo_x:
                    Ae[k, (j - Ne) * M] = i / Ne
                elif hi_x:
                    Ae[k, M - 1 + (j - Ne) * M] = rx_hi / Ne
                if lo_y:
                    Ae[k, i - Ne] = j / Ne
                elif hi_y:
                    Ae[k, i - Ne + (N - 1) * M] = ry_hi / Ne
                if inx and iny:
                    Ae[k, (i - Ne) + (j - Ne) * M] = 1.0
        self.Ae = Ae
        return Ae

    def evalAdv(self, par):
        """Advection velocities per face (E: wx, N: wy, W: wx, S: wy) (``:267-275``)."""
        npb = self.Nbs2
        res = np.stack([self.bsA[:, 0, :] @ par[:npb], self.bsA[:, 1, :] @ par[npb:],
                        self.bsA[:, 2, :] @ par[:npb], self.bsA[:, 3, :] @ par[npb:]], axis=1)
        if self.isExtended:
            res = self._ext_map() @ res
        return res
