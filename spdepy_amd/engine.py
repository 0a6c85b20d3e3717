"""GPU evaluation engine of one space-time lattice: assembly of Q, the two
Cholesky factorisations, log-determinants, mu_c, the log-likelihood and the
per-cell gradient adjoints (``logLike``, advection_idiffusion2D.py:167-207 and
var_advection_var_diffusion2D.py:201-237).  Everything runs in the native
library on the current HIP stream; torch supplies device memory only."""
import ctypes

import numpy as np

from . import _native as nat
from .plan import Plan, _ptr, _stream

_PLANS = {}

# column layout of the per-cell adjoints written by spde_grad_cells
OUT_KAPPA, OUT_H, OUT_DGX, OUT_DGY, OUT_KAPPA0, OUT_H0, OUT_HC, OUT_WX, OUT_WY = 0, 1, 17, 21, 25, 26, 30, 34, 35
OUT_DOTGQ, OUT_DOT00, OUT_TAUTR, OUT_ADVDIR, OUT_WIDTH = 36, 37, 38, 39, 40


def get_plan(M, N, T, bc, leaf_max=128, nslots=2, wide=False):
    key = (M, N, T, bc, leaf_max, nslots, wide)
    if key not in _PLANS:
        _PLANS[key] = Plan(M, N, T, bc, leaf_max=leaf_max, nslots=nslots, wide=wide)
    return _PLANS[key]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("spdepy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _obs_device(holder, torch, n, obs_cols, y2):
    """Device copies of the observations: column indices, data, counts per
    lattice entry and S^T y (summed on the host in observation order).  Kept
    on `holder` while the same observations come back, so an evaluation only
    scales them by tau on the device."""
    c = getattr(holder, "_obs", None)
    if (c is not None and c[0].shape == obs_cols.shape and c[1].shape == y2.shape and np.array_equal(c[0], obs_cols)
            and np.array_equal(c[1], y2)):
        return c[2]
    counts = np.bincount(obs_cols, minlength=n).astype(np.float64)
    sty = np.zeros((n, y2.shape[1]))
    np.add.at(sty, obs_cols, y2)
    dev = dict(cols=torch.as_tensor(obs_cols).cuda(), y=torch.as_tensor(y2).cuda(),
               counts=torch.as_tensor(counts).cuda(), sty=torch.as_tensor(sty).cuda())
    holder._obs = (obs_cols.copy(), y2.copy(), dev)
    return dev


class Coefficients:
    """spde_coeffs plus the device tensors it points to (kept alive here)."""

    def __init__(self, grid, *, variable, kappa, H, G, sigma, kappa0, H0, a0_variable=False, mix=0,
                 kappa_c=None, H_c=None, G_c=None):
        """variable=True: per-cell kappa[Ns], H[Ns,4,2,2], G[Ns,4] (or device
        tensors); mix=1 takes the constant advection G_c instead of G, mix=2
        the constant kappa_c and H_c instead of kappa and H."""
        torch = _torch()
        c = nat.Coeffs()
        c.variable = int(variable)
        c.hx, c.hy, c.dt, c.V, c.sigma = grid.hx, grid.hy, grid.dt, grid.V, float(sigma)
        self.keep = []
        if variable:
            def dev(a):   # device tensors (fields evaluated on the GPU) are used as they are
                if a is None:
                    return None
                if isinstance(a, torch.Tensor):
                    return a.contiguous()
                return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64).reshape(-1)).cuda()
            self.kappa, self.H, self.G = dev(kappa), dev(H), dev(G)
            c.kappa = self.kappa.data_ptr() if self.kappa is not None else None
            c.H = self.H.data_ptr() if self.H is not None else None
            c.G = self.G.data_ptr() if self.G is not None else None
            c.mix = int(mix)
            if mix == 1:
                c.G_c[0], c.G_c[1] = float(G_c[0]), float(G_c[1])
            elif mix == 2:
                c.kappa_c = float(kappa_c)
                Hc = np.asarray(H_c, dtype=np.float64).reshape(4)
                for q in range(4):
                    c.H_c[q] = Hc[q]
        else:
            c.kappa_c = float(kappa)
            Hc = np.asarray(H, dtype=np.float64).reshape(4)
            for q in range(4):
                c.H_c[q] = Hc[q]
            c.G_c[0], c.G_c[1] = float(G[0]), float(G[1])
        c.kappa0 = float(kappa0)
        c.a0_variable = int(a0_variable)
        H0 = np.asarray(H0, dtype=np.float64).reshape(4)
        for q in range(4):
            c.H0[q] = H0[q]
        self.c = c


class Engine:
    """Space-time grids: Q on the M x N x T lattice.  Spatial grids
    (Whittle-Matern models, whittle_matern2D.py:59-126): Q = Q0 = A0^T A0 / V
    on the one-slice lattice, from the A0 coefficients (kappa0, H0)."""

    def __init__(self, grid, bc, leaf_max=128):
        shape = list(grid.shape)
        self.is_spatial = len(shape) == 2
        M, N, T = (shape + [1]) if self.is_spatial else shape
        self.grid, self.bc = grid, bc
        self.plan = get_plan(M, N, T, bc, leaf_max)
        self.n, self.Ns, self.nnz = grid.n, grid.Ns, self.plan.nnz
        self._spatial = None

    @property
    def spatial(self):
        """One-slice plan of the lattice (slots: 0 Q0, 1 B^T B) for the
        closed-form log|Q|."""
        if self._spatial is None:
            M, N, _ = self.plan.shape
            self._spatial = get_plan(M, N, 1, self.bc, nslots=2)
        return self._spatial

    # -------------------------------------------------------------- assembly
    def assemble(self, coeffs):
        torch = _torch()
        q = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        self.plan.assemble(coeffs.c, q)
        if self.is_spatial:  # the Q0 block of the templates is the model's Q
            self.plan.assemble_spatial(self.plan, 0, q)
        return q

    def to_scipy(self, qval):
        """The reference-pattern CSC (exact zeros dropped as scipy does)."""
        from scipy import sparse
        indptr, indices, data = self.plan.prune(qval)
        n = self.n
        # Q is symmetric: its CSR arrays are the CSC arrays of Q^T = Q
        return sparse.csc_matrix((data.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))

    def iterative(self):
        """The IC(0) / PCG context of the space-time plan (created once)."""
        if getattr(self, "_iter", None) is None:
            from .krylov import IterativeSolver
            self._iter = IterativeSolver(self.plan)
        return self._iter

    def _side_stream(self, torch):
        dev = torch.cuda.current_device()
        if getattr(self, "_side", None) is None or self._side_dev != dev:
            self._side = torch.cuda.Stream()     # default priority
            self._side_dev = dev
        return self._side

    # ------------------------------------------------------- log-likelihood
    def loglik(self, coeffs, obs_cols, data, log_tau, grad=True, trace="exact", probes=None, logdet="closed",
               contract=None, adv_dir=None):
        """Returns (like, parts) where like is the normalised -loglik/(m r) and
        parts holds the per-cell gradient adjoints and the scalar pieces.

        logdet="closed" (exact trace only): log|Q| and its gradient come from
        Q = F^T D F (F block bidiagonal in time): log|Q| = log|Q0| + (T-1)
        log|B^T B| + (T-1) sum_k log(q_k c), B = A / V, with the spatial
        matrices factorised and selectively inverted on the one-slice plan;
        only Q_c is factorised at full size.  logdet="factor" factorises Q
        as the reference does (advection_idiffusion2D.py:173).

        contract(cells_device) -> (coef_grad, column_sums): when given, the
        per-cell adjoints stay on the device and only the contraction onto
        the model's coefficients and the 40 column sums come back.
        adv_dir (Ns x 4 faces): column 39 of the cells receives the derivative
        along the scaling of the advection field (cov-advection models)."""
        torch = _torch()
        y = np.asarray(data, dtype=np.float64)
        y2 = y.reshape(y.shape[0], -1)
        m, r = y2.shape
        tau = float(np.exp(log_tau))
        n = self.n
        exact = trace == "exact"
        closed = logdet == "closed" and (exact or not grad) and not self.is_spatial
        qval = self.assemble(coeffs)
        obs = _obs_device(self, torch, n, np.asarray(obs_cols), y2)
        diag_add = obs["counts"] * tau
        # Two streams: the selected inverse of Q_c (the longest piece) runs on
        # a side stream while mu_c, the data terms and the log|Q| pieces run
        # on the current one.  A cross-stream wait is enqueued right after the
        # event it waits on is recorded.
        main = torch.cuda.current_stream()
        side = self._side_stream(torch)
        sel = grad and exact
        sp = self.spatial if closed else None
        if sel:
            sqc = torch.empty(self.plan.nnz, dtype=torch.float64, device="cuda")
            if closed:
                s_q0 = torch.empty(sp.nnz, dtype=torch.float64, device="cuda")
                s_gram = torch.empty(sp.nnz, dtype=torch.float64, device="cuda")
            else:
                sq = torch.empty(self.plan.nnz, dtype=torch.float64, device="cuda")
        b_dev = obs["sty"] * tau
        if closed:
            # Critical path: factorisation of Q_c on this stream, then its
            # selected inverse on the side stream (waiting on an event
            # recorded just before).  The solve for mu_c and the one-slice
            # log|Q| work follow here, concurrent with the selected inverse.
            self.plan.factor_async(1, qval, diag_add)
            if sel:
                factored = torch.cuda.Event()
                factored.record(main)
                side.wait_event(factored)
                with torch.cuda.stream(side):
                    self.plan.selinv(1, sqc)
            mu = self.plan.solve(1, b_dev)
            qmu = self.plan.spmv(qval, mu)
            q0 = sp.assemble_spatial(self.plan, 0)
            gram = sp.assemble_spatial(self.plan, 1)
            sp.factor_async(0, q0, owner=self)
            sp.factor_async(1, gram, owner=self)
            if sel:
                sp.selinv(0, s_q0)
                sp.selinv(1, s_gram)
        else:
            # the reference's route: Q factorised (and inverted) on the side stream
            ready = torch.cuda.Event()
            ready.record(main)
            side.wait_event(ready)
            self.plan.factor_async(1, qval, diag_add)
            factored = torch.cuda.Event()
            factored.record(main)
            with torch.cuda.stream(side):
                self.plan.factor_async(0, qval)
                if sel:
                    self.plan.selinv(0, sq)
            if sel:
                side.wait_event(factored)
                with torch.cuda.stream(side):
                    self.plan.selinv(1, sqc)
            mu = self.plan.solve(1, b_dev)
            qmu = self.plan.spmv(qval, mu)
        resid = obs["y"] - mu[obs["cols"]]
        done = torch.cuda.Event()
        done.record(side)
        main.wait_event(done)
        T = self.plan.shape[2]
        keep = []
        if grad:
            # the per-cell adjoints need only device data: they are enqueued
            # before the host reads the scalar pieces back, so the GPU works
            # on them meanwhile
            g = nat.GradArgs()
            g.r = r
            g.qval = qval.data_ptr()
            mu_cm = mu.t().contiguous()            # n x r column-major
            g.mu = mu_cm.data_ptr()
            g.diag_add = diag_add.data_ptr()
            g.coeffs = coeffs.c
            keep.append(mu_cm)
            if adv_dir is not None:
                adv = torch.as_tensor(np.ascontiguousarray(adv_dir, dtype=np.float64).reshape(-1)).cuda()
                g.adv_dir = adv.data_ptr()
                keep.append(adv)
            if exact:
                g.mode, g.nh1 = 0, 0
                g.sigma_qc = sqc.data_ptr()
                keep.append(sqc)
                if closed:
                    g.sigma_q = None
                    g.sigma_gram, g.sigma_q0, g.spatial = s_gram.data_ptr(), s_q0.data_ptr(), sp._h
                    keep += [s_gram, s_q0]
                else:
                    g.sigma_q = sq.data_ptr()
                    keep.append(sq)
            else:
                V = torch.as_tensor(np.ascontiguousarray(probes, dtype=np.float64)).cuda()
                trq = self.plan.solve(0, V).t().contiguous()
                trqc = self.plan.solve(1, V).t().contiguous()
                Vt = V.t().contiguous()
                g.mode, g.nh1 = 1, V.shape[1]
                g.trq, g.trqc, g.probes = trq.data_ptr(), trqc.data_ptr(), Vt.data_ptr()
                keep += [trq, trqc, Vt]
            out = torch.empty(self.Ns * OUT_WIDTH, dtype=torch.float64, device="cuda")
            nat.check(nat.lib().spde_grad_cells(self.plan._h, g, _ptr(out), _stream(torch)))
        ld_qc = self.plan.factor_finish(1)
        if closed:
            ld_q0 = sp.factor_finish(0)
            ld_gram = sp.factor_finish(1)
            logq = self.plan.logq_sum(1.0 / (coeffs.c.dt * coeffs.c.sigma))
            ld_q = ld_q0 + (T - 1) * (ld_gram + logq)
        else:
            ld_q = self.plan.factor_finish(0)
        quad = float((mu * qmu).sum())
        rss = float((resid * resid).sum())
        like = 0.5 * ld_q * r + m * r * np.log(tau) / 2 - 0.5 * ld_qc * r - 0.5 * quad - tau / 2 * rss
        parts = dict(m=m, r=r, tau=tau, rss=rss, ld_q=ld_q, ld_qc=ld_qc, qval=qval, mu=mu, like_raw=like)
        if not grad:
            return -like / (m * r), parts
        if contract is not None:
            parts["coef_grad"], colsum = contract(out)
        else:
            cells = out.cpu().numpy().reshape(self.Ns, OUT_WIDTH)
            parts["cells"] = cells
            colsum = cells.sum(0)
        parts["colsum"] = colsum
        dot_gq = colsum[OUT_DOTGQ]
        if exact:
            # <G, Q> without its cancellation: tr(Q_c^-1 Q) = n - tr(Q_c^-1 tau S^T S)
            # (the trace term falls from ~n to O(m) in the sum over all entries)
            dot_gq = -0.5 * r * (n - colsum[OUT_TAUTR]) - 0.5 * quad
            if not closed:
                dot_gq += 0.5 * r * n      # + 1/2 r tr(Q^-1 Q)
        parts["g_sigma"] = -(dot_gq - colsum[OUT_DOT00])
        if exact and closed:
            # 1/2 r d log|Q| / d log sigma = -1/2 r (T-1) Ns
            parts["g_sigma"] -= 0.5 * r * (T - 1) * self.plan.shape[0] * self.plan.shape[1]
        parts["g_tau"] = m * r / 2 - 0.5 * r * colsum[OUT_TAUTR] - tau / 2 * rss
        del keep
        return -like / (m * r), parts


def loglik_fused(engine, coeffs, obs_cols, data, log_tau, contract=None):
    """The exact-trace evaluation through the single C entry point
    spde_loglik_grad (closed-form log|Q|): assembly, both factorisations,
    the selected inverse, mu_c, every reduction and the per-cell adjoints
    run inside the library; returns (like, parts) like
    Engine.loglik(..., trace="exact", logdet="closed").  contract as there."""
    torch = _torch()
    y = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(len(obs_cols), -1).T)  # r x m = m x r col-major
    m = len(obs_cols)
    r = y.shape[0]
    cols = np.ascontiguousarray(obs_cols, dtype=np.int64)
    cells = torch.empty(engine.Ns * OUT_WIDTH, dtype=torch.float64, device="cuda")
    out = nat.LoglikOut()
    nat.check(nat.lib().spde_loglik_grad(engine.plan._h, engine.spatial._h, ctypes.byref(coeffs.c),
                                         cols.ctypes.data, y.ctypes.data, m, r, float(log_tau), _ptr(cells),
                                         ctypes.byref(out), _stream(torch)))
    parts = dict(m=m, r=r, tau=float(np.exp(log_tau)), ld_q=out.logdet_q, ld_qc=out.logdet_qc, like_raw=out.like_raw,
                 g_sigma=out.g_sigma, g_tau=out.g_tau)
    if contract is not None:
        parts["coef_grad"], parts["colsum"] = contract(cells)
    else:
        h = cells.cpu().numpy().reshape(engine.Ns, OUT_WIDTH)
        parts["cells"] = h
        parts["colsum"] = h.sum(0)
    return out.like, parts


class KronEngine:
    """Separable space-time models Q = Qt (x) Qs (seperable_spatial_temporal2D.py:
    64-124): the space-time plan couples neighbouring slices through the 5x5
    window (a wide-time plan); Qs lives on the one-slice plan of the same
    lattice.  Only Q_c is factorised at full size: log|Q| = Ns log|Qt| +
    T log|Qs|, and Q^-1 = Qt^-1 (x) Qs^-1 gives the log|Q| part of the
    gradient from the selected inverse of Qs."""

    def __init__(self, grid, bc, leaf_max=128):
        M, N, T = grid.shape
        self.grid, self.bc = grid, bc
        self.plan = get_plan(M, N, T, bc, leaf_max, wide=True)
        self.spatial = get_plan(M, N, 1, bc, leaf_max)
        self.n, self.Ns, self.nnz, self.T = grid.n, grid.Ns, self.plan.nnz, T

    def _qt_dev(self, qt):
        torch = _torch()
        return torch.as_tensor(np.ascontiguousarray(qt, dtype=np.float64)).cuda()

    def assemble(self, qs, qt):
        """Q values on the wide plan from Qs (one-slice values) and qt[2T-1]."""
        torch = _torch()
        q = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        qt_d = self._qt_dev(qt)
        nat.check(nat.lib().spde_assemble_kron(self.plan._h, self.spatial._h, _ptr(qs), _ptr(qt_d), _ptr(q),
                                               _stream(torch)))
        return q

    to_scipy = Engine.to_scipy
    _side_stream = Engine._side_stream

    @staticmethod
    def qt_dense(qt, T):
        return np.diag(qt[:T]) + np.diag(qt[T:], 1) + np.diag(qt[T:], -1)

    def loglik(self, qs, qt, dqts, spatial_coeffs, obs_cols, data, log_tau, grad=True, contract=None):
        """(-loglik/(m r), parts) for Q = Qt (x) Qs with the exact trace.
        dqts: the derivatives of qt (each 2T-1) of the temporal parameters;
        parts["g_temporal"] their gradient entries, parts["cells"] (or
        "coef_grad") the per-cell adjoints of the spatial operator A0 (the
        variable Whittle-Matern layout), parts["g_tau"]."""
        torch = _torch()
        y = np.asarray(data, dtype=np.float64)
        y2 = y.reshape(y.shape[0], -1)
        m, r = y2.shape
        tau = float(np.exp(log_tau))
        n, Ns, T = self.n, self.Ns, self.T
        sp = self.spatial
        qval = self.assemble(qs, qt)
        obs = _obs_device(self, torch, n, np.asarray(obs_cols), y2)
        diag_add = obs["counts"] * tau
        main = torch.cuda.current_stream()
        side = self._side_stream(torch)
        if grad:
            sig_s = torch.empty(sp.nnz, dtype=torch.float64, device="cuda")
            sig_c = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        ready = torch.cuda.Event()
        ready.record(main)
        side.wait_event(ready)
        self.plan.factor_async(1, qval, diag_add, owner=self)
        factored = torch.cuda.Event()
        factored.record(main)
        with torch.cuda.stream(side):
            sp.factor_async(0, qs, owner=self)
            if grad:
                sp.selinv(0, sig_s)
        if grad:
            side.wait_event(factored)
            with torch.cuda.stream(side):
                self.plan.selinv(1, sig_c)
        mu = self.plan.solve(1, obs["sty"] * tau)
        qmu = self.plan.spmv(qval, mu)
        resid = obs["y"] - mu[obs["cols"]]
        done = torch.cuda.Event()
        done.record(side)
        main.wait_event(done)
        ld_qc = self.plan.factor_finish(1)
        ld_s = sp.factor_finish(0)
        Qt = self.qt_dense(np.asarray(qt, dtype=np.float64), T)
        sign, ld_t = np.linalg.slogdet(Qt)
        if sign <= 0:
            raise nat.NotPositiveDefinite(nat.SPDE_ERR_NOT_SPD, "the temporal precision Qt is not positive definite")
        ld_q = Ns * ld_t + T * ld_s
        quad = float((mu * qmu).sum())
        rss = float((resid * resid).sum())
        like = 0.5 * ld_q * r + m * r * np.log(tau) / 2 - 0.5 * ld_qc * r - 0.5 * quad - tau / 2 * rss
        parts = dict(m=m, r=r, tau=tau, rss=rss, ld_q=ld_q, ld_qc=ld_qc, qval=qval, mu=mu, like_raw=like)
        if not grad:
            return -like / (m * r), parts
        mu_cm = mu.t().contiguous()
        weights = torch.empty(sp.nnz, dtype=torch.float64, device="cuda")
        blocks = torch.empty(9 * T, dtype=torch.float64, device="cuda")
        qt_d = self._qt_dev(qt)
        nat.check(nat.lib().spde_kron_reduce(self.plan._h, sp._h, _ptr(qs), _ptr(qt_d), _ptr(sig_s), 0.5 * r * T,
                                             _ptr(sig_c), _ptr(mu_cm), r, _ptr(diag_add), _ptr(weights),
                                             _ptr(blocks), _stream(torch)))
        g = nat.GradArgs()
        g.mode, g.r, g.nh1 = 0, r, 0
        g.qval = qs.data_ptr()
        g.weights = weights.data_ptr()
        g.coeffs = spatial_coeffs.c
        out = torch.empty(Ns * OUT_WIDTH, dtype=torch.float64, device="cuda")
        nat.check(nat.lib().spde_grad_cells(sp._h, g, _ptr(out), _stream(torch)))
        if contract is not None:
            parts["coef_grad"], parts["colsum"] = contract(out)
        else:
            parts["cells"] = out.cpu().numpy().reshape(Ns, OUT_WIDTH)
            parts["colsum"] = parts["cells"].sum(0)
        B = blocks.cpu().numpy().reshape(T, 3, 3)
        iQt = np.linalg.inv(Qt)
        gt = []
        for dqt in dqts:
            dqt = np.asarray(dqt, dtype=np.float64)
            dQt = self.qt_dense(dqt, T)
            coef = np.zeros((T, 3))
            coef[:, 1] = dqt[:T]
            coef[1:, 0] = dqt[T:]
            coef[:-1, 2] = dqt[T:]
            gt.append(0.5 * r * Ns * float(np.sum(iQt * dQt)) - 0.5 * r * float(np.sum(coef * B[:, :, 0]))
                      - 0.5 * float(np.sum(coef * B[:, :, 1])))
        parts["g_temporal"] = gt
        parts["g_tau"] = m * r / 2 - 0.5 * r * float(B[:, 1, 2].sum()) - tau / 2 * rss
        del mu_cm, weights, blocks, sig_s, sig_c
        return -like / (m * r), parts
