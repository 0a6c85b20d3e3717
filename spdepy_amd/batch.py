"""Batched exact evaluation of many models of uneven size (the loop over
models of the reference's kriging drivers, examples/server/get_results.py:6-51,
done as one batch): ``spde_loglik_batch`` enqueues every model's exact
log-likelihood + gradient on a pool of HIP streams before finishing any, so
small lattices share the chip instead of running one after another.  Each
lattice shape has its own plan (symbolic analysis, launch schedules, arenas),
optionally with the supernode leaf size the autotuner chose for it
(``spdepy_amd.tuner``); conditional samples with many right-hand sides come
from each model's factor of Q_c."""
import ctypes

import numpy as np

from . import _native as nat
from .engine import Engine, OUT_WIDTH
from .plan import _ptr


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("spdepy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class BatchEvaluator:
    """models: drop-in space-time models (stationary or B-spline advection-
    diffusion) whose data are set (initFit); leaf: optional per-model
    supernode leaf sizes (the tuner's choice); nstreams: HIP streams the
    batch spreads over; chunk: right-hand sides per solve of the conditional
    samples (each plan keeps front vectors for that many)."""

    def __init__(self, models, nstreams=8, leaf=None, chunk=256):
        torch = _torch()
        self.chunk = int(chunk)
        self.models = list(models)
        for i, m in enumerate(self.models):
            if leaf is not None and leaf[i] is not None:
                m._eng = Engine(m.grid, m.bc, leaf_max=int(leaf[i]))
            m._engine()
        self.streams = [torch.cuda.Stream() for _ in range(max(1, nstreams))]   # default priority
        self._obs = []
        for m in self.models:
            cols = np.ascontiguousarray(m._obs_cols(), dtype=np.int64)
            y = np.asarray(m.data, dtype=np.float64).reshape(cols.size, -1)
            self._obs.append((cols, np.ascontiguousarray(y.T)))   # m x r column-major
        rs = {o[1].shape[0] for o in self._obs}
        if len(rs) != 1:
            raise ValueError("all models of a batch need the same number of realisations")
        self.r = rs.pop()

    def loglik(self, pars, grad=True):
        """[(like, jac)] for every model (jac None when grad is False), the
        values logLike(par) returns; the parts of each model in self.parts."""
        torch = _torch()
        B = len(self.models)
        coeffs, cells = [], []
        for m, par in zip(self.models, pars):
            coeffs.append(m._coeffs(np.asarray(par, dtype=np.float64)))
            cells.append(torch.empty(m.grid.Ns * OUT_WIDTH, dtype=torch.float64, device="cuda"))
        torch.cuda.synchronize()        # the coefficient fields are ready before other streams read them
        VP = ctypes.c_void_p * B
        plans = VP(*[m._engine().plan._h.value for m in self.models])
        sps = VP(*[m._engine().spatial._h.value for m in self.models])
        cs = (nat.Coeffs * B)(*[c.c for c in coeffs])
        obs = VP(*[o[0].ctypes.data for o in self._obs])
        ys = VP(*[o[1].ctypes.data for o in self._obs])
        ms = (ctypes.c_int64 * B)(*[o[0].size for o in self._obs])
        taus = (ctypes.c_double * B)(*[float(p[-1]) for p in pars])
        cp = VP(*[c.data_ptr() for c in cells])
        outs = (nat.LoglikOut * B)()
        st = (ctypes.c_void_p * len(self.streams))(*[s.cuda_stream for s in self.streams])
        status = (ctypes.c_int32 * B)()
        nat.check(nat.lib().spde_loglik_batch(B, plans, sps, cs, obs, ys, ms, self.r, taus, cp, outs, st,
                                              len(self.streams), status))
        res, self.parts = [], []
        for i, (m, par) in enumerate(zip(self.models, pars)):
            o = outs[i]
            parts = dict(m=int(ms[i]), r=self.r, tau=float(np.exp(par[-1])), ld_q=o.logdet_q, ld_qc=o.logdet_qc,
                         like_raw=o.like_raw, g_sigma=o.g_sigma, g_tau=o.g_tau)
            if grad:
                h = cells[i].cpu().numpy().reshape(m.grid.Ns, OUT_WIDTH)
                parts["cells"] = h
                parts["colsum"] = h.sum(0)
                if hasattr(m, "_fields_dev"):
                    m._fields = m.fields(np.asarray(par, dtype=np.float64))[:4]
                g = m._contract(np.asarray(par, dtype=np.float64), coeffs[i], parts)
                res.append((o.like, -g / (parts["m"] * parts["r"])))
            else:
                res.append((o.like, None))
            self.parts.append(parts)
        return res

    def conditional_samples(self, i, nsamples, seed=0):
        """nsamples draws of x | y for model i after loglik: mu_c + P^T L^-T z
        with the factor of Q_c of the last evaluation and z from the
        library's generator (spde_gaussian with seed; n x nsamples device)."""
        torch = _torch()
        from .krylov import gaussian
        m = self.models[i]
        plan = m._engine().plan
        n = plan.n
        Z = gaussian(n, nsamples, seed)
        # right-hand sides in chunks of at most self.chunk: the solve's front
        # vectors scale with the chunk, and every plan of a batch keeps its own
        X = torch.empty_like(Z)
        for c0 in range(0, nsamples, self.chunk):
            c1 = min(nsamples, c0 + self.chunk)
            X[:, c0:c1] = plan.solve(1, Z[:, c0:c1].contiguous(), mode=nat.SOLVE_LT)
        mu = torch.empty(n * self.r, dtype=torch.float64, device="cuda")
        nat.check(nat.lib().spde_loglik_mu(plan._h, _ptr(mu), self.r, ctypes.c_void_p(
            torch.cuda.current_stream().cuda_stream)))
        return X + mu[:n, None]
