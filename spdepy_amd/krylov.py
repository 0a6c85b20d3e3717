"""Mixed-precision iterative evaluation of the log-likelihood and its gradient
(the body of ``logLike``, advection_idiffusion2D.py:167-207 and
var_advection_var_diffusion2D.py:201-237) without factorising Q_c:

  * log|Q| and its gradient exactly, from the spatial matrices of
    Q = F^T D F on the one-slice plan (as the exact path);
  * mu_c = Q_c^-1 tau S^T y by FP32 PCG with an IC(0) preconditioner in an
    18-colour ordering and FP64 iterative refinement to ||r|| <= tol ||b||;
  * log|Q_c| = log|L L^T| + tr log(L^-1 Q_c L^-T), the trace by stochastic
    Lanczos quadrature over Rademacher probes (its standard error reported);
  * tr(Q_c^-1 dQ_i) by the Hutchinson estimate of Q_c^-1 on Q's pattern from
    PCG solves with Rademacher probes (the reference's estimator, :182-189),
    contracted through the same adjoint kernels as the exact path.

The preconditioner may carry a learned correction (``spdepy_amd.learned``).
All arithmetic runs in the native library; torch provides device buffers."""
import ctypes
import threading

import numpy as np

from . import _native as nat
from .plan import _ptr, _stream


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("spdepy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def cg_lanczos(alpha, beta, iters):
    """The Lanczos tridiagonals (diagonal, off-diagonal; k x cap, as
    slq_quadrature takes them) of L^-1 A L^-T from each column's PCG step
    coefficients alpha_j, beta_j (Saad, Iterative Methods for Sparse Linear
    Systems, 2nd ed., sec. 6.7.3): d_1 = 1/alpha_1, d_j = 1/alpha_j +
    beta_(j-1)/alpha_(j-1), e_j = sqrt(beta_j)/alpha_j; a column stops at
    its last iteration (its off-diagonal there is zero)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    beta = np.asarray(beta, dtype=np.float64)
    k, cap = alpha.shape
    d = np.zeros((k, cap))
    e = np.zeros((k, cap))
    for r in range(k):
        m = int(min(iters[r], cap))
        a, b = alpha[r, :m], beta[r, :m]
        if m == 0:
            continue
        d[r, :m] = 1.0 / a
        d[r, 1:m] += b[:m - 1] / a[:m - 1]
        e[r, :m - 1] = np.sqrt(np.maximum(b[:m - 1], 0.0)) / a[:m - 1]
    return d, e


def slq_quadrature(alpha, beta, n):
    """Gauss quadrature of log over each probe's Lanczos tridiagonal:
    n * sum_i (e_1^T y_i)^2 log(theta_i); returns (mean, standard error,
    per-probe values).  A probe's tridiagonal stops at its first zero beta."""
    vals = []
    for a, b in zip(alpha, beta):
        m = len(a)
        for j in range(m - 1):
            if not b[j] > 0:
                m = j + 1
                break
        T = np.diag(a[:m]) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
        th, Y = np.linalg.eigh(T)
        vals.append(n * float(np.sum(Y[0] ** 2 * np.log(np.maximum(th, 1e-300)))))
    vals = np.asarray(vals)
    se = float(vals.std(ddof=1) / np.sqrt(len(vals))) if len(vals) > 1 else float("inf")
    return float(vals.mean()), se, vals


class IterativeSolver:
    """The native IC(0) / PCG context of one space-time plan."""

    def __init__(self, plan):
        self.plan = plan
        h = ctypes.c_void_p()
        nat.check(nat.lib().spde_ic_create(plan._h, ctypes.byref(h)))
        self._h = h
        nl, nc = ctypes.c_int64(), ctypes.c_int32()
        nat.check(nat.lib().spde_ic_info(h, ctypes.byref(nl), ctypes.byref(nc)))
        self.nlow, self.ncolours = int(nl.value), int(nc.value)
        self.n = plan.n
        self.shift = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is None or not h.value:
            return
        self._h = None
        try:
            nat.lib().spde_ic_destroy(h)
        except (AttributeError, TypeError):  # at interpreter exit the modules are gone; the process frees the device
            pass

    def view(self):
        """A second handle on this factor with its own work buffers, for a
        solve on another stream at the same time (spde_ic_view); it reads the
        factor of this solver's latest factorisation."""
        v = IterativeSolver.__new__(IterativeSolver)
        h = ctypes.c_void_p()
        nat.check(nat.lib().spde_ic_view(self._h, ctypes.byref(h)))
        v._h, v._parent = h, self
        v.plan, v.nlow, v.ncolours, v.n, v.shift = self.plan, self.nlow, self.ncolours, self.n, self.shift
        return v

    def export(self):
        n, nl = self.n, self.nlow
        out = dict(colour=np.zeros(n, np.int32), lptr=np.zeros(n + 1, np.int64), lcol=np.zeros(max(nl, 1), np.int32),
                   lch=np.zeros(max(nl, 1), np.int64), uptr=np.zeros(n + 1, np.int64),
                   ucol=np.zeros(max(nl, 1), np.int32))
        nat.check(nat.lib().spde_ic_export(self._h, *[out[k].ctypes.data for k in
                                                      ("colour", "lptr", "lcol", "lch", "uptr", "ucol")]))
        return out

    def factor(self, qval, diag_add=None, shift=0.1, max_shift=10.0):
        """IC(0) of Q_c with the diagonal scaled by (1 + shift); on a
        non-positive pivot the shift is tripled (up to max_shift).  Returns
        log|L L^T|; the shift used is kept in self.shift."""
        torch = _torch()
        ld = ctypes.c_double()
        while True:
            try:
                nat.check(nat.lib().spde_ic_factor(self._h, _ptr(qval), _ptr(diag_add), float(shift),
                                                   ctypes.byref(ld), _stream(torch)))
                break
            except nat.NotPositiveDefinite:
                if shift * 3 > max_shift:
                    raise
                shift *= 3
        self.shift = shift
        self._keep = (qval, diag_add)
        return ld.value

    def correct(self, corr):
        """Applies a learned correction (device float tensor, n x 22) to the
        current factor; returns the new log|L L^T|."""
        torch = _torch()
        ld = ctypes.c_double()
        nat.check(nat.lib().spde_ic_correct(self._h, _ptr(corr), ctypes.byref(ld), _stream(torch)))
        return ld.value

    def values(self):
        torch = _torch()
        lo = torch.empty(max(self.nlow, 1), dtype=torch.float64, device="cuda")
        dg = torch.empty(self.n, dtype=torch.float64, device="cuda")
        ad = torch.empty(self.n, dtype=torch.float64, device="cuda")
        nat.check(nat.lib().spde_ic_values(self._h, _ptr(lo), _ptr(dg), _ptr(ad), _stream(torch)))
        return lo[:self.nlow], dg, ad

    @staticmethod
    def _k(X):
        k = 1 if X.dim() == 1 else X.shape[1]
        if k & (k - 1) or k > 64:
            raise ValueError("right-hand side count must be a power of two <= 64")
        return k

    def apply(self, R):
        torch = _torch()
        Z = torch.empty_like(R)
        nat.check(nat.lib().spde_ic_apply(self._h, _ptr(R), _ptr(Z), self._k(R), _stream(torch)))
        return Z

    def sweep(self, which, X):
        """which = 1: L^-1 X; which = 2: L^-T X."""
        torch = _torch()
        Y = torch.empty_like(X)
        nat.check(nat.lib().spde_ic_sweep(self._h, which, _ptr(X), _ptr(Y), self._k(X), _stream(torch)))
        return Y

    def spmm(self, X):
        torch = _torch()
        Y = torch.empty_like(X)
        nat.check(nat.lib().spde_ic_spmm(self._h, _ptr(X), _ptr(Y), self._k(X), _stream(torch)))
        return Y

    def solve(self, B, tol=1e-6, maxit=5000):
        """FP32 PCG: X (n x k float32, row-major) and the iterations per column."""
        torch = _torch()
        B = B.contiguous()
        k = self._k(B)
        X = torch.empty_like(B)
        it = np.zeros(k, np.int32)
        nat.check(nat.lib().spde_pcg_solve(self._h, _ptr(B), _ptr(X), k, float(tol), int(maxit), it.ctypes.data,
                                           _stream(torch)))
        return X, it

    def solve_coeffs(self, B, tol=1e-6, maxit=5000, cap=50):
        """solve() with each column's first `cap` PCG step coefficients
        (k x cap alpha and beta, zero past its iterations)."""
        torch = _torch()
        B = B.contiguous()
        k = self._k(B)
        X = torch.empty_like(B)
        it = np.zeros(k, np.int32)
        a = np.zeros((k, cap))
        b = np.zeros((k, cap))
        nat.check(nat.lib().spde_pcg_solve_coeffs(self._h, _ptr(B), _ptr(X), k, float(tol), int(maxit),
                                                  it.ctypes.data, a.ctypes.data, b.ctypes.data, int(cap),
                                                  _stream(torch)))
        return X, it, a, b

    def lmul(self, W):
        """L W for the current FP32 factor (n x k float32)."""
        torch = _torch()
        W = W.contiguous()
        Y = torch.empty_like(W)
        nat.check(nat.lib().spde_ic_lmul(self._h, _ptr(W), _ptr(Y), self._k(W), _stream(torch)))
        return Y

    def refine(self, b, tol=1e-10, inner_tol=1e-5, max_rounds=10, maxit=5000):
        """FP64 solution of Q_c x = b by iterative refinement around FP32 PCG."""
        torch = _torch()
        b = b.contiguous()
        x = torch.empty_like(b)
        it, rounds, rel = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        nat.check(nat.lib().spde_pcg_refine(self._h, _ptr(b), _ptr(x), float(tol), float(inner_tol),
                                            int(max_rounds), int(maxit), ctypes.byref(it), ctypes.byref(rounds),
                                            ctypes.byref(rel), _stream(torch)))
        return x, dict(iters=int(it.value), rounds=int(rounds.value), relres=float(rel.value))

    def replace(self, b, tol=1e-10, delta=0.1, maxit=10000):
        """FP64 solution of Q_c x = b by FP32 PCG with FP64 residual
        replacement whenever the recursive residual has fallen by delta."""
        torch = _torch()
        b = b.contiguous()
        x = torch.empty_like(b)
        it, upd, rel = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        nat.check(nat.lib().spde_pcg_replace(self._h, _ptr(b), _ptr(x), float(tol), float(delta), int(maxit),
                                             ctypes.byref(it), ctypes.byref(upd), ctypes.byref(rel),
                                             _stream(torch)))
        return x, dict(iters=int(it.value), updates=int(upd.value), relres=float(rel.value))

    def replace_block(self, B, tol=1e-10, delta=0.1, maxit=10000):
        """FP64 X = Q_c^-1 B for k right-hand sides (as refine_block) by the
        FP32 block PCG with FP64 residual replacement; info holds per-column
        iterations, ||b - Q_c x|| / ||b|| and the number of replacements."""
        torch = _torch()
        B = B.contiguous()
        k = self._k(B)
        X = torch.empty_like(B)
        its = np.zeros(k, np.int32)
        rel = np.zeros(k)
        upd = ctypes.c_int32()
        nat.check(nat.lib().spde_pcg_replace_block(self._h, _ptr(B), _ptr(X), k, float(tol), float(delta), int(maxit),
                                                   its.ctypes.data, ctypes.byref(upd), rel.ctypes.data,
                                                   _stream(torch)))
        return X, dict(iters=its, updates=int(upd.value), relres=rel)

    def refine_block(self, B, tol=1e-10, inner_tol=1e-5, max_rounds=10, maxit=5000):
        """FP64 X = Q_c^-1 B for k right-hand sides (B: n x k float64 device,
        row-major; k a power of two <= 64) by iterative refinement around the
        FP32 block PCG; info holds per-column iterations and ||b - Q_c x|| / ||b||."""
        torch = _torch()
        B = B.contiguous()
        k = self._k(B)
        X = torch.empty_like(B)
        its = np.zeros(k, np.int32)
        rel = np.zeros(k)
        rounds = np.zeros(1, np.int32)
        nat.check(nat.lib().spde_pcg_refine_block(self._h, _ptr(B), _ptr(X), k, float(tol), float(inner_tol),
                                                  int(max_rounds), int(maxit), its.ctypes.data, rounds.ctypes.data,
                                                  rel.ctypes.data, _stream(torch)))
        return X, dict(iters=its, rounds=int(rounds[0]), relres=rel)

    def rademacher(self, k, seed, scale=1.0):
        torch = _torch()
        V = torch.empty((self.n, k), dtype=torch.float32, device="cuda")
        nat.check(nat.lib().spde_rademacher(self.n, k, ctypes.c_uint64(seed), ctypes.c_float(scale), _ptr(V),
                                            _stream(torch)))
        return V

    def slq(self, k, steps, seed):
        torch = _torch()
        a = np.zeros((k, steps))
        b = np.zeros((k, steps))
        nat.check(nat.lib().spde_slq(self._h, k, steps, ctypes.c_uint64(seed), a.ctypes.data, b.ctypes.data,
                                     _stream(torch)))
        return a, b

    def hutch_pattern(self, U, V, out=None, alpha=1.0, beta=0.0):
        """out = beta out + alpha (estimate of Q_c^-1 on the pattern from U = Q_c^-1 V)."""
        torch = _torch()
        sig = torch.empty(self.plan.nnz, dtype=torch.float64, device="cuda") if out is None else out
        nat.check(nat.lib().spde_hutch_pattern(self._h, _ptr(U), _ptr(V), self._k(U), float(alpha), float(beta),
                                               _ptr(sig), _stream(torch)))
        return sig


def gaussian(n, k, seed):
    """n x k standard normal draws (float64 device, row-major) from the
    library's counter-based generator (spde_gaussian)."""
    torch = _torch()
    Z = torch.empty((n, k), dtype=torch.float64, device="cuda")
    nat.check(nat.lib().spde_gaussian(int(n), int(k), ctypes.c_uint64(seed), _ptr(Z), _stream(torch)))
    return Z


def conditional_rhs(engine, coeffs, obs_cols, y, log_tau, k, seed):
    """B (n x k float64 device, row-major) = F^T D^(1/2) z1 + S^T (tau y +
    sqrt(tau) z2) with z1 ~ N(0, I_n) (spde_gaussian with seed) and z2 ~
    N(0, I_m) (numpy default_rng(seed)): E[B] = tau S^T y = Q_c mu_c and
    Cov[B] = Q + tau S^T S = Q_c, so Q_c^-1 B holds k draws of x | y.  The
    operators are those of the last assembly of coeffs on the engine's plan."""
    torch = _torch()
    n = engine.n
    cols = np.asarray(obs_cols, dtype=np.int64)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    if y.size != cols.size:
        raise ValueError("conditional samples take one realisation (y of length m)")
    if cols.size and (cols.min() < 0 or cols.max() >= n):
        raise ValueError("observation index outside the lattice")
    tau = float(np.exp(log_tau))
    Z = gaussian(n, k, seed)
    B = torch.empty_like(Z)
    nat.check(nat.lib().spde_fform_apply_t(engine.plan._h, ctypes.byref(coeffs.c), _ptr(Z), _ptr(B), int(k),
                                           _stream(torch)))
    z2 = np.random.default_rng(seed).normal(size=(cols.size, k))
    rows, inv = np.unique(cols, return_inverse=True)
    w = np.zeros((rows.size, k))
    np.add.at(w, inv, tau * y[:, None] + np.sqrt(tau) * z2)
    rows_d = torch.as_tensor(rows).cuda()
    w_d = torch.as_tensor(np.ascontiguousarray(w)).cuda()
    nat.check(nat.lib().spde_add_rows(int(n), int(k), int(rows.size), _ptr(rows_d), _ptr(w_d), _ptr(B),
                                      _stream(torch)))
    torch.cuda.current_stream().synchronize()
    return B


def conditional_samples_pcg(engine, coeffs, obs_cols, y, log_tau, k=64, seed=0, tol=1e-10, shift=0.1,
                            correction=None, inner_tol=1e-5, solver="replace", delta=0.1):
    """k draws of x | y (n x k float64 device, row-major) without a factor:
    Q_c^-1 B for the right-hand sides of conditional_rhs, solved by the FP32
    block PCG with IC(0) (or its learned correction) and FP64 residual
    replacement (solver="replace", every `delta` of residual fall) or FP64
    refinement rounds of restarted solves to `inner_tol` (solver="refine"),
    to ||B - Q_c X|| <= tol ||B|| per column; returns (X, info)."""
    torch = _torch()
    from .engine import _obs_device
    it = engine.iterative()
    qval = engine.assemble(coeffs)
    yy = np.asarray(y, dtype=np.float64).reshape(-1)
    obs = _obs_device(engine, torch, engine.n, np.asarray(obs_cols), yy.reshape(-1, 1))
    diag_add = obs["counts"] * float(np.exp(log_tau))
    it.factor(qval, diag_add, shift=shift)
    if correction is not None:
        it.correct(correction(engine, coeffs, qval, diag_add))
    B = conditional_rhs(engine, coeffs, obs_cols, yy, log_tau, k, seed)
    if solver == "replace":
        X, info = it.replace_block(B, tol=tol, delta=delta)
    elif solver == "refine":
        X, info = it.refine_block(B, tol=tol, inner_tol=inner_tol)
    else:
        raise ValueError("solver must be 'replace' or 'refine'")
    info["B"] = B
    return X, info


def _side_solver(it):
    """The view of `it` and the stream that the conditional mean's solves
    use beside the probes (made once per solver)."""
    torch = _torch()
    if getattr(it, "_side", None) is None:
        it._side = (it.view(), torch.cuda.Stream())
    return it._side


def _start_side(it, job):
    """Runs job(view of it) on the side stream from a second host thread,
    after the work enqueued so far on the current stream; returns a function
    that joins the thread, makes the current stream wait for the side stream
    and returns the job's result (tensors marked for the current stream), or
    raises the job's exception."""
    torch = _torch()
    view, stream = _side_solver(it)
    main = torch.cuda.current_stream()
    stream.wait_stream(main)
    box = {}

    def run():
        try:
            with torch.cuda.stream(stream):
                box["res"] = job(view)
        except BaseException as e:
            box["err"] = e
    worker = threading.Thread(target=run, name="spdepy-side-solve")
    worker.start()

    def join():
        worker.join()
        if "err" in box:
            raise box["err"]
        main.wait_stream(stream)
        for x, _ in box["res"]:
            x.record_stream(main)
        return box["res"]
    return join


def _hutchinson_cg(it, probes, seed, tol, steps):
    """The estimate of Q_c^-1 on the pattern and the Lanczos tridiagonals
    of L^-1 Q_c L^-T (first `steps` steps) from the same PCG solves: probes
    v = L w for Rademacher w in batches of up to 64 (batch j from seed + 1 +
    its first index), u = Q_c^-1 v and y = L^-T w, so E[u y^T] = Q_c^-1;
    returns (sigma, iterations per probe, diagonals, off-diagonals)."""
    sqc, iters, D, E, done = None, [], [], [], 0
    while done < probes:
        k = min(64, probes - done)
        k = 1 << (k.bit_length() - 1)
        W = it.rademacher(k, seed + 1 + done)
        U, its, a, b = it.solve_coeffs(it.lmul(W), tol=tol, cap=steps)
        Y = it.sweep(2, W)
        iters += list(its)
        sqc = it.hutch_pattern(U, Y, out=sqc, alpha=k / probes, beta=0.0 if sqc is None else 1.0)
        d, e = cg_lanczos(a, b, its)
        D.append(d)
        E.append(e)
        done += k
    return sqc, iters, np.vstack(D), np.vstack(E)


def _hutchinson(it, probes, seed, tol):
    """Estimate of Q_c^-1 on the pattern from `probes` Rademacher probes, in
    batches of up to 64 (the probes of batch j from seed + 1 + its first
    index); returns (sigma, iterations per probe)."""
    sqc, iters, done = None, [], 0
    while done < probes:
        k = min(64, probes - done)
        k = 1 << (k.bit_length() - 1)
        V = it.rademacher(k, seed + 1 + done)
        U, its = it.solve(V, tol=tol)
        iters += list(its)
        sqc = it.hutch_pattern(U, V, out=sqc, alpha=k / probes, beta=0.0 if sqc is None else 1.0)
        done += k
    return sqc, iters


def loglik_pcg(engine, coeffs, obs_cols, data, log_tau, grad=True, slq_probes=32, slq_steps=50, hutch_probes=64,
               seed=0, tol=1e-10, probe_tol=1e-3, shift=0.1, correction=None, contract=None, adv_dir=None,
               tol_like=None, inner_tol=1e-5, mean_solver="replace", delta=0.1, slq="cg", pilot_steps=20):
    """(-loglik/(m r), parts) as ``Engine.loglik`` returns them, with log|Q_c|
    from stochastic Lanczos quadrature and the trace part of the gradient from
    the Hutchinson estimate.  correction(engine, coeffs, qval, diag_add) ->
    device float tensor (n x 22) applies a learned preconditioner correction.
    parts gains "logdet_qc_se" (standard error of the SLQ estimate),
    "pcg" (iteration counts) and "ic_shift".

    mean_solver: "replace" (default) solves for mu_c by FP32 PCG with FP64
    residual replacement each time the residual has fallen by `delta`
    (1,088 iterations at 100 x 100 x 20); "refine" by restarted FP32 solves
    to `inner_tol` inside FP64 refinement rounds (1,383).  Both to `tol`.

    slq: "cg" (default when the gradient is wanted and no tolerance waits
    on the value) takes the Lanczos quadrature from the Hutchinson probes'
    own PCG solves -- probes v = L w for Rademacher w, whose PCG coefficients
    give the Lanczos tridiagonal of L^-1 Q_c L^-T started at w, and whose
    solutions u = Q_c^-1 L w with y = L^-T w estimate Q_c^-1 (E[u y^T] =
    Q_c^-1) -- so log|Q_c| comes from hutch_probes probes x slq_steps steps
    at no extra cost (the batched-CG estimator of Gardner et al., NeurIPS
    2018); "lanczos" runs slq_probes separate Lanczos probes first.

    probe_tol: relative residual of the FP32 solves Q_c U = V of the
    Hutchinson probes.  At 100 x 100 x 20 with 64 probes, 1e-3 instead of
    1e-6 moves the gradient by 2.4e-3 of its norm while the estimator's own
    error (probe variance) is 0.75 of it, for half the iterations.

    tol_like: the relative standard error the value must meet.  The SLQ
    estimate runs first; when 1/2 r se exceeds tol_like times an upper bound
    of |loglik| (mu_c minimises mu^T Q mu + tau ||y - S mu||^2, so
    0 <= mu^T Q mu + tau ||y - S mu||^2 <= tau ||y||^2), the tolerance cannot
    be met and the call returns (None, parts) with parts["infeasible"] set,
    before the solves for mu_c and the Hutchinson probes.  A pilot of 4
    probes of pilot_steps Lanczos steps decides first when even a tenth of
    its error, scaled to slq_probes, is out of reach (parts["pilot"]); the
    spread of the probes' quadratures, which sets that error, is fixed
    after the first steps."""
    torch = _torch()
    from .engine import OUT_DOT00, OUT_TAUTR, OUT_WIDTH, _obs_device
    if engine.is_spatial:
        raise ValueError("the iterative path is for space-time lattices")
    y = np.asarray(data, dtype=np.float64)
    y2 = y.reshape(y.shape[0], -1)
    m, r = y2.shape
    tau = float(np.exp(log_tau))
    n, T = engine.n, engine.plan.shape[2]
    it = engine.iterative()
    qval = engine.assemble(coeffs)
    obs = _obs_device(engine, torch, n, np.asarray(obs_cols), y2)
    diag_add = obs["counts"] * tau
    sp = engine.spatial
    q0 = sp.assemble_spatial(engine.plan, 0)
    gram = sp.assemble_spatial(engine.plan, 1)
    sp.factor_async(0, q0, owner=engine)
    sp.factor_async(1, gram, owner=engine)
    s_q0 = s_gram = None
    ld_m = it.factor(qval, diag_add, shift=shift)
    if correction is not None:
        ld_m = it.correct(correction(engine, coeffs, qval, diag_add))
    b_dev = obs["sty"] * tau

    if mean_solver not in ("replace", "refine"):
        raise ValueError("mean_solver must be 'replace' or 'refine'")

    def mean_solves(solver):
        if mean_solver == "replace":
            return [solver.replace(b_dev[:, c].contiguous(), tol=tol, delta=delta) for c in range(r)]
        return [solver.refine(b_dev[:, c].contiguous(), tol=tol, inner_tol=inner_tol) for c in range(r)]

    # mu_c (one right-hand side: short, latency-bound launches) runs on a
    # second stream from a second host thread beside the SLQ and the
    # Hutchinson probes' 64-column solves on this one, each handle with its
    # own work buffers; with a tolerance on the value it starts only once the
    # SLQ has shown the tolerance within reach
    side, results = None, None
    try:
        if grad and hutch_probes > 0 and tol_like is None:
            side = _start_side(it, mean_solves)
        ld_q0 = sp.factor_finish(0)
        ld_gram = sp.factor_finish(1)
        logq = engine.plan.logq_sum(1.0 / (coeffs.c.dt * coeffs.c.sigma))
        ld_q = ld_q0 + (T - 1) * (ld_gram + logq)

        def out_of_reach(ld_qc, se, margin=1.0):
            """(bound, whether margin * (1/2) r se exceeds tol_like * bound)."""
            known = 0.5 * ld_q * r + m * r * np.log(tau) / 2 - 0.5 * ld_qc * r
            bound = abs(known) + 0.5 * tau * float((y2 * y2).sum())
            return bound, margin * 0.5 * r * se > tol_like * bound

        def infeasible(ld_qc, se, bound, pilot):
            return None, dict(m=m, r=r, tau=tau, ld_q=ld_q, ld_qc=ld_qc, logdet_qc_se=se, logdet_m=ld_m,
                              ic_shift=it.shift, infeasible=True, like_bound=bound, pilot=pilot,
                              pcg=dict(mu=[]))
        if tol_like is not None and slq_probes > 4:
            # a pilot of 4 probes: when a tenth of its standard error, scaled to
            # slq_probes probes, already exceeds the tolerance, the full
            # estimate cannot meet it (the tenth covers the pilot's own spread)
            a_p, b_p = it.slq(4, min(pilot_steps, slq_steps), seed + 7919)
            tr_p, se_p, _ = slq_quadrature(a_p, b_p, n)
            se_est = se_p * np.sqrt(4.0 / slq_probes)
            bound, out = out_of_reach(ld_m + tr_p, se_est, margin=0.1)
            if out:
                return infeasible(ld_m + tr_p, se_est, bound, True)
        cg_slq = slq == "cg" and side is not None
        if not cg_slq:
            alpha, beta = it.slq(slq_probes, slq_steps, seed)
            trlog, se, _ = slq_quadrature(alpha, beta, n)
            ld_qc = ld_m + trlog
        if tol_like is not None:
            bound, out = out_of_reach(ld_qc, se)
            if out:
                return infeasible(ld_qc, se, bound, False)
            if grad and hutch_probes > 0:
                side = _start_side(it, mean_solves)
        if grad:
            # the spatial selected inverses of the log|Q| part of the gradient,
            # enqueued once the value is known to be wanted
            s_q0 = torch.empty(sp.nnz, dtype=torch.float64, device="cuda")
            s_gram = torch.empty(sp.nnz, dtype=torch.float64, device="cuda")
            sp.selinv(0, s_q0)
            sp.selinv(1, s_gram)
        if cg_slq:
            sqc, iters, alpha, beta = _hutchinson_cg(it, hutch_probes, seed, probe_tol, slq_steps)
        elif side is not None:
            sqc, iters = _hutchinson(it, hutch_probes, seed, probe_tol)
        else:
            results = mean_solves(it)
    finally:
        if side is not None:       # joined on every exit, so the view's buffers are free again
            results = side()
    if cg_slq:
        trlog, se, _ = slq_quadrature(alpha, beta, n)
        ld_qc = ld_m + trlog
    mus = [x for x, _ in results]
    info = [inf for _, inf in results]
    mu = torch.stack(mus, 1)
    qmu = engine.plan.spmv(qval, mu)
    resid = obs["y"] - mu[obs["cols"]]
    quad = float((mu * qmu).sum())
    rss = float((resid * resid).sum())
    like = 0.5 * ld_q * r + m * r * np.log(tau) / 2 - 0.5 * ld_qc * r - 0.5 * quad - tau / 2 * rss
    parts = dict(m=m, r=r, tau=tau, rss=rss, ld_q=ld_q, ld_qc=ld_qc, qval=qval, mu=mu, like_raw=like,
                 logdet_qc_se=se, logdet_m=ld_m, ic_shift=it.shift, pcg=dict(mu=info))
    if not grad:
        return -like / (m * r), parts
    if side is None:
        sqc, iters = _hutchinson(it, hutch_probes, seed, probe_tol)
    parts["pcg"]["probes"] = np.asarray(iters)
    g = nat.GradArgs()
    g.mode, g.nh1, g.r = 0, 0, r
    g.qval = qval.data_ptr()
    mu_cm = mu.t().contiguous()
    g.mu = mu_cm.data_ptr()
    g.diag_add = diag_add.data_ptr()
    g.coeffs = coeffs.c
    g.sigma_qc = sqc.data_ptr()
    g.sigma_q = None
    g.sigma_gram, g.sigma_q0, g.spatial = s_gram.data_ptr(), s_q0.data_ptr(), sp._h
    keep = [mu_cm, sqc, s_gram, s_q0]
    if adv_dir is not None:
        adv = torch.as_tensor(np.ascontiguousarray(adv_dir, dtype=np.float64).reshape(-1)).cuda()
        g.adv_dir = adv.data_ptr()
        keep.append(adv)
    out = torch.empty(engine.Ns * OUT_WIDTH, dtype=torch.float64, device="cuda")
    nat.check(nat.lib().spde_grad_cells(engine.plan._h, g, _ptr(out), _stream(torch)))
    if contract is not None:
        parts["coef_grad"], colsum = contract(out)
    else:
        cells = out.cpu().numpy().reshape(engine.Ns, OUT_WIDTH)
        parts["cells"] = cells
        colsum = cells.sum(0)
    parts["colsum"] = colsum
    # <G, Q> through tr(Q_c^-1 Q) = n - tr(Q_c^-1 tau S^T S): the estimate then
    # needs only the observed diagonal of Q_c^-1
    dot_gq = -0.5 * r * (n - colsum[OUT_TAUTR]) - 0.5 * quad
    parts["g_sigma"] = -(dot_gq - colsum[OUT_DOT00]) - 0.5 * r * (T - 1) * engine.plan.shape[0] * \
        engine.plan.shape[1]
    parts["g_tau"] = m * r / 2 - 0.5 * r * colsum[OUT_TAUTR] - tau / 2 * rss
    del keep
    return -like / (m * r), parts
