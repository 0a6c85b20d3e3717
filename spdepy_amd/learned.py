"""Learned preconditioner for the iterative path (no reference counterpart;
it serves ``logLike(par, method="pcg")``, which replaces the CHOLMOD
factorisation of Q_c, advection_idiffusion2D.py:173-189).

A 4-layer convolutional network (3x3 kernels, 32 channels, ReLU) reads seven
Jacobi-scaled stencil coefficients of Q_c per cell of each time slice and
predicts a correction of the IC(0) factor in the scaled space: 21 pair
channels (the 12 forward neighbours in the slice and the 3x3 window of the
next slice) and one diagonal channel (a log-scale of L_ii).  It is trained on
the GPU by its own BF16 forward and backward kernels (``learned.hip``) to
minimise the symmetrised divergence between p = N(0, Q_c^-1) and
q = N(0, (L' L'^T)^-1), J = KL(p||q) + KL(q||p) = 1/2 E||L'^T x||^2 +
1/2 E[y^T Q_c y] - n (x samples of the exact Cholesky path, y = L'^-T z for
Rademacher z) = sum over the spectrum of L'^-1 Q_c L'^-T of
(lambda + 1/lambda)/2 - 1: both ends of the spectrum, which bound the PCG
iterations and the variance of the stochastic log-determinant, are pushed
towards 1.  (KL(p||q) alone, whose minimiser on the pattern is the
KL-optimal sparse factor, lowers the divergence while the smallest
eigenvalues fall, and with them the convergence of PCG.)
Weights are kept in FP32 (Adam); the forward pass multiplies BF16 operands
with FP32 accumulation."""
import ctypes
import os

import numpy as np

from . import _native as nat
from .plan import _ptr, _stream

LAYERS = ((8, True), (32, True), (32, True), (32, False))   # (input channels, ReLU)
NCH = 32          # hidden and padded output channels
NPAIR = 22        # correction channels used
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("spdepy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def param_layout():
    """Offsets of (W_l, b_l) in the flat parameter vector: W_l has 9 cin x 32
    entries (row k = tap * cin + ci), b_l 32."""
    out, off = [], 0
    for cin, _ in LAYERS:
        w = 9 * cin * NCH
        out.append((off, w, off + w, NCH))
        off += w + NCH
    return out, off


def init_params(seed=0):
    """He-normal hidden layers; the output layer starts at a small scale so
    the first corrections are close to plain IC(0); output channels >= 22 are
    zero and stay fixed."""
    rng = np.random.default_rng(seed)
    lay, total = param_layout()
    th = np.zeros(total, np.float32)
    for li, ((cin, _), (wo, wn, bo, bn)) in enumerate(zip(LAYERS, lay)):
        std = np.sqrt(2.0 / (9 * cin)) * (0.01 if li == len(LAYERS) - 1 else 1.0)
        W = (std * rng.normal(size=(9 * cin, NCH))).astype(np.float32)
        if li == len(LAYERS) - 1:
            W[:, NPAIR:] = 0.0
        th[wo:wo + wn] = W.ravel()
    return th


def fixed_mask():
    """1 for trainable entries, 0 for the unused output channels."""
    lay, total = param_layout()
    mask = np.ones(total, np.uint8)
    wo, wn, bo, bn = lay[-1]
    W = mask[wo:wo + wn].reshape(-1, NCH)
    W[:, NPAIR:] = 0
    mask[bo + NPAIR:bo + bn] = 0
    return mask


class LearnedPreconditioner:
    """The network on the device: FP32 master parameters, Adam moments, packed
    BF16 operands; forward (inference), backward and training steps."""

    def __init__(self, params=None, seed=0, lr=1e-3):
        torch = _torch()
        th = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.layout, total = param_layout()
        assert th.size == total
        self.theta = torch.as_tensor(th).cuda()
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.mask = torch.as_tensor(fixed_mask()).cuda()
        self.grad = torch.zeros_like(self.theta)
        self.step_count = 0
        self.lr = lr
        self._frag_f, self._frag_b = [], []
        for cin, _ in LAYERS:
            hf, hb = ctypes.c_int64(), ctypes.c_int64()
            nat.check(nat.lib().spde_conv_packed_size(cin, 0, ctypes.byref(hf)))
            nat.check(nat.lib().spde_conv_packed_size(cin, 1, ctypes.byref(hb)))
            self._frag_f.append(torch.empty(int(hf.value), dtype=torch.int16, device="cuda"))
            self._frag_b.append(torch.empty(int(hb.value), dtype=torch.int16, device="cuda"))
        self._work = {}
        self.pack()

    # ----------------------------------------------------------- parameters
    def _views(self, t):
        return [(t[wo:wo + wn], t[bo:bo + bn]) for (wo, wn, bo, bn) in self.layout]

    def pack(self):
        torch = _torch()
        for (cin, _), (W, _b), ff, fb in zip(LAYERS, self._views(self.theta), self._frag_f, self._frag_b):
            nat.check(nat.lib().spde_conv_pack(cin, 0, _ptr(W), _ptr(ff), _stream(torch)))
            nat.check(nat.lib().spde_conv_pack(cin, 1, _ptr(W), _ptr(fb), _stream(torch)))

    def params(self):
        return self.theta.cpu().numpy()

    def save(self, path):
        np.savez_compressed(path, theta=self.params(), steps=np.array(self.step_count))

    @classmethod
    def load(cls, path, **kw):
        d = np.load(path)
        net = cls(params=d["theta"], **kw)
        net.step_count = int(d["steps"]) if "steps" in d.files else 0
        return net

    # ----------------------------------------------------------- passes
    def forward(self, feat, shape):
        """Activations [feat, Y1, Y2, Y3, Y4] (each P x channels, float32)."""
        torch = _torch()
        M, N, T = shape
        P = M * N * T
        acts = [feat]
        x = feat
        for (cin, relu), (W, b), ff in zip(LAYERS, self._views(self.theta), self._frag_f):
            y = torch.empty((P, NCH), dtype=torch.float32, device="cuda")
            nat.check(nat.lib().spde_conv_forward(M, N, T, cin, _ptr(x), None, _ptr(ff), _ptr(b), int(relu), _ptr(y),
                                                  _stream(torch)))
            acts.append(y)
            x = y
        return acts

    def backward(self, acts, dout, shape):
        """Gradient of the loss for the parameters from dL/d(output channels)
        dout (P x 22 float32); returns the flat gradient (device)."""
        torch = _torch()
        M, N, T = shape
        P = M * N * T
        G = torch.zeros((P, NCH), dtype=torch.float32, device="cuda")
        G[:, :NPAIR] = dout
        grads = self._views(self.grad)
        for li in range(len(LAYERS) - 1, -1, -1):
            cin, relu = LAYERS[li]
            X, Y = acts[li], acts[li + 1]
            mask = Y if relu else None
            key = (P, cin)
            if key not in self._work:
                wf = ctypes.c_int64()
                nat.check(nat.lib().spde_conv_wgrad_workspace(P, cin, ctypes.byref(wf)))
                self._work[key] = torch.empty(int(wf.value), dtype=torch.float32, device="cuda")
            dW, db = grads[li]
            nat.check(nat.lib().spde_conv_backward(M, N, T, cin, _ptr(X), _ptr(G), _ptr(mask), _ptr(dW), _ptr(db),
                                                   _ptr(self._work[key]), _stream(torch)))
            if li > 0:
                Gp = torch.empty((P, NCH), dtype=torch.float32, device="cuda")
                nat.check(nat.lib().spde_conv_forward(M, N, T, NCH, _ptr(G), _ptr(mask), _ptr(self._frag_b[li]), None,
                                                      0, _ptr(Gp), _stream(torch)))
                G = Gp
        return self.grad

    def adam_step(self, grad=None, lr=None, b1=0.9, b2=0.999, eps=1e-8):
        torch = _torch()
        self.step_count += 1
        g = self.grad if grad is None else grad
        nat.check(nat.lib().spde_adam(self.theta.numel(), _ptr(g), _ptr(self.mask), float(self.lr if lr is None else lr),
                                      b1, b2, eps, self.step_count, _ptr(self.theta), _ptr(self.m), _ptr(self.v),
                                      _stream(torch)))
        self.pack()

    # ----------------------------------------------------------- use
    def features(self, it):
        torch = _torch()
        feat = torch.empty((it.n, 8), dtype=torch.float32, device="cuda")
        nat.check(nat.lib().spde_ic_features(it._h, _ptr(feat), _stream(torch)))
        return feat

    def correction_for(self, it, shape):
        """Correction (n x 22 float32) for the factor currently in `it`."""
        acts = self.forward(self.features(it), shape)
        return acts[-1][:, :NPAIR].contiguous()

    def corrector(self):
        """The hook loglik_pcg takes: (engine, coeffs, qval, diag_add) ->
        correction of the factor just computed by the engine's solver."""
        def hook(engine, coeffs, qval, diag_add):
            return self.correction_for(engine.iterative(), engine.plan.shape)
        return hook

    def train_step(self, it, samples, shape, lr=None, loss="jeffreys", seed=0):
        """One Adam step at the IC(0) factor currently in `it` (just factored,
        uncorrected) with exact samples (n x K float64, row-major, lattice
        order).  loss="jeffreys": the symmetrised divergence
        KL(p||q) + KL(q||p) (p = N(0, Q_c^-1), q = N(0, (L' L'^T)^-1), with
        K Rademacher probes of seed `seed` for the second term), which bounds
        both ends of the preconditioned spectrum; loss="kl": KL(p||q) alone;
        loss="kaporin": Kaporin's condition number log(mean eig) - mean log eig
        of L'^-1 Q_c L'^-T, which bounds the CG iteration count and, like CG,
        ignores the scale of L'; "kaporin+jeffreys" adds 0.1 of the divergence.
        Returns the loss per row before the step."""
        torch = _torch()
        acts = self.forward(self.features(it), shape)
        ld_m = it.correct(acts[-1][:, :NPAIR].contiguous())
        dout = torch.empty((it.n, NPAIR), dtype=torch.float32, device="cuda")
        val = ctypes.c_double()
        K = samples.shape[1]
        if loss.startswith("kaporin"):
            # Kaporin's condition number per row, log(t / n) + log|L' L'^T| / n with
            # t = mean z^T L'^-1 A L'^-T z: the Jeffreys gradient with zero samples
            # is dt / 2, and d log L'_ii / d corr[i, 0] = 1.  Scale-invariant (as
            # CG is); "kaporin+jeffreys" adds 0.1 of the per-row divergence against
            # the exact samples.
            n = it.n
            z = it.rademacher(K, seed)
            zeros = torch.zeros_like(samples)
            nat.check(nat.lib().spde_ic_jeffreys_grad(it._h, _ptr(zeros), _ptr(z), K, 1.0, _ptr(dout),
                                                      ctypes.byref(val), _stream(torch)))
            t = 2.0 * (val.value + n)
            total = float(np.log(t / n) + ld_m / n)
            dout *= 2.0 / t
            dout[:, 0] += 2.0 / n
            if loss == "kaporin+jeffreys":
                dj = torch.empty_like(dout)
                nat.check(nat.lib().spde_ic_jeffreys_grad(it._h, _ptr(samples), _ptr(z), K, 0.1 / n, _ptr(dj),
                                                          ctypes.byref(val), _stream(torch)))
                dout += dj
                total += 0.1 * val.value / n
            self.backward(acts, dout, shape)
            self.adam_step(lr=lr)
            return total
        if loss == "jeffreys":
            z = it.rademacher(K, seed)
            nat.check(nat.lib().spde_ic_jeffreys_grad(it._h, _ptr(samples), _ptr(z), K, 1.0 / it.n, _ptr(dout),
                                                      ctypes.byref(val), _stream(torch)))
        else:
            nat.check(nat.lib().spde_ic_kl_grad(it._h, _ptr(samples), K, 1.0 / it.n, _ptr(dout), ctypes.byref(val),
                                                _stream(torch)))
        self.backward(acts, dout, shape)
        self.adam_step(lr=lr)
        return val.value / it.n

_DEFAULT = {}


def default_weights_path():
    return os.path.join(WEIGHTS_DIR, "learned_ic.npz")


def default_preconditioner():
    """The network trained by ``python -m spdepy_amd.train_learned`` (weights
    in spdepy_amd/weights/learned_ic.npz); raises if it was never trained."""
    path = default_weights_path()
    if path not in _DEFAULT:
        if not os.path.exists(path):
            raise FileNotFoundError("no trained weights at %s: run python -m spdepy_amd.train_learned" % path)
        _DEFAULT[path] = LearnedPreconditioner.load(path)
    return _DEFAULT[path]


def jeffreys_loss(it, samples, seed=12345):
    """Symmetrised KL divergence per row of the factor currently in `it`
    against N(0, Q_c^-1), from exact samples and as many Rademacher probes."""
    torch = _torch()
    dout = torch.empty((it.n, NPAIR), dtype=torch.float32, device="cuda")
    val = ctypes.c_double()
    z = it.rademacher(samples.shape[1], seed)
    nat.check(nat.lib().spde_ic_jeffreys_grad(it._h, _ptr(samples), _ptr(z), samples.shape[1], 1.0, _ptr(dout),
                                              ctypes.byref(val), _stream(torch)))
    return val.value / it.n


def kl_loss(it, samples):
    """KL loss per row of the factor currently in `it` (1/2 mean ||L^T x||^2 -
    sum log L_ii, over n) for exact samples."""
    torch = _torch()
    dout = torch.empty((it.n, NPAIR), dtype=torch.float32, device="cuda")
    loss = ctypes.c_double()
    nat.check(nat.lib().spde_ic_kl_grad(it._h, _ptr(samples), samples.shape[1], 1.0, _ptr(dout), ctypes.byref(loss),
                                        _stream(torch)))
    return loss.value / it.n


def draw_parameters(Np, rng):
    """A training or evaluation draw of the B-spline model's parameters:
    log kappa ~ N(0, 0.3^2), log gamma ~ log 0.5 + N(0, 0.3^2), vx, vy, wx,
    wy ~ N(0, 0.3^2) per coefficient, log sigma = 0, log tau = log 100."""
    return np.hstack([0.3 * rng.normal(size=Np), np.log(0.5) + 0.3 * rng.normal(size=Np),
                      0.3 * rng.normal(size=4 * Np), [0.0, np.log(100.0)]])


def exact_samples(engine, qval, diag_add, K, seed):
    """K samples of N(0, Q_c^-1) from the exact FP64 Cholesky of Q_c (slot 1
    of the engine's plan): x = P^T L^-T z, returned n x K row-major."""
    torch = _torch()
    from . import _native as nat2
    engine.plan.factor(1, qval, diag_add)
    g = torch.Generator(device="cuda").manual_seed(int(seed))
    Z = torch.randn((engine.n, K), dtype=torch.float64, device="cuda", generator=g)
    X = engine.plan.solve(1, Z, mode=nat2.SOLVE_LT)
    return X.contiguous()


def train(net, model, draws, K=16, shift=0.1, lr=1e-3, obs_frac=0.025, seed=0, log=None, loss="jeffreys"):
    """Trains `net` on `draws` parameter settings of the B-spline model
    `model` (its grid and boundary condition): for each draw, Q_c with
    observations at a random obs_frac of the lattice (tau = 100), K exact
    samples from the FP64 Cholesky path (for the divergence losses; the
    Kaporin loss uses K Rademacher probes and no samples), the IC(0) factor,
    one Adam step on `loss` (see LearnedPreconditioner.train_step).  Returns
    the per-draw losses."""
    torch = _torch()
    eng = model._engine()
    it = eng.iterative()
    shape = eng.plan.shape
    n = eng.n
    losses = []
    rng = np.random.default_rng(seed)
    for d in range(draws):
        par = draw_parameters(model.Np, rng)
        coeffs = model._coeffs(par)
        qval = eng.assemble(coeffs)
        obs = rng.choice(n, max(1, int(obs_frac * n)), replace=False)
        dadd = np.zeros(n)
        dadd[obs] = np.exp(par[-1])
        dadd_d = torch.as_tensor(dadd).cuda()
        # Kaporin's loss needs Q_c and the factor only (its K probes are
        # Rademacher); the divergences compare against exact samples
        x = torch.zeros((n, K), dtype=torch.float64, device="cuda") if loss == "kaporin" else \
            exact_samples(eng, qval, dadd_d, K, seed * 100003 + d)
        it.factor(qval, dadd_d, shift=shift)
        losses.append(net.train_step(it, x, shape, lr=lr, loss=loss, seed=seed * 100003 + d + 7))
        if log is not None and (d % 50 == 0 or d == draws - 1):
            log("draw %d loss %.6f" % (d, losses[-1]))
    return np.asarray(losses)
