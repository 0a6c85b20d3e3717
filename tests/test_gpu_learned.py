"""GPU checks of the learned preconditioner (learned.hip, the feature and KL
kernels of pcg.hip) against the oracle (oracle/learned.py) on the same
inputs, and what a trained network gains over plain IC(0)."""
import numpy as np
import pytest
from scipy import sparse

from oracle import fixtures as F, krylov as K, learned as Le
from tests.gpu_util import product_case, step

pytestmark = pytest.mark.gpu


def test_bf16_matrix_instruction_lane_layout():
    import torch
    from spdepy_amd import _native as nat
    from spdepy_amd.plan import _ptr, _stream
    out = torch.zeros(256, dtype=torch.float32, device="cuda")
    step("bf16 MFMA probe")
    nat.check(nat.lib().spde_debug_bf16_probe(_ptr(out), _stream(torch)))
    got = out.cpu().numpy().reshape(64, 4)
    lane = np.arange(64)[:, None]
    j = np.arange(4)[None, :]
    # D = A: lane holds rows 4 (lane / 16) + j of column lane % 16
    assert np.array_equal(got, (16 * (4 * (lane // 16) + j) + lane % 16).astype(np.float32))


def _conv_case(shape, cin, seed):
    import torch
    rng = np.random.default_rng(seed)
    P = int(np.prod(shape))
    X = rng.normal(size=(P, cin)).astype(np.float32)
    W = rng.normal(scale=np.sqrt(2 / (9 * cin)), size=(9 * cin, 32)).astype(np.float32)
    b = rng.normal(scale=0.1, size=32).astype(np.float32)
    G = rng.normal(size=(P, 32)).astype(np.float32)
    d = lambda a: torch.as_tensor(a).cuda()
    return X, W, b, G, d(X), d(W), d(b), d(G)


@pytest.mark.parametrize("cin,shape", [(8, (13, 7, 3)), (32, (9, 10, 2))])
def test_conv_forward_backward_match_oracle(cin, shape):
    import ctypes
    import torch
    from spdepy_amd import _native as nat
    from spdepy_amd.plan import _ptr, _stream
    L = nat.lib()
    st = _stream(torch)
    M, N, T = shape
    P = M * N * T
    X, W, b, G, Xd, Wd, bd, Gd = _conv_case(shape, cin, 11 + cin)
    hf, hb = ctypes.c_int64(), ctypes.c_int64()
    nat.check(L.spde_conv_packed_size(cin, 0, ctypes.byref(hf)))
    nat.check(L.spde_conv_packed_size(cin, 1, ctypes.byref(hb)))
    ff = torch.empty(hf.value, dtype=torch.int16, device="cuda")
    fb = torch.empty(hb.value, dtype=torch.int16, device="cuda")
    nat.check(L.spde_conv_pack(cin, 0, _ptr(Wd), _ptr(ff), st))
    nat.check(L.spde_conv_pack(cin, 1, _ptr(Wd), _ptr(fb), st))
    step("conv forward cin=%d %s" % (cin, shape))
    Y = torch.empty((P, 32), dtype=torch.float32, device="cuda")
    nat.check(L.spde_conv_forward(M, N, T, cin, _ptr(Xd), None, _ptr(ff), _ptr(bd), 1, _ptr(Y), st))
    Yo = Le.conv(X, W, b, shape, True)
    Yg = Y.cpu().numpy()
    assert np.allclose(Yg, Yo, rtol=1e-5, atol=1e-5 * np.abs(Yo).max())
    step("conv backward cin=%d" % cin)
    wf = ctypes.c_int64()
    nat.check(L.spde_conv_wgrad_workspace(P, cin, ctypes.byref(wf)))
    work = torch.empty(wf.value, dtype=torch.float32, device="cuda")
    dW = torch.empty(9 * cin * 32, dtype=torch.float32, device="cuda")
    db = torch.empty(32, dtype=torch.float32, device="cuda")
    nat.check(L.spde_conv_backward(M, N, T, cin, _ptr(Xd), _ptr(Gd), _ptr(Y), _ptr(dW), _ptr(db), _ptr(work), st))
    dXo, dWo, dbo = Le.conv_backward(X, W, G, Yg, shape)
    assert np.allclose(dW.cpu().numpy().reshape(9 * cin, 32), dWo, rtol=1e-4, atol=1e-5 * np.abs(dWo).max())
    assert np.allclose(db.cpu().numpy(), dbo, rtol=1e-4, atol=1e-5 * np.abs(dbo).max())
    if cin == 32:
        dX = torch.empty((P, 32), dtype=torch.float32, device="cuda")
        nat.check(L.spde_conv_forward(M, N, T, 32, _ptr(Gd), _ptr(Y), _ptr(fb), None, 0, _ptr(dX), st))
        assert np.allclose(dX.cpu().numpy(), dXo, rtol=1e-4, atol=1e-5 * np.abs(dXo).max())


@pytest.fixture(scope="module")
def ic_case():
    import torch
    step("learned: setup 12x10x6")
    c = F.small(12, 10, 6, bc=1, m=60)
    mod = product_case(c)
    mod.initFit(c["y"], idx=c["idx"], fitQ0=False)
    par = c["par"]
    eng = mod._engine()
    qval = eng.assemble(mod._coeffs(par))
    cols = mod._obs_cols()
    dadd = np.bincount(cols, minlength=eng.n).astype(np.float64) * np.exp(par[-1])
    e = eng.plan.export()
    data = qval.cpu().numpy().copy()
    rows = np.repeat(np.arange(eng.n), np.diff(e["indptr"]))
    data[np.nonzero(rows == e["indices"])[0]] += dadd
    A = sparse.csr_matrix((data, e["indices"], e["indptr"]), shape=(eng.n, eng.n))
    return dict(mod=mod, eng=eng, qval=qval, dadd=torch.as_tensor(dadd).cuda(), A=A, e=e)


def test_features_and_pair_channels(ic_case):
    from spdepy_amd.learned import LearnedPreconditioner
    eng = ic_case["eng"]
    it = eng.iterative()
    it.factor(ic_case["qval"], ic_case["dadd"], shift=0.1)
    ex = it.export()
    M, N, T = eng.plan.shape
    rows = np.repeat(np.arange(eng.n), np.diff(ex["lptr"]))
    assert np.array_equal(ex["lch"][:it.nlow], Le.pair_channels(M, N, T, rows, ex["lcol"][:it.nlow]))
    step("features")
    net = LearnedPreconditioner(seed=0)
    f = net.features(it).cpu().numpy()
    A = ic_case["A"]
    ad = A.diagonal()
    code = ic_case["e"]["code"]
    want = np.zeros((eng.n, 8))
    want[:, 0] = ic_case["dadd"].cpu().numpy() / ad
    Ac = A.tocoo()
    slot = {13: 1, 11: 2, 17: 3, 7: 4, 38: 5, 29: 6}
    for r_, c_, v, cd in zip(Ac.row, Ac.col, Ac.data, code):
        if cd in slot:
            want[r_, slot[cd]] = v / np.sqrt(ad[r_] * ad[c_])
    assert np.allclose(f, want, rtol=1e-6, atol=1e-7)


def test_network_and_kl_gradient_match_oracle(ic_case):
    import torch
    from spdepy_amd.learned import LearnedPreconditioner, NPAIR
    from spdepy_amd import _native as nat
    from spdepy_amd.plan import _ptr, _stream
    eng = ic_case["eng"]
    shape = eng.plan.shape
    it = eng.iterative()
    it.factor(ic_case["qval"], ic_case["dadd"], shift=0.1)
    net = LearnedPreconditioner(seed=3)
    # a larger output layer than the trained start, so the correction matters
    th = net.params()
    th[-(9 * 32 * 32 + 32):] *= 30.0
    net = LearnedPreconditioner(params=th)
    step("network forward")
    feat = net.features(it)
    acts = net.forward(feat, shape)
    acts_o = Le.forward(th, feat.cpu().numpy(), shape)
    for a, ao in zip(acts[1:], acts_o[1:]):
        assert np.allclose(a.cpu().numpy(), ao, rtol=1e-4, atol=1e-4 * np.abs(ao).max())
    corr = acts[-1][:, :NPAIR].contiguous()
    lo0, dg0, ad = [t.cpu().numpy() for t in it.values()]
    it.correct(corr)
    lo1, dg1, _ = [t.cpu().numpy() for t in it.values()]
    ex = it.export()
    lptr, lcol, lch = ex["lptr"], ex["lcol"][:it.nlow], ex["lch"][:it.nlow]
    lo_o, dg_o = Le.corrected_factor(lo0, dg0, lptr, lcol, lch, ad, corr.cpu().numpy())
    assert np.allclose(lo1, lo_o, rtol=1e-12, atol=1e-12 * np.abs(lo_o).max())
    assert np.allclose(dg1, dg_o, rtol=1e-12)
    step("KL gradient")
    x = torch.randn((eng.n, 16), dtype=torch.float64, device="cuda") * 0.05
    dout = torch.empty((eng.n, NPAIR), dtype=torch.float32, device="cuda")
    import ctypes
    loss = ctypes.c_double()
    nat.check(nat.lib().spde_ic_kl_grad(it._h, _ptr(x), 16, 1.0, _ptr(dout), ctypes.byref(loss), _stream(torch)))
    loss_o, dout_o = Le.kl_grad(lo1, dg1, lptr, lcol, lch, ad, x.cpu().numpy())
    assert abs(loss.value - loss_o) <= 1e-10 * abs(loss_o)
    assert np.allclose(dout.cpu().numpy(), dout_o, rtol=1e-5, atol=1e-6 * np.abs(dout_o).max())
    step("Jeffreys gradient")
    z = it.rademacher(16, seed=4)
    dj = torch.empty((eng.n, NPAIR), dtype=torch.float32, device="cuda")
    nat.check(nat.lib().spde_ic_jeffreys_grad(it._h, _ptr(x), _ptr(z), 16, 1.0, _ptr(dj), ctypes.byref(loss),
                                              _stream(torch)))
    M, N, T = shape
    lj, dj_o = Le.jeffreys_grad(lo1, dg1, lptr, lcol, lch, ic_case["A"], K.colours(M, N, T), x.cpu().numpy(),
                                z.cpu().numpy().astype(np.float64))
    assert abs(loss.value - lj) <= 1e-4 * abs(lj) + 1e-6 * eng.n
    assert np.allclose(dj.cpu().numpy(), dj_o, rtol=1e-3, atol=1e-3 * np.abs(dj_o).max())
    step("network backward")
    g = net.backward(acts, dout, shape).cpu().numpy()
    g_o = Le.backward(th, acts_o, dout.cpu().numpy().astype(np.float64), shape)
    assert np.allclose(g, g_o, rtol=1e-3, atol=1e-4 * np.abs(g_o).max())


def test_training_beats_plain_ic0():
    """Train on 300 draws at 20 x 20 x 6 (the GPU Cholesky path supplies the
    samples), then on held-out draws compare the symmetrised divergence and
    the PCG iterations with the trained correction against plain IC(0)."""
    import torch
    import spdepy_amd as sp
    from spdepy_amd import spdes
    from spdepy_amd.learned import LearnedPreconditioner, train, draw_parameters, exact_samples, jeffreys_loss
    ax = (np.arange(20) + 0.5) / 20
    g = sp.grid(x=ax, y=ax, t=np.arange(6.0), Nbs=4)
    g0 = sp.grid(x=ax, y=ax, Nbs=4)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=1)
    mod = spdes.VarAdvectionVarDiffusion2D(grid=g, mod0=mod0, bc=1)
    net = LearnedPreconditioner(seed=0, lr=3e-3)
    step("train 300 draws 20x20x6")
    losses = train(net, mod, 300, K=16, seed=1)
    print("  loss first %.5f last %.5f" % (losses[:10].mean(), losses[-10:].mean()))
    assert losses[-20:].mean() < losses[:20].mean()
    eng = mod._engine()
    it = eng.iterative()
    rng = np.random.default_rng(99)
    gains = []
    for d in range(3):
        par = draw_parameters(mod.Np, rng)
        qval = eng.assemble(mod._coeffs(par))
        dadd = np.zeros(eng.n)
        dadd[rng.choice(eng.n, eng.n // 40, replace=False)] = 100.0
        dd = torch.as_tensor(dadd).cuda()
        x = exact_samples(eng, qval, dd, 64, 7 + d)
        B = torch.randn((eng.n, 8), dtype=torch.float32, device="cuda")
        it.factor(qval, dd, shift=0.1)
        j_ic = jeffreys_loss(it, x)
        _, it_ic = it.solve(B, tol=1e-6)
        it.factor(qval, dd, shift=0.1)
        it.correct(net.correction_for(it, eng.plan.shape))
        j_ln = jeffreys_loss(it, x)
        _, it_ln = it.solve(B, tol=1e-6)
        print("  draw %d: divergence/row IC %.4f learned %.4f; PCG iterations IC %.1f learned %.1f"
              % (d, j_ic, j_ln, it_ic.mean(), it_ln.mean()))
        gains.append((j_ic - j_ln, it_ic.mean() / it_ln.mean()))
    gains = np.array(gains)
    assert np.all(gains[:, 0] > 0)          # lower divergence on every held-out draw
    assert gains[:, 1].mean() > 1.0         # fewer PCG iterations on average


def test_learned_route_and_cholesky_fallback():
    """logLike(method="learned") with the shipped weights: without a
    tolerance the estimate lies within 4 standard errors of the exact value;
    at the default 1e-6 the stochastic log-determinant cannot meet it, the
    route sees that from a 4-probe pilot of the Lanczos quadrature and returns
    the exact Cholesky value and gradient; at a loose tolerance neither the
    pilot nor the full estimate falls back."""
    c = F.small(12, 10, 6, bc=1, m=60)
    mod = product_case(c)
    mod.initFit(c["y"], idx=c["idx"], fitQ0=False)
    par = c["par"]
    step("exact logLike 12x10x6")
    l_ex, j_ex = mod.logLike(par)
    step("learned route without a tolerance")
    l_ln, j_ln = mod.logLike(par, method="learned", seed=5, tol_like=np.inf, hutch_probes=256)
    p = mod.last_parts
    se = 0.5 * p["r"] * p["logdet_qc_se"] / (p["m"] * p["r"])
    print("  exact %.10f learned %.10f (se %.2e)" % (l_ex, l_ln, se))
    assert not p["fallback"] and abs(l_ln - l_ex) <= 4 * se + 1e-9 * abs(l_ex)
    step("learned route at 1e-6: Cholesky fallback")
    l_fb, j_fb = mod.logLike(par, method="learned", seed=5)
    p = mod.last_parts
    assert p["fallback"] and p["infeasible"] and p["pilot"] and p["like_rel_se"] > 1e-6
    assert l_fb == l_ex and np.array_equal(j_fb, j_ex)
    step("learned route at 0.5: no fallback")
    l_lo, _ = mod.logLike(par, method="learned", seed=5, tol_like=0.5, hutch_probes=64)
    p = mod.last_parts
    assert not p["fallback"] and not p.get("infeasible") and p["like_rel_se"] <= 0.5
    assert abs(l_lo - l_ex) <= 4 * 0.5 * p["r"] * p["logdet_qc_se"] / (p["m"] * p["r"]) + 1e-9 * abs(l_ex)


def _held_out_iterations(mod, net, ndraws, seed, extra=()):
    """Mean PCG iterations (tolerance 1e-5, 8 right-hand sides) with plain
    IC(0) and with the network's correction, on held-out draws of the
    B-spline model (observations on 2.5% of the lattice, tau = 100)."""
    import torch
    from spdepy_amd.learned import draw_parameters
    eng = mod._engine()
    it = eng.iterative()
    n = eng.n
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(ndraws):
        par = draw_parameters(mod.Np, rng)
        qval = eng.assemble(mod._coeffs(par)).clone()
        dadd = np.zeros(n)
        dadd[rng.choice(n, n // 40, replace=False)] = 100.0
        cases.append((qval, torch.as_tensor(dadd).cuda()))
    for par, cols in extra:
        qval = eng.assemble(mod._coeffs(par)).clone()
        dadd = np.zeros(n)
        dadd[cols] = np.exp(par[-1])
        cases.append((qval, torch.as_tensor(dadd).cuda()))
    B = torch.as_tensor(np.random.default_rng(seed + 1).standard_normal((n, 8)).astype(np.float32)).cuda()
    out = []
    for qval, dd in cases:
        it.factor(qval, dd, shift=0.1)
        _, plain = it.solve(B, tol=1e-5)
        it.factor(qval, dd, shift=0.1)
        it.correct(net.correction_for(it, eng.plan.shape))
        _, learned = it.solve(B, tol=1e-5)
        out.append((plain.mean(), learned.mean()))
    return np.array(out)


def test_kaporin_training_beats_plain_ic0():
    """Training on Kaporin's condition number (log of the mean over the mean
    log of the eigenvalues of L'^-1 Q_c L'^-T, which bounds CG's iteration
    count) for 300 draws at 20 x 20 x 6: fewer PCG iterations than plain
    IC(0) on every held-out draw."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    from spdepy_amd.learned import LearnedPreconditioner, train
    ax = (np.arange(20) + 0.5) / 20
    g = sp.grid(x=ax, y=ax, t=np.arange(6.0), Nbs=4)
    g0 = sp.grid(x=ax, y=ax, Nbs=4)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=1)
    mod = spdes.VarAdvectionVarDiffusion2D(grid=g, mod0=mod0, bc=1)
    net = LearnedPreconditioner(seed=0)
    step("train 300 draws 20x20x6 (Kaporin)")
    train(net, mod, 300, K=16, seed=1, lr=1e-3, loss="kaporin")
    res = _held_out_iterations(mod, net, 3, 99)
    for p, l in res:
        print("  PCG iterations IC %.1f learned %.1f" % (p, l))
    assert np.all(res[:, 1] < res[:, 0])


def test_shipped_weights_beat_plain_ic0_at_100x100x20():
    """The shipped network (spdepy_amd/weights/learned_ic.npz) on held-out
    draws at 100 x 100 x 20 and on the benchmark workload's parameters and
    observation sites: fewer PCG iterations than plain IC(0) on every case."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    from spdepy_amd.learned import default_preconditioner
    M = N = 100
    ax = (np.arange(M) + 0.5) / M
    g = sp.grid(x=ax, y=ax, t=np.arange(20.0), Nbs=8)
    g0 = sp.grid(x=ax, y=ax, Nbs=8)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=1)
    mod = spdes.VarAdvectionVarDiffusion2D(grid=g, mod0=mod0, bc=1)
    Np = g.Nbs2
    rb = np.random.default_rng(3)
    pb = np.hstack([np.zeros(Np), np.full(Np, np.log(0.5)), 0.3 * rb.normal(size=4 * Np), [0.0, np.log(100.0)]])
    cols = np.sort(np.random.default_rng(4).choice(M * N * 20, 5000, replace=False))
    step("shipped weights vs plain IC(0) at 100x100x20")
    res = _held_out_iterations(mod, default_preconditioner(), 2, 77, extra=[(pb, cols)])
    for p, l in res:
        print("  PCG iterations IC %.1f learned %.1f (ratio %.3f)" % (p, l, l / p))
    assert np.all(res[:, 1] < res[:, 0])
