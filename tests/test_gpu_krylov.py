"""GPU checks of the iterative path (pcg.hip through the C ABI) against the
oracle (oracle/krylov.py) on the same inputs: IC(0) values, triangular sweeps,
FP32 PCG residuals, FP64 refinement, Rademacher probes, Lanczos coefficients
and the SLQ log-determinant, the Hutchinson pattern estimate, and the
iterative log-likelihood against the exact path."""
import numpy as np
import pytest
from scipy import sparse
from scipy.sparse import linalg as spla

from oracle import fixtures as F, krylov as K
from tests.gpu_util import product_case, step

pytestmark = pytest.mark.gpu


def _setup(M=12, N=10, T=6, m=60, bc=1):
    import torch
    c = F.small(M, N, T, bc=bc, m=m)
    mod = product_case(c)
    mod.initFit(c["y"], idx=c["idx"], fitQ0=False)
    par = c["par"]
    eng = mod._engine()
    coeffs = mod._coeffs(par)
    qval = eng.assemble(coeffs)
    cols = mod._obs_cols()
    tau = np.exp(par[-1])
    dadd = np.bincount(cols, minlength=eng.n).astype(np.float64) * tau
    dadd_d = torch.as_tensor(dadd).cuda()
    # Q_c on the structural pattern with its explicit zeros kept (scipy's
    # sparse addition would drop them, and IC(0) depends on the pattern)
    e = eng.plan.export()
    data = qval.cpu().numpy().copy()
    rows = np.repeat(np.arange(eng.n), np.diff(e["indptr"]))
    data[np.nonzero(rows == e["indices"])[0]] += dadd
    A = sparse.csr_matrix((data, e["indices"], e["indptr"]), shape=(eng.n, eng.n))
    return dict(c=c, mod=mod, eng=eng, coeffs=coeffs, qval=qval, dadd=dadd_d, A=A, col=K.colours(M, N, T),
                coeffs_keep=coeffs)


@pytest.fixture(scope="module")
def case():
    step("iterative path: setup 12x10x6")
    return _setup()


def test_ic0_values_match_oracle(case):
    eng = case["eng"]
    it = eng.iterative()
    step("spde_ic_factor 12x10x6")
    it.factor(case["qval"], case["dadd"], shift=0.1)
    assert it.shift == 0.1
    lo, dg, ad = it.values()
    ex = it.export()
    L, d = K.ic0(case["A"], case["col"], 0.1)
    lo = lo.cpu().numpy()
    rows = np.repeat(np.arange(eng.n), np.diff(ex["lptr"]))
    ref = np.asarray(L[rows, ex["lcol"][:it.nlow]]).ravel()
    scale = np.abs(ref).max()
    assert np.allclose(lo, ref, rtol=1e-11, atol=1e-12 * scale)
    assert np.allclose(dg.cpu().numpy(), d, rtol=1e-12)
    assert np.allclose(ad.cpu().numpy(), case["A"].diagonal(), rtol=1e-15)
    assert np.array_equal(ex["colour"], case["col"])


def test_sweeps_apply_and_spmm(case):
    import torch
    eng = case["eng"]
    it = eng.iterative()
    it.factor(case["qval"], case["dadd"], shift=0.1)
    L, d = K.ic0(case["A"], case["col"], 0.1)
    Fm = K.full_factor(L, d).toarray()   # lower in colour order, indexed by lattice indices
    R = torch.randn(eng.n, 4, dtype=torch.float32, device="cuda")
    Rh = R.cpu().numpy().astype(np.float64)
    order = np.lexsort((np.arange(eng.n), case["col"]))
    Fo = Fm[np.ix_(order, order)]
    step("sweeps")
    y = it.sweep(1, R).cpu().numpy()
    yo = np.empty_like(Rh)
    yo[order] = np.linalg.solve(Fo, Rh[order])
    assert np.allclose(y, yo, rtol=1e-4, atol=1e-4 * np.abs(yo).max())
    x = it.sweep(2, R).cpu().numpy()
    xo = np.empty_like(Rh)
    xo[order] = np.linalg.solve(Fo.T, Rh[order])
    assert np.allclose(x, xo, rtol=1e-4, atol=1e-4 * np.abs(xo).max())
    z = it.apply(R).cpu().numpy()
    zo = np.empty_like(Rh)
    zo[order] = np.linalg.solve(Fo.T, np.linalg.solve(Fo, Rh[order]))
    assert np.allclose(z, zo, rtol=1e-4, atol=1e-4 * np.abs(zo).max())
    Y = it.spmm(R).cpu().numpy()
    assert np.allclose(Y, case["A"] @ Rh, rtol=1e-5, atol=1e-5 * np.abs(case["A"] @ Rh).max())


@pytest.mark.parametrize("k", [1, 8, 64])
def test_pcg_solve_residual(case, k):
    import torch
    eng = case["eng"]
    it = eng.iterative()
    it.factor(case["qval"], case["dadd"], shift=0.1)
    B = torch.randn(eng.n, k, dtype=torch.float32, device="cuda")
    step("pcg k=%d" % k)
    X, iters = it.solve(B, tol=1e-6, maxit=3000)
    Bh = B.cpu().numpy().astype(np.float64)
    Xh = X.cpu().numpy().astype(np.float64)
    res = np.linalg.norm(Bh - case["A"] @ Xh, axis=0) / np.linalg.norm(Bh, axis=0)
    print("  iterations", iters.min(), iters.max(), "max relres", res.max())
    assert np.all(iters > 0) and np.all(iters < 3000)
    # the recursive FP32 residual reaches 1e-6; the true residual, evaluated in
    # FP64, is what FP32 arithmetic attains at this conditioning (refinement
    # in FP64 takes it further)
    assert res.max() < 2e-4


def test_refinement_reaches_fp64(case):
    import torch
    eng = case["eng"]
    it = eng.iterative()
    it.factor(case["qval"], case["dadd"], shift=0.1)
    b = torch.randn(eng.n, dtype=torch.float64, device="cuda")
    step("refine")
    x, info = it.refine(b, tol=1e-10)
    bh, xh = b.cpu().numpy(), x.cpu().numpy()
    rel = np.linalg.norm(bh - case["A"] @ xh) / np.linalg.norm(bh)
    print("  refinement", info, "host relres %.2e" % rel)
    assert info["relres"] <= 1e-10 and rel <= 2e-10
    xd = spla.spsolve(case["A"].tocsc(), bh)
    assert np.allclose(xh, xd, rtol=1e-7, atol=1e-8 * np.abs(xd).max())


@pytest.mark.parametrize("delta", [0.1, 1e-3])
def test_residual_replacement_reaches_fp64(case, delta):
    """spde_pcg_replace: FP32 PCG with the FP64 residual replacing the
    recursive one reaches the FP64 tolerance and the direct solution, also
    with long segments (delta 1e-3), where a segment that makes the true
    residual worse is taken back."""
    import torch
    eng = case["eng"]
    it = eng.iterative()
    it.factor(case["qval"], case["dadd"], shift=0.1)
    b = torch.randn(eng.n, dtype=torch.float64, device="cuda")
    step("residual replacement delta %g" % delta)
    x, info = it.replace(b, tol=1e-10, delta=delta)
    _, info_r = it.refine(b, tol=1e-10)
    bh, xh = b.cpu().numpy(), x.cpu().numpy()
    rel = np.linalg.norm(bh - case["A"] @ xh) / np.linalg.norm(bh)
    print("  replacement", info, "restarted refinement", info_r, "host relres %.2e" % rel)
    assert info["relres"] <= 1e-10 and rel <= 2e-10 and info["updates"] >= 1
    xd = spla.spsolve(case["A"].tocsc(), bh)
    assert np.allclose(xh, xd, rtol=1e-7, atol=1e-8 * np.abs(xd).max())


def test_residual_replacement_edge_cases(case):
    """A zero right-hand side returns zero with no iteration; a zero column
    of a block stops at once while the others converge; an iteration cap
    returns the residual reached instead of failing."""
    import torch
    eng = case["eng"]
    it = eng.iterative()
    it.factor(case["qval"], case["dadd"], shift=0.1)
    step("residual replacement: zero right-hand side, zero column, iteration cap")
    x, info = it.replace(torch.zeros(eng.n, dtype=torch.float64, device="cuda"))
    assert info["iters"] == 0 and info["relres"] == 0.0 and float(x.abs().max()) == 0.0
    B = torch.randn(eng.n, 4, dtype=torch.float64, device="cuda")
    B[:, 2] = 0.0
    X, info = it.replace_block(B, tol=1e-10)
    print("  block with a zero column:", info["iters"], info["relres"])
    assert info["iters"][2] == 0 and info["relres"][2] == 0.0 and float(X[:, 2].abs().max()) == 0.0
    assert np.all(np.delete(info["relres"], 2) <= 1e-10)
    b = torch.randn(eng.n, dtype=torch.float64, device="cuda")
    x, info = it.replace(b, tol=1e-10, maxit=8)
    print("  capped at 8 iterations:", info)
    assert info["iters"] == 8 and 0.0 < info["relres"] <= 1.0


def test_view_solves_beside_its_parent(case):
    """spde_ic_view: a second handle on the factor with its own work buffers
    gives the parent's results bit for bit, also while both solve at once on
    two streams from two host threads; a view cannot refactor, and it keeps
    working after its parent's handle is destroyed.  On a created stream the
    refinement's inner iterations replay a captured HIP graph (kernel by
    kernel on the default stream): every call, each of many replays with host
    polls between them, equals the launches one by one."""
    import threading
    import torch
    from spdepy_amd import _native as nat
    from spdepy_amd.krylov import IterativeSolver
    eng = case["eng"]
    it = IterativeSolver(eng.plan)
    it.factor(case["qval"], case["dadd"], shift=0.1)
    v = it.view()
    g = torch.Generator(device="cuda").manual_seed(5)
    b = torch.randn(eng.n, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(eng.n, 64, dtype=torch.float32, device="cuda", generator=g)
    step("parent and view, one after the other")
    Bd = B[:, :4].double().contiguous()
    x0, i0 = it.refine(b, tol=1e-10)
    X0, j0 = it.solve(B, tol=1e-6)
    Y0, k0 = it.refine_block(Bd, tol=1e-10)
    x1, i1 = v.refine(b, tol=1e-10)
    assert torch.equal(x0, x1) and i0 == i1
    print("  refinement", i0, "block", k0["iters"])
    step("parent and view at once on two streams")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    box = {}

    def run():
        with torch.cuda.stream(side):
            box["x"] = [v.refine(b, tol=1e-10) for _ in range(2)]
            box["Y"] = [v.refine_block(Bd, tol=1e-10) for _ in range(2)]
    th = threading.Thread(target=run)
    th.start()
    X2, j2 = it.solve(B, tol=1e-6)
    th.join()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    for x, inf in box["x"]:
        assert torch.equal(x, x0) and inf == i0
    for Y, inf in box["Y"]:
        assert torch.equal(Y, Y0) and np.array_equal(inf["iters"], k0["iters"])
    assert torch.equal(X2, X0) and np.array_equal(j2, j0)
    with pytest.raises(nat.SpdeError, match="view"):
        v.factor(case["qval"], case["dadd"], shift=0.1)
    step("view after its parent's handle is destroyed")
    nat.lib().spde_ic_destroy(it._h)
    it._h = None
    x3, _ = v.refine(b, tol=1e-10)
    assert torch.equal(x3, x0)
    del v


def test_rademacher_and_hutch_pattern(case):
    import torch
    eng = case["eng"]
    it = eng.iterative()
    step("rademacher + hutch pattern")
    V = it.rademacher(8, seed=123)
    assert np.array_equal(V.cpu().numpy(), K.rademacher(eng.n, 8, 123).astype(np.float32))
    U = torch.randn(eng.n, 8, dtype=torch.float32, device="cuda")
    sig = it.hutch_pattern(U, V).cpu().numpy()
    e = eng.plan.export()
    Uh, Vh = U.cpu().numpy().astype(np.float64), V.cpu().numpy().astype(np.float64)
    rows = np.repeat(np.arange(eng.n), np.diff(e["indptr"]))
    cols = e["indices"]
    ref = 0.5 * ((Uh[rows] * Vh[cols]).sum(1) + (Uh[cols] * Vh[rows]).sum(1)) / 8
    assert np.allclose(sig, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_lanczos_and_slq(case):
    eng = case["eng"]
    it = eng.iterative()
    ld_m = it.factor(case["qval"], case["dadd"], shift=0.1)
    L, d = K.ic0(case["A"], case["col"], 0.1)
    assert abs(ld_m - 2 * np.log(d).sum()) <= 1e-11 * abs(ld_m)
    step("slq 32 x 40")
    a, b = it.slq(32, 40, seed=5)
    n = eng.n
    order = np.lexsort((np.arange(n), case["col"]))
    Fo = K.full_factor(L, d).toarray()[np.ix_(order, order)]
    Ao = case["A"].toarray()[np.ix_(order, order)]
    Bm = np.linalg.solve(Fo, np.linalg.solve(Fo, Ao).T)
    V = K.rademacher(n, 32, 5)[order] / np.sqrt(n)
    a0, b0 = K.lanczos(lambda x: Bm @ x, V[:, 0], 6)
    assert np.allclose(a[0, :6], a0, rtol=2e-4)
    assert np.allclose(b[0, :5], b0[:5], rtol=2e-3)
    est, se, _ = K.slq_logdet(a, b, n)
    exact = np.linalg.slogdet(Ao)[1] - ld_m
    print("  SLQ %.6f +- %.6f exact %.6f" % (est, se, exact))
    assert abs(est - exact) <= 4 * se + 1e-5 * abs(np.linalg.slogdet(Ao)[1])


def test_lanczos_from_pcg_coefficients(case):
    """The PCG step coefficients of Q_c u = L w give the Lanczos tridiagonal
    of L^-1 Q_c L^-T started at w: equal to the explicit Lanczos (spde_slq,
    same Rademacher probes) over the first steps, and their quadrature
    estimates log|Q_c| - log|L L^T| within its standard error."""
    from spdepy_amd.krylov import cg_lanczos, slq_quadrature
    eng = case["eng"]
    it = eng.iterative()
    ld_m = it.factor(case["qval"], case["dadd"], shift=0.1)
    step("PCG coefficients vs explicit Lanczos")
    W = it.rademacher(8, seed=21)
    U, its, a, b = it.solve_coeffs(it.lmul(W), tol=1e-6, cap=30)
    d, e = cg_lanczos(a, b, its)
    la, lb = it.slq(8, 30, seed=21)
    print("  iterations", its, "diag cg", d[0, :4], "lanczos", la[0, :4])
    assert np.allclose(d[:, :6], la[:, :6], rtol=2e-3)
    assert np.allclose(e[:, :5], lb[:, :5], rtol=2e-3)
    est, se, _ = slq_quadrature(d, e, eng.n)
    exact = np.linalg.slogdet(case["A"].toarray())[1] - ld_m
    print("  log|Q_c| - log|M|: CG quadrature %.6f +- %.6f, exact %.6f" % (est, se, exact))
    assert abs(est - exact) <= 4 * se + 1e-6 * abs(exact)


def test_pcg_loglike_against_exact(case):
    mod, par = case["mod"], case["c"]["par"]
    step("logLike exact vs pcg 12x10x6")
    l_ex, j_ex = mod.logLike(par)
    l_it, j_it = mod.logLike(par, method="pcg", seed=3, hutch_probes=512)
    parts = mod.last_parts
    m, r = parts["m"], parts["r"]
    se_like = 0.5 * r * parts["logdet_qc_se"] / (m * r)
    rel_g = np.linalg.norm(j_it - j_ex) / np.linalg.norm(j_ex)
    print("  like exact %.10f pcg %.10f (se %.2e)  grad rel err %.3f  pcg %s" % (l_ex, l_it, se_like, rel_g,
                                                                              parts["pcg"]["mu"]))
    assert abs(l_it - l_ex) <= 4 * se_like + 1e-9 * abs(l_ex)
    assert parts["pcg"]["mu"][0]["relres"] <= 1e-10
    assert rel_g < 0.3     # 512 Hutchinson probes on n = 720
    # mu_c from the refined PCG solve equals the exact path's
    mu_it = parts["mu"].cpu().numpy()
    mod.logLike(par, grad=False)
    eng = mod._engine()
    like_ex, p_ex = eng.loglik(mod._coeffs(par), mod._obs_cols(), mod.data, par[-1], grad=False)
    mu_ex = p_ex["mu"].cpu().numpy()
    assert np.allclose(mu_it, mu_ex, rtol=1e-8, atol=1e-9 * np.abs(mu_ex).max())


def test_gaussian_draws_match_oracle():
    from spdepy_amd.krylov import gaussian
    step("spde_gaussian 3000 x 8")
    z = gaussian(3000, 8, seed=21).cpu().numpy()
    ref = K.gaussian(3000, 8, seed=21)
    # the device log/cos and numpy's differ by at most a few ulp
    assert np.allclose(z, ref, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("bc", [1, 3])
def test_fform_square_root_matches_oracle(bc):
    """spde_fform_apply_t against the oracle's F^T D^(1/2) z, and R R^T = Q
    (R applied to the identity) against the assembled Q."""
    import torch
    from spdepy_amd import _native as nat
    from spdepy_amd.krylov import gaussian
    from spdepy_amd.plan import _ptr, _stream
    c = F.small(7, 6, 4, bc=bc, m=20)
    mod = product_case(c)
    g, model, par = c["grid"], c["model"], c["par"]
    eng = mod._engine()
    coeffs = mod._coeffs(par)
    qval = eng.assemble(coeffs)
    n = eng.n

    def apply(z):
        out = torch.empty_like(z)
        nat.check(nat.lib().spde_fform_apply_t(eng.plan._h, coeffs.c, _ptr(z), _ptr(out), z.shape[1],
                                               _stream(torch)))
        return out.cpu().numpy()
    step("F^T D^1/2 z bc=%d" % bc)
    Z = gaussian(n, 3, seed=5)
    A, q, sigma = model.spatial_operator(par)
    A0 = c["mod0"].A0(c["mod0"].par)[0]
    ref = K.fform_t(A, q, A0, g.V, g.shape[2], sigma, g.dt, Z.cpu().numpy())
    got = apply(Z)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
    R = apply(torch.eye(n, dtype=torch.float64, device="cuda"))
    Qd = eng.to_scipy(qval).toarray()
    assert np.allclose(R @ R.T, Qd, rtol=1e-12, atol=1e-12 * np.abs(Qd).max())


@pytest.mark.parametrize("solver", ["replace", "refine"])
def test_conditional_samples_by_refined_pcg(case, solver):
    """Q_c^-1 B for the conditional right-hand sides by the block PCG with
    FP64 residual replacement, and by FP64 refinement rounds: every column
    within 1e-10 relative residual, equal to the sparse direct solve of the
    same B; E[B] = tau S^T y checked through the draws' mean against mu_c."""
    from spdepy_amd.krylov import conditional_samples_pcg
    mod, par, eng = case["mod"], case["c"]["par"], case["eng"]
    step("conditional samples by %s PCG 12x10x6 (k=16)" % solver)
    X, info = conditional_samples_pcg(eng, mod._coeffs(par), mod._obs_cols(), mod.data, par[-1], k=16, seed=4,
                                      solver=solver)
    print("  %s, iterations %s, relres max %.2e" % (
        "%d replacements" % info["updates"] if solver == "replace" else "%d rounds" % info["rounds"],
        info["iters"], info["relres"].max()))
    assert info["relres"].max() <= 1e-10
    Xh, Bh = X.cpu().numpy(), info["B"].cpu().numpy()
    A = case["A"].tocsc()
    res = np.linalg.norm(Bh - A @ Xh, axis=0) / np.linalg.norm(Bh, axis=0)
    assert res.max() <= 2e-10
    Xd = spla.spsolve(A, Bh)
    assert np.allclose(Xh, Xd, rtol=1e-7, atol=1e-8 * np.abs(Xd).max())
    # the right-hand sides: B - tau S^T y is F^T D^1/2 z1 + sqrt(tau) S^T z2, zero-mean
    cols = mod._obs_cols()
    tau = np.exp(par[-1])
    sty = np.zeros(eng.n)
    np.add.at(sty, cols, tau * np.asarray(mod.data).ravel())
    mu = spla.spsolve(A, sty)
    dev = Xh - mu[:, None]
    # (x - mu)^T Q_c (x - mu) is chi^2_n per draw: mean n, sd sqrt(2n)
    chi = np.einsum("is,is->s", dev, A @ dev)
    assert abs(chi.mean() - eng.n) < 4 * np.sqrt(2 * eng.n / dev.shape[1])
