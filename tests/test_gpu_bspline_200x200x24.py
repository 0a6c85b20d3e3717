"""The mixed-precision configuration at 200x200x24 (n = 960,000): the FP32
PCG path with IC(0), FP64 iterative refinement to a relative residual of
1e-10 for the conditional mean (FP64 residual replacement, and restarted
refinement rounds) and 64 conditional samples, and the
stochastic Lanczos log-determinant (32 Rademacher probes x 50 steps, FP32
Lanczos vectors, FP64 tridiagonals), each with its error against the FP64
Cholesky path on the same inputs (non-stationary advection-diffusion, 8x8
B-spline fields with coefficients ~ N(0, 0.3^2), 24,000 observations of a
smooth synthetic field with noise sd 0.1)."""
import numpy as np
import pytest

from tests.gpu_util import step

pytestmark = pytest.mark.gpu


def _model(M=200, N=200, T=24, Nbs=8, m=24000, bc=1):
    import spdepy_amd as sp
    from spdepy_amd import spdes
    x = (np.arange(M) + 0.5) / M
    y = (np.arange(N) + 0.5) / N
    g = sp.grid(x=x, y=y, t=np.arange(float(T)), Nbs=Nbs)
    g0 = sp.grid(x=x, y=y, Nbs=Nbs)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array([0.0, np.log(0.5), 0.0]), bc=bc)
    mod = spdes.VarAdvectionVarDiffusion2D(grid=g, mod0=mod0, bc=bc)
    Np = g.Nbs2
    rng = np.random.default_rng(3)
    par = np.hstack([np.zeros(Np), np.full(Np, np.log(0.5)), 0.3 * rng.normal(size=4 * Np),
                     [0.0, np.log(100.0)]])
    idx = np.sort(np.random.default_rng(4).choice(M * N * T, m, replace=False))
    t, rem = idx // (M * N), idx % (M * N)
    xi, yi = x[rem % M], y[rem // M]
    field = np.sin(2 * np.pi * xi) * np.cos(2 * np.pi * yi) * np.cos(0.2 * t)
    data = field + 0.1 * np.random.default_rng(7).normal(size=m)
    mod.initFit(data, idx=idx, fitQ0=False)
    return mod, par


def test_iterative_path_against_cholesky_at_200x200x24():
    import torch
    from spdepy_amd.krylov import conditional_samples_pcg, loglik_pcg
    step("model and plans 200x200x24")
    mod, par = _model()
    eng = mod._engine()
    coeffs = mod._coeffs(par)
    cols, data = mod._obs_cols(), mod.data
    step("exact: FP64 Cholesky of Q_c at n = 960,000")
    like_ex, p_ex = eng.loglik(coeffs, cols, data, par[-1], grad=False)
    mu_ex = p_ex["mu"].cpu().numpy().ravel()
    step("iterative: IC(0), PCG with FP64 residual replacement for mu_c, SLQ 32 x 50")
    like_it, p_it = loglik_pcg(eng, coeffs, cols, data, par[-1], grad=False, slq_probes=32, slq_steps=50, seed=1)
    mu_it = p_it["mu"].cpu().numpy().ravel()
    mu_err = np.linalg.norm(mu_it - mu_ex) / np.linalg.norm(mu_ex)
    ld_err = p_it["ld_qc"] - p_ex["ld_qc"]
    like_rel = abs(like_it - like_ex) / abs(like_ex)
    print("  log|Q_c| exact %.6f SLQ %.6f (error %.3e, %.2f standard errors; relative %.2e)" % (
        p_ex["ld_qc"], p_it["ld_qc"], ld_err, ld_err / p_it["logdet_qc_se"], ld_err / abs(p_ex["ld_qc"])))
    print("  logLike exact %.12f iterative %.12f relative error %.2e" % (like_ex, like_it, like_rel))
    print("  mu_c: PCG %s, relative error vs Cholesky %.2e" % (p_it["pcg"]["mu"], mu_err))
    assert p_it["pcg"]["mu"][0]["relres"] <= 1e-10
    assert mu_err <= 1e-6
    step("mu_c by restarted refinement rounds")
    _, p_rf = loglik_pcg(eng, coeffs, cols, data, par[-1], grad=False, slq_probes=2, slq_steps=5, seed=1,
                         mean_solver="refine")
    mu_rf = p_rf["mu"].cpu().numpy().ravel()
    err_rf = np.linalg.norm(mu_rf - mu_ex) / np.linalg.norm(mu_ex)
    print("  mu_c: restarted rounds %s, relative error vs Cholesky %.2e" % (p_rf["pcg"]["mu"], err_rf))
    assert p_rf["pcg"]["mu"][0]["relres"] <= 1e-10 and err_rf <= 1e-6
    assert abs(ld_err) <= 4 * p_it["logdet_qc_se"] + 1e-9 * abs(p_ex["ld_qc"])
    for solver in ("replace", "refine"):
        step("64 conditional samples by %s PCG" % solver)
        X, info = conditional_samples_pcg(eng, coeffs, cols, data, par[-1], k=64, seed=2, solver=solver)
        print("  %s, iterations per sample %d..%d, relres max %.2e" % (
            "%d replacements" % info["updates"] if solver == "replace" else "%d rounds" % info["rounds"],
            info["iters"].min(), info["iters"].max(), info["relres"].max()))
        assert info["relres"].max() <= 1e-10
        step("the same 64 right-hand sides by the Cholesky factor")
        Xc = eng.plan.solve(1, info["B"])
        err = (torch.linalg.norm(X - Xc, dim=0) / torch.linalg.norm(Xc, dim=0)).cpu().numpy()
        print("  samples: relative error vs Cholesky max %.2e median %.2e" % (err.max(), np.median(err)))
        assert err.max() <= 1e-6
        del X, Xc


def test_exact_path_with_gradient_at_200x200x24():
    """The exact FP64 log-likelihood and gradient at n = 960,000 (Cholesky and
    selected inverse of Q_c, log|Q| from the spatial factors) fit one GPU's
    memory: log|Q| equals the oracle's closed form, the selected inverse
    satisfies sum_j (Q_c)_ij (Q_c^-1)_ij = 1 on every row, and the gradient
    agrees with Richardson central differences of the value."""
    import torch
    from oracle import fixtures as F, grid as G, models as Mo
    step("model 200x200x24, exact logLike + gradient")
    mod, par = _model()
    like, jac = mod.logLike(par)
    parts = mod.last_parts
    eng = mod._engine()
    import time
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    like2, jac2 = mod.logLike(par)
    torch.cuda.synchronize()
    print("  one evaluation with gradient: %.3f s" % (time.perf_counter() - t0))
    assert like2 == like and np.array_equal(jac2, jac)
    info = eng.plan.info
    print("  like %.12f; plan: nnz(L) %.3e, front doubles %.3e (%.0f GB per slot), selected-inverse arena "
          "the same" % (like, info["nnz_L"], info["front_doubles"], 8e-9 * info["front_doubles"]))
    step("oracle closed-form log|Q| at 200x200x24")
    M, N, T = 200, 200, 24
    g = G.STGrid(F.axis(M), F.axis(N), np.arange(float(T)), Nbs=8)
    mod0 = F.mod0_iso(M, N, Nbs=8)
    A, q, s = Mo.VarAdvVarDiff(g, mod0, bc=1).spatial_operator(par)
    A0 = mod0.A0(mod0.par)[0]
    cf = Mo.logdet_closed_form(A, q, A0, g.V, T, s, g.dt)
    print("  log|Q| GPU %.10f oracle %.10f" % (parts["ld_q"], cf))
    assert abs(parts["ld_q"] - cf) <= 1e-11 * abs(cf)
    step("selected inverse identity on every row")
    coeffs = mod._coeffs(par)
    qval = eng.assemble(coeffs)
    cols = np.asarray(mod._obs_cols())
    tau = np.exp(par[-1])
    dadd = torch.as_tensor(np.bincount(cols, minlength=eng.n).astype(np.float64) * tau).cuda()
    sig = eng.plan.selinv(1)
    indptr, indices = eng.plan.structure()
    rows = torch.as_tensor(np.repeat(np.arange(eng.n), np.diff(indptr))).cuda()
    qc = qval.clone()
    diag = rows == torch.as_tensor(indices).cuda()
    qc[diag] += dadd
    row_sum = torch.zeros(eng.n, dtype=torch.float64, device="cuda").index_add_(0, rows, qc * sig)
    err = float((row_sum - 1.0).abs().max())
    print("  max |sum_j (Q_c)_ij Sigma_ij - 1| = %.3e" % err)
    assert err < 1e-8
    del sig, rows, qc, diag, row_sum

    def value(p):
        return mod.logLike(p, grad=False)
    scale = np.abs(jac).max()
    Np = mod.Np
    for i in [3 * Np + 11, 6 * Np + 1]:
        step("central differences, parameter %d" % i)

        def central(h):
            pp, pm = par.copy(), par.copy()
            pp[i] += h
            pm[i] -= h
            return (value(pp) - value(pm)) / (2 * h)
        fd = (4 * central(1e-4) - central(2e-4)) / 3
        print("  param %d grad %+.10e fd %+.10e err/scale %.2e" % (i, jac[i], fd, abs(fd - jac[i]) / scale))
        assert abs(fd - jac[i]) <= 2e-6 * abs(jac[i]) + 2e-6 * scale
