"""GPU parity of the native path against the oracle (MI355X)."""
import os

import numpy as np
import pytest
from scipy import sparse

from oracle import fixtures as F, models as Mo, stencils as OS
from tests.gpu_util import product_case, step

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "stencils.npz"))


def _names():
    return sorted({k.split("/")[0] for k in GOLD.files if "/" in k})


def test_stencils_bitexact_vs_reference_golden():
    from spdepy_amd import spdes
    hx, hy = float(GOLD["hx"]), float(GOLD["hy"])
    bad = []
    for name in _names():
        kind, rest = name.split("_", 1)
        dims, b = rest.split("_b")
        M, N = map(int, dims.split("x"))
        bc = int(b[0])
        step("stencil " + name)
        if kind in ("AHv", "AHiso", "AHc", "AHciso"):
            got = spdes.stencil_diffusion((M, N), GOLD[name + "/in"], hx, hy, bc, kind in ("AHv", "AHiso"))
        elif kind in ("Awc", "Awc0"):
            got = spdes.stencil_advection((M, N), GOLD[name + "/in"], hx, hy, int(rest.split("_d")[1]), bc, False)
        else:
            got = spdes.stencil_advection((M, N), GOLD[name + "/G"], hx, hy, int(rest.split("_d")[1]), bc, True,
                                          GOLD[name + "/dG"])
        row = GOLD[name + "/row"]
        keep = row != M * N
        ok = (np.array_equal(got[0], row) and np.array_equal(got[1][keep], GOLD[name + "/col"][keep])
              and np.array_equal(got[2][keep], GOLD[name + "/val"][keep], equal_nan=True))
        if not ok:
            bad.append(name)
    assert not bad, bad


CASES = {
    "var_b1": lambda: F.small(9, 8, 4, bc=1, m=40),
    "var_b3": lambda: F.small(8, 9, 3, bc=3, m=40),
    "var_b1_ext": lambda: F.small(7, 6, 3, bc=1, m=30, ne=2),
    "iso_b1": lambda: F.stationary(10, 9, 4, bc=1, m=40),
    "ani_b3": lambda: F.stationary(9, 10, 3, bc=3, kind="ani", m=40),
    "ha_b1": lambda: F.stationary(8, 9, 4, bc=1, kind="ha", m=40),
    "ani_b2": lambda: F.stationary(9, 10, 3, bc=2, kind="ani", m=40),
    "var_b2": lambda: F.small(8, 9, 3, bc=2, m=40),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_assembly_pattern_bitexact_values(case):
    c = CASES[case]()
    step("assemble " + case)
    m = product_case(c)
    Qg, fac, _ = m.makeQ(c["par"], grad=False)
    Qo, _ = c["model"].makeQ(c["par"], grad=False)
    Qo = Qo.tocsc()
    Qo.sort_indices()
    Qg.sort_indices()
    assert np.array_equal(Qg.indptr, Qo.indptr)
    assert np.array_equal(Qg.indices, Qo.indices)
    assert np.allclose(Qg.data, Qo.data, rtol=1e-12, atol=1e-13 * np.abs(Qo.data).max())
    step("factor " + case)
    ld = fac.logdet()
    ref = np.linalg.slogdet(Qo.toarray())[1]
    assert abs(ld - ref) <= 1e-11 * abs(ref)


@pytest.mark.parametrize("leaf", [8, 40, 1000])
def test_factor_solve_selinv_small(leaf):
    """Multi-level trees (small leaves) and a single dense front (large leaf)."""
    import torch
    from spdepy_amd.engine import Coefficients, Engine
    c = CASES["var_b1"]()
    m = product_case(c)
    g = m.grid
    eng = Engine(g, 1, leaf_max=leaf)
    step("factor leaf=%d" % leaf)
    qval = eng.assemble(m._coeffs(c["par"]))
    Q = eng.to_scipy(qval).toarray()
    ld = eng.plan.factor(0, qval)
    assert abs(ld - np.linalg.slogdet(Q)[1]) <= 1e-11 * abs(np.linalg.slogdet(Q)[1])
    rng = np.random.default_rng(1)
    B = rng.normal(size=(g.n, 3))
    step("solve leaf=%d" % leaf)
    X = eng.plan.solve(0, torch.as_tensor(B).cuda()).cpu().numpy()
    assert np.allclose(Q @ X, B, rtol=1e-9, atol=1e-9 * np.abs(B).max())
    Y = eng.plan.solve(0, torch.as_tensor(B).cuda(), 1)
    X2 = eng.plan.solve(0, Y, 2).cpu().numpy()
    assert np.allclose(X2, X, rtol=1e-9, atol=1e-12)
    step("selinv leaf=%d" % leaf)
    s = eng.plan.selinv(0).cpu().numpy()
    indptr, indices = eng.plan.structure()
    rows = np.repeat(np.arange(g.n), np.diff(indptr))
    inv = np.linalg.inv(Q)
    assert np.allclose(s, inv[rows, indices], rtol=1e-9, atol=1e-12 * np.abs(inv).max())


@pytest.mark.parametrize("case", sorted(CASES))
def test_loglike_exact_gradient(case):
    c = CASES[case]()
    g = c["grid"]
    S = g.getS(c["idx"])
    Qo, dQo = c["model"].makeQ(c["par"], grad=True)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, S, c["y"], c["par"][-1], trace="exact")
    m = product_case(c)
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)
    step("logLike exact " + case)
    like, jac = m.logLike(c["par"])
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    assert jac.shape == jac_o.shape
    scale = np.abs(jac_o).max()
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale


@pytest.mark.parametrize("case", ["var_b1", "iso_b1", "ha_b1"])
def test_loglike_exact_gradient_several_realisations(case):
    """Data with r = 3 columns (independent realisations at the same
    locations, the reference's 2-D data): mu_c and the data terms take three
    right-hand sides (the narrow solve path), the trace term is scaled by r."""
    c = CASES[case]()
    g = c["grid"]
    S = g.getS(c["idx"])
    rng = np.random.default_rng(21)
    y3 = np.column_stack([c["y"], c["y"][::-1] * 0.7, 0.2 * rng.normal(size=c["y"].shape[0])])
    Qo, dQo = c["model"].makeQ(c["par"], grad=True)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, S, y3, c["par"][-1], trace="exact")
    m = product_case(c)
    m.initFit(y3, idx=c["idx"], fitQ0=False)
    step("logLike exact r=3 " + case)
    like, jac = m.logLike(c["par"])
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    scale = np.abs(jac_o).max()
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale


@pytest.mark.parametrize("case", ["var_b1", "ani_b3"])
def test_loglike_repeated_locations(case):
    """Several observations of one location (repeated rows of S): Q_c gets
    tau times the count on the diagonal and S^T y sums the repeats."""
    c = CASES[case]()
    g = c["grid"]
    rng = np.random.default_rng(33)
    idx = np.sort(np.concatenate([c["idx"], c["idx"][::4]]))
    y = rng.normal(size=idx.shape[0])
    S = g.getS(idx)
    assert S.shape[0] > len(set(idx.tolist()))
    Qo, dQo = c["model"].makeQ(c["par"], grad=True)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, S, y, c["par"][-1], trace="exact")
    m = product_case(c)
    m.initFit(y, idx=idx, fitQ0=False)
    step("logLike exact repeated locations " + case)
    like, jac = m.logLike(c["par"])
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    scale = np.abs(jac_o).max()
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale


@pytest.mark.parametrize("case", ["var_b1", "ani_b3", "iso_b1"])
def test_closed_form_and_factorised_logdet_agree(case):
    """The two routes to log|Q| and its gradient (spatial matrices of
    Q = F^T D F, or the full factorisation of Q) give the same result."""
    c = CASES[case]()
    m = product_case(c)
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)
    step("logLike closed vs factor " + case)
    l1, j1 = m.logLike(c["par"], logdet="closed")
    l2, j2 = m.logLike(c["par"], logdet="factor")
    assert abs(l1 - l2) <= 1e-11 * abs(l2)
    assert np.allclose(j1, j2, rtol=1e-10, atol=1e-10 * np.abs(j2).max())


def test_loglike_hutchinson_same_probes():
    c = CASES["var_b1"]()
    g = c["grid"]
    S = g.getS(c["idx"])
    Qo, dQo = c["model"].makeQ(c["par"], grad=True)
    V = Mo.reference_probes(g.n, 20, seed=11)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, S, c["y"], c["par"][-1], trace="hutchinson", probes=V)
    m = product_case(c)
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)
    np.random.seed(11)
    step("logLike hutchinson")
    like, jac = m.logLike(c["par"], nh1=20, trace="hutchinson")
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    scale = np.abs(jac_o).max()
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale)


def test_loglike_with_initial_state_parameters():
    """fitQ0: mod0 parameters inside par, their gradient from the Q0 block."""
    c = CASES["iso_b1"]()
    g = c["grid"]
    S = g.getS(c["idx"])
    par = np.hstack([c["par"][:-1], c["mod0"].par[:-1], c["par"][-1:]])
    Qo, dQo = c["model"].makeQ(par, grad=True)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, S, c["y"], par[-1], trace="exact")
    m = product_case(c)
    m.initFit(c["y"], idx=c["idx"])
    step("logLike fitQ0")
    like, jac = m.logLike(par)
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    scale = np.abs(jac_o).max()
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale)


def test_model_sample_and_update():
    """Model.sample: x = P^T L^-T z + mu with our ordering (Q x - mu part = P^T L z
    is not observable, so check Q x equals the solve of z's factor image);
    Model.update: mu = tau Q_c^-1 S^T y (model.py:120-127)."""
    import spdepy_amd as sp
    c = CASES["var_b1"]()
    g = c["grid"]
    m = product_case(c)
    m.setQ(par=c["par"])
    M = sp.Model.__new__(sp.Model)
    M.grid, M.mod, M.useCov = m.grid, m, False
    M.mu = np.zeros(g.n)
    M.setModel()
    step("Model.sample")
    data = M.sample(n=3, simple=True, seed=4)
    assert data.shape == (g.n, 3)
    Qd = c["model"].makeQ(c["par"], grad=False)[0].toarray()
    z = np.random.default_rng(4).normal(size=g.n * 3).reshape(g.n, 3)
    # covariance identity: x^T Q x has expectation n; with 3 draws check the
    # exact relation L^T P x = z through the factor's own solve
    x = data
    fac = M.Q_fac
    back = fac.pt_solve_Lt(fac.solve_L_p(Qd @ x))
    assert np.allclose(np.asarray(back), x, rtol=1e-8, atol=1e-10)
    step("Model.update")
    S = g.getS(c["idx"])
    M.update(c["y"], c["idx"], tau=np.exp(c["par"][-1]))
    tau = np.exp(c["par"][-1])
    Qc = Qd + tau * (S.T @ S).toarray()
    mu_ref = np.linalg.solve(Qc, tau * (S.T @ c["y"]))
    assert np.allclose(M.mu, mu_ref, rtol=1e-9, atol=1e-12)


def test_factor_cholmod_calls():
    """The CHOLMOD Factor calls (model.py:79-80,125-126) with CHOLMOD's
    meaning: apply_P / apply_Pt permute (P from the plan's ordering),
    solve_L / solve_Lt solve with L alone in elimination order, and the
    fused calls equal their compositions; L L^T = P Q P^T is checked
    against the dense matrix."""
    c = CASES["var_b1"]()
    g = c["grid"]
    m = product_case(c)
    step("factor for the CHOLMOD calls")
    Q, fac, _ = m.makeQ(c["par"], grad=False)
    Qd = Q.toarray()
    perm = m._engine().plan.export()["perm"]
    b = np.random.default_rng(9).normal(size=(g.n, 3))
    pb = fac.apply_P(b)
    assert np.array_equal(pb, b[perm])
    assert np.array_equal(fac.apply_Pt(pb), b)
    assert np.array_equal(fac.apply_P(b[:, 0]), b[perm, 0])
    y = fac.solve_L(pb)
    assert np.allclose(y, fac.solve_L_p(b), rtol=1e-13, atol=1e-13 * np.abs(y).max())
    x = fac.apply_Pt(fac.solve_Lt(y))
    assert np.allclose(x, np.linalg.solve(Qd, b), rtol=1e-9, atol=1e-11 * np.abs(x).max())
    assert np.allclose(x, fac.solve_A(b), rtol=1e-12, atol=1e-13 * np.abs(x).max())
    assert np.allclose(fac.pt_solve_Lt(y), x, rtol=1e-13, atol=1e-13 * np.abs(x).max())
    # L from L^-1 applied to the identity: L L^T = P Q P^T
    Linv = fac.solve_L(np.eye(g.n))
    L = np.linalg.inv(Linv)
    Pq = Qd[np.ix_(perm, perm)]
    assert np.allclose(L @ L.T, Pq, rtol=1e-9, atol=1e-11 * np.abs(Pq).max())
    assert np.allclose(np.triu(L, 1), 0, atol=1e-11 * np.abs(L).max())


def test_model_qinv_selected_inverse():
    """Model.qinv (model.py:89-118): Q^-1 on the pattern of Q and its diagonal
    against the dense inverse; after an update the conditional precision's."""
    import spdepy_amd as sp
    c = CASES["var_b1"]()
    g = c["grid"]
    m = product_case(c)
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)
    m.setQ(par=c["par"])
    mod = sp.Model.__new__(sp.Model)
    mod.grid, mod.mod, mod.Q, mod.mvar, mod.useCov = m.grid, m, None, None, False
    mod.setModel()
    step("qinv")
    mvar = mod.qinv()
    Qd = mod.Q.toarray()
    inv = np.linalg.inv(Qd)
    assert np.allclose(mvar, np.diag(inv), rtol=1e-10)
    P = mod.pQinv.tocoo()
    assert np.allclose(P.data, inv[P.row, P.col], rtol=1e-9, atol=1e-12 * np.abs(inv).max())
    assert set(zip(*mod.Q.nonzero())) == set(zip(P.row, P.col))
    step("qinv after update")
    mod.update(c["y"][:5], c["idx"][:5])
    mvar2 = mod.qinv()
    inv2 = np.linalg.inv(mod.Q.toarray())
    assert np.allclose(mvar2, np.diag(inv2), rtol=1e-10)


def test_model_fit_lowers_the_objective():
    """Model.fit (model.py:49-71): the optimiser over the GPU logLike lowers
    the normalised negative log-likelihood from a perturbed start."""
    import spdepy_amd as sp
    from spdepy_amd.optim import Optimize
    c = CASES["var_b1"]()
    m = product_case(c)
    mod = sp.Model.__new__(sp.Model)
    mod.grid, mod.mod, mod.useCov = m.grid, m, False
    mod.optim = Optimize(m.logLike)
    x0 = c["par"] + 0.2
    step("Model.fit")
    res = mod.fit(c["y"], idx=c["idx"], fitQ0=False, x0=x0.copy(), stepType="adam", lr=0.02, max_steps=12, pol=1)
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)   # fit ends with setQ, which resets S as the reference does
    f0 = m.logLike(x0, grad=False)
    f1 = m.logLike(res["x"], grad=False)
    assert np.isfinite(f1) and f1 < f0, (f0, f1)
    assert len(mod.optim.histF) == 12


@pytest.mark.parametrize("kind,bc", [("iso", 1), ("ani", 3), ("ha", 1)])
def test_spatial_whittle_matern_loglike(kind, bc):
    """Spatial Whittle-Matern logLike (whittle_matern2D.py:90-126) on the
    one-slice plan: value and gradient against the oracle (exact trace)."""
    from oracle import grid as OG
    import spdepy_amd as sp
    from spdepy_amd import spdes
    M, N = 11, 9
    g0 = OG.SpatialGrid(F.axis(M), F.axis(N), Nbs=3)
    par = np.array([0.2, np.log(0.7), -0.1]) if kind == "iso" else np.array([0.2, np.log(0.7), 0.3, -0.4, -0.1])
    par[-1] = np.log(50.0)
    wm = Mo.WhittleMatern(g0, kind, par=par.copy(), bc=bc)
    Q, dQ = wm.makeQ(par, grad=True)
    rng = np.random.default_rng(9)
    idx = np.sort(rng.choice(M * N, 40, replace=False))
    x = np.linalg.solve(np.linalg.cholesky(Q.toarray()).T, rng.normal(size=M * N))
    y = x[g0.n2e(idx)] + 0.1 * rng.normal(size=idx.size)
    like_o, jac_o, _ = Mo.loglike(Q, dQ, g0.getS(idx), y, par[-1], trace="exact")
    pg = sp.grid(x=F.axis(M), y=F.axis(N), Nbs=3)
    cls = {"iso": spdes.WhittleMatern2D, "ani": spdes.WhittleMaternAnisotropic2D, "ha": spdes.WhittleMaternHa2D}[kind]
    m = cls(grid=pg, par=par.copy(), bc=bc)
    m.initFit(y, idx=idx)
    step("spatial logLike " + kind)
    like, jac = m.logLike(par)
    assert abs(like - like_o) <= 1e-9 * abs(like_o), (like, like_o)
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * np.abs(jac_o).max()), (jac, jac_o)


@pytest.mark.parametrize("kind,bc", [("iso", 1), ("ani", 3), ("ha", 1)])
def test_var_whittle_matern_q_and_loglike(kind, bc):
    """Variable Whittle-Matern models (var_whittle_matern*2D.py): Q (pattern
    and values) and the exact-trace logLike value and gradient against the
    oracle, with A0 assembled from the fields on the one-slice plan."""
    from oracle import grid as OG
    import spdepy_amd as sp
    from spdepy_amd import spdes
    M, N = 11, 9
    g0 = OG.SpatialGrid(F.axis(M), F.axis(N), Nbs=3)
    wm = Mo.VarWhittleMatern(g0, kind, bc=bc)
    rng = np.random.default_rng(4)
    par = 0.3 * rng.normal(size=wm.Np * (2 if kind == "iso" else 4) + 1)
    par[-1] = np.log(50.0)
    Qo, dQo = wm.makeQ(par)
    idx = np.sort(rng.choice(M * N, 45, replace=False))
    x = np.linalg.solve(np.linalg.cholesky(Qo.toarray()).T, rng.normal(size=M * N))
    y = x[g0.n2e(idx)] + 0.1 * rng.normal(size=idx.size)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, g0.getS(idx), y, par[-1], trace="exact")
    pg = sp.grid(x=F.axis(M), y=F.axis(N), Nbs=3)
    cls = {"iso": spdes.VarWhittleMatern2D, "ani": spdes.VarWhittleMaternAnisotropic2D,
           "ha": spdes.VarWhittleMaternHa2D}[kind]
    m = cls(grid=pg, par=par.copy(), bc=bc)
    step("var Whittle-Matern Q " + kind)
    Q = m.Q.tocsc()
    Qo = Qo.tocsc()
    Qo.sort_indices(); Q.sort_indices()
    assert np.array_equal(Q.indptr, Qo.indptr) and np.array_equal(Q.indices, Qo.indices)
    assert np.allclose(Q.data, Qo.data, rtol=1e-12, atol=1e-12 * abs(Qo.data).max())
    m.initFit(y, idx=idx)
    step("var Whittle-Matern logLike " + kind)
    like, jac = m.logLike(par)
    assert abs(like - like_o) <= 1e-9 * abs(like_o), (like, like_o)
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * np.abs(jac_o).max())


FIELD_CASES = [(k, dv, gv, 3) for k in ("iso", "ani", "ha") for dv, gv in ((True, False), (False, True), (True, True))]
FIELD_CASES += [("iso", True, False, 1), ("ha", False, True, 1), ("ani", True, True, 1)]


@pytest.mark.parametrize("kind,dvar,gvar,bc", FIELD_CASES)
def test_field_family_q_and_loglike(kind, dvar, gvar, bc):
    """Families with some coefficients as B-spline fields and the rest
    constant (advection_var_*, var_advection_*, var_advection_var_*): Q
    (pattern and values) and the exact-trace logLike value and gradient
    against the oracle."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    c = F.field_model(9, 8, 4, kind=kind, dvar=dvar, gvar=gvar, bc=bc, m=45)
    g, par = c["grid"], c["par"]
    Qo, dQo = c["model"].makeQ(par, grad=True)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, g.getS(c["idx"]), c["y"], par[-1], trace="exact")
    pg = sp.grid(x=g.x, y=g.y, t=g.t, Nbs=3)
    mod0 = spdes.WhittleMatern2D(grid=sp.grid(x=g.x, y=g.y, Nbs=3), par=np.array(c["mod0"].par), bc=bc)
    spde = {(True, False): "advection-var-diffusion", (False, True): "var-advection-diffusion",
            (True, True): "var-advection-var-diffusion"}[(dvar, gvar)]
    m = spdes.spde_init(spde, pg, ani=kind != "iso", ha=kind == "ha", bc=bc, mod0=mod0)
    step("field family Q %s dvar=%d gvar=%d bc=%d" % (kind, dvar, gvar, bc))
    Qg, _, _ = m.makeQ(par, grad=False)
    Qo = Qo.tocsc()
    Qo.sort_indices()
    Qg.sort_indices()
    assert np.array_equal(Qg.indptr, Qo.indptr) and np.array_equal(Qg.indices, Qo.indices)
    assert np.allclose(Qg.data, Qo.data, rtol=1e-12, atol=1e-13 * np.abs(Qo.data).max())
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)
    step("field family logLike %s dvar=%d gvar=%d bc=%d" % (kind, dvar, gvar, bc))
    like, jac = m.logLike(par)
    assert abs(like - like_o) <= 1e-9 * abs(like_o), (like, like_o)
    scale = np.abs(jac_o).max()
    assert jac.shape == jac_o.shape
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale


@pytest.mark.parametrize("kind,T,bc", [("ani", 5, 3), ("ani", 4, 1), ("iso", 10, 3), ("ha", 10, 1)])
def test_separable_q_and_loglike(kind, T, bc):
    """Separable models Q = Qt (x) Qs (seperable_spatial_temporal*2D.py) on the
    wide-time plan: Q (pattern and values) and the exact-trace logLike value
    and gradient against the oracle (dense inverses)."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    from oracle import grid as OG
    M, N = 8, 7
    g = OG.STGrid(F.axis(M), F.axis(N), np.arange(float(T)), Nbs=3)
    om = Mo.Separable(g, kind, bc=bc)
    rng = np.random.default_rng(12)
    tp = [np.log(0.6)] if kind == "ani" else [0.5, np.log(2.0)]
    par = np.hstack([0.3 * rng.normal(size=om.ns), tp, [np.log(50.0)]])
    Qo, dQo = om.makeQ(par)
    L = np.linalg.cholesky(Qo.toarray())
    x = np.linalg.solve(L.T, rng.normal(size=(g.n, 2)))
    idx = np.sort(rng.choice(g.n, 60, replace=False))
    y = x[g.n2e(idx)] + 0.1 * rng.normal(size=(60, 2))
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, g.getS(idx), y, par[-1], trace="exact")
    pg = sp.grid(x=g.x, y=g.y, t=g.t, Nbs=3)
    m = spdes.spde_init("seperable-spatial-temporal", pg, ani=kind != "iso", ha=kind == "ha", bc=bc)
    step("separable Q %s T=%d bc=%d" % (kind, T, bc))
    Qg, fac, _ = m.makeQ(par, grad=False)
    Qo = Qo.tocsc()
    Qo.sort_indices()
    Qg.sort_indices()
    assert np.array_equal(Qg.indptr, Qo.indptr) and np.array_equal(Qg.indices, Qo.indices)
    assert np.allclose(Qg.data, Qo.data, rtol=1e-12, atol=1e-13 * np.abs(Qo.data).max())
    step("separable factor %s" % kind)
    ref = np.linalg.slogdet(Qo.toarray())[1]
    assert abs(fac.logdet() - ref) <= 1e-11 * abs(ref)
    m.initFit(y, idx=idx)
    step("separable logLike %s T=%d bc=%d" % (kind, T, bc))
    like, jac = m.logLike(par)
    assert abs(like - like_o) <= 1e-9 * abs(like_o), (like, like_o)
    scale = np.abs(jac_o).max()
    assert jac.shape == jac_o.shape
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale


def test_library_first_then_torch_in_a_fresh_process():
    """A script whose first GPU work is a native plan (no torch call before
    it) still gets a working torch device afterwards."""
    import subprocess
    import sys
    code = ("import numpy as np\n"
            "from spdepy_amd.plan import Plan\n"
            "p = Plan(6, 5, 3, 3, nslots=1, wide=True)\n"
            "import torch\n"
            "x = torch.ones(4, device='cuda')\n"
            "print(float(x.sum()))\n")
    step("fresh process: plan before torch")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=180,
                         cwd=os.path.dirname(os.path.dirname(__file__)))
    assert res.returncode == 0, res.stderr[-2000:]
    assert res.stdout.strip().endswith("4.0")


@pytest.mark.parametrize("kind,dvar,bc,lam", [("iso", False, 3, 0.8), ("ani", False, 1, 0.6), ("ha", False, 3, -0.5),
                                              ("iso", True, 1, 0.7), ("ha", True, 3, 0.9)])
def test_cov_family_q_and_loglike(kind, dvar, bc, lam):
    """Cov-advection families (advection lambda * ww from assimilate_adv of
    synthetic currents; SINMOD currents are not available): Q and the
    exact-trace logLike value and gradient against the oracle, whose lambda
    derivative uses the stencil of ww as the reference does (also for
    lambda < 0, where it is not the derivative)."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    c = F.field_model(9, 8, 4, kind=kind, dvar=dvar, gvar=False, bc=bc, m=45, cov=lam)
    g, par = c["grid"], c["par"]
    Qo, dQo = c["model"].makeQ(par, grad=True)
    like_o, jac_o, _ = Mo.loglike(Qo, dQo, g.getS(c["idx"]), c["y"], par[-1], trace="exact")
    pg = sp.grid(x=g.x, y=g.y, t=g.t, Nbs=3)
    mod0 = spdes.WhittleMatern2D(grid=sp.grid(x=g.x, y=g.y, Nbs=3), par=np.array(c["mod0"].par), bc=bc)
    spde = "cov-advection-var-diffusion" if dvar else "cov-advection-diffusion"
    m = spdes.spde_init(spde, pg, ani=kind != "iso", ha=kind == "ha", bc=bc, mod0=mod0)
    X, Y = np.meshgrid(g.x, g.y)
    ww = pg.assimilate_adv(*F.currents(g))
    assert np.array_equal(ww, c["ww"])
    m.initFit(c["y"], idx=c["idx"], fitQ0=False, ww=ww)
    step("cov family Q %s dvar=%d bc=%d" % (kind, dvar, bc))
    Qg, _, _ = m.makeQ(par, grad=False)
    Qo = Qo.tocsc()
    Qo.sort_indices()
    Qg.sort_indices()
    assert np.array_equal(Qg.indptr, Qo.indptr) and np.array_equal(Qg.indices, Qo.indices)
    assert np.allclose(Qg.data, Qo.data, rtol=1e-12, atol=1e-13 * np.abs(Qo.data).max())
    step("cov family logLike %s dvar=%d bc=%d" % (kind, dvar, bc))
    like, jac = m.logLike(par)
    assert abs(like - like_o) <= 1e-9 * abs(like_o), (like, like_o)
    scale = np.abs(jac_o).max()
    assert jac.shape == jac_o.shape
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale


@pytest.mark.parametrize("family", ["separable", "field", "cov"])
def test_model_sample_and_update_other_families(family):
    """Model.setModel/sample/update (model.py:73-127) on the separable, field
    and cov-advection families: L^T P x = z for the samples and
    mu = tau Q_c^-1 S^T y against dense algebra on the oracle's Q."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    from oracle import grid as OG
    rng = np.random.default_rng(21)
    if family == "separable":
        g = OG.STGrid(F.axis(8), F.axis(7), np.arange(5.0), Nbs=3)
        om = Mo.Separable(g, "ani", bc=3)
        par = np.hstack([0.3 * rng.normal(size=om.ns), [np.log(0.6)], [np.log(50.0)]])
        pg = sp.grid(x=g.x, y=g.y, t=g.t, Nbs=3)
        m = spdes.SeperableSpatialTemporal2D(grid=pg, par=par.copy(), bc=3)
        Qd = om.makeQ(par, grad=False)[0].toarray()
    else:
        c = F.field_model(8, 7, 4, kind="ani", dvar=True, gvar=family == "field", bc=3, m=30,
                          cov=0.7 if family == "cov" else None)
        g, par = c["grid"], c["par"]
        pg = sp.grid(x=g.x, y=g.y, t=g.t, Nbs=3)
        mod0 = spdes.WhittleMatern2D(grid=sp.grid(x=g.x, y=g.y, Nbs=3), par=np.array(c["mod0"].par), bc=3)
        spde = "cov-advection-var-diffusion" if family == "cov" else "var-advection-var-diffusion"
        m = spdes.spde_init(spde, pg, ani=True, ha=False, bc=3, mod0=mod0)
        if family == "cov":
            m.ww = c["ww"]
        m.setQ(par=par.copy())
        Qd = c["model"].makeQ(par, grad=False)[0].toarray()
    M = sp.Model.__new__(sp.Model)
    M.grid, M.mod, M.useCov = pg, m, False
    M.mu = np.zeros(g.n)
    M.setModel()
    step("Model.sample " + family)
    x = M.sample(n=2, simple=True, seed=3)
    fac = M.Q_fac
    back = fac.pt_solve_Lt(fac.solve_L_p(Qd @ x))
    assert np.allclose(np.asarray(back), x, rtol=1e-8, atol=1e-10)
    step("Model.update " + family)
    idx = np.sort(rng.choice(g.n, 40, replace=False))
    y = rng.normal(size=40)
    tau = np.exp(par[-1])
    M.update(y, idx, tau=tau)
    S = g.getS(idx)
    mu_ref = np.linalg.solve(Qd + tau * (S.T @ S).toarray(), tau * (S.T @ y))
    assert np.allclose(M.mu, mu_ref, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("case,r", [("var_b1", 1), ("iso_b1", 2), ("ani_b3", 1)])
def test_fused_entry_point_matches_the_engine(case, r):
    """spde_loglik_grad (one C call) against the Python-orchestrated engine
    path: the same value, log-determinants, per-cell adjoints and sigma/tau
    gradient entries; the model's gradient from the fused cells equals the
    oracle's."""
    from spdepy_amd import engine as E
    c = CASES[case]()
    m = product_case(c)
    y = c["y"] if r == 1 else np.stack([c["y"], 0.5 * c["y"] + 0.01], axis=1)
    m.initFit(y, idx=c["idx"], fitQ0=False)
    par = c["par"]
    eng = m._engine()
    coeffs = m._coeffs(par)
    step("engine path " + case)
    like_e, parts_e = eng.loglik(coeffs, m._obs_cols(), m.data, par[-1], grad=True)
    step("fused entry point " + case)
    like_f, parts_f = E.loglik_fused(eng, coeffs, m._obs_cols(), m.data, par[-1])
    assert abs(like_f - like_e) <= 1e-12 * abs(like_e)
    assert abs(parts_f["ld_qc"] - parts_e["ld_qc"]) <= 1e-12 * abs(parts_e["ld_qc"])
    assert abs(parts_f["ld_q"] - parts_e["ld_q"]) <= 1e-12 * abs(parts_e["ld_q"])
    scale = np.abs(parts_e["cells"]).max()
    assert np.allclose(parts_f["cells"], parts_e["cells"], rtol=1e-11, atol=1e-12 * scale)
    for k in ("g_sigma", "g_tau"):
        assert abs(parts_f[k] - parts_e[k]) <= 1e-10 * max(1.0, abs(parts_e[k])), k
    if r == 1:
        import torch
        contract = m._device_contract()
        if contract is not None:
            parts_f["coef_grad"], parts_f["colsum"] = contract(torch.as_tensor(parts_f["cells"].ravel()).cuda())
        g = m._contract(par, coeffs, parts_f)
        jac = -g / (parts_f["m"] * parts_f["r"])
        S = c["grid"].getS(c["idx"])
        Qo, dQo = c["model"].makeQ(par, grad=True)
        like_o, jac_o, _ = Mo.loglike(Qo, dQo, S, c["y"], par[-1], trace="exact")
        assert abs(like_f - like_o) <= 1e-9 * abs(like_o)
        assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * np.abs(jac_o).max())
