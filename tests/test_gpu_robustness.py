"""Failure paths and trajectories on the GPU: an indefinite precision is
reported with its lattice index (the reference raises CHOLMOD's
not-positive-definite error), non-finite parameters or values raise instead
of returning numbers, and 12 Adam steps of Model.fit over the GPU logLike
follow the oracle's iterates (optim/__init__.py:26-62, adam.py:12-26)."""
import numpy as np
import pytest

from oracle import fixtures as F, models as Mo, optim as O
from tests.gpu_util import product_case, step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    c = F.small(9, 8, 4, bc=1, m=40)
    m = product_case(c)
    m.initFit(c["y"], idx=c["idx"], fitQ0=False)
    return c, m


def test_indefinite_q_reports_the_lattice_index(small):
    from spdepy_amd import _native as nat
    c, m = small
    eng = m._engine()
    qval = eng.assemble(m._coeffs(c["par"]))
    e = eng.plan.export()
    k = 137
    row = np.arange(e["indptr"][k], e["indptr"][k + 1])
    pos = int(row[e["indices"][row] == k][0])
    bad = qval.clone()
    bad[pos] = -1e3 * abs(float(qval[pos]))
    step("factor an indefinite Q")
    with pytest.raises(nat.NotPositiveDefinite) as err:
        eng.plan.factor(1, bad)
    assert "lattice index %d" % k in str(err.value)
    # the plan stays usable: the valid matrix factors afterwards
    ld = eng.plan.factor(1, qval)
    assert np.isfinite(ld)


def test_non_finite_parameters_and_values_raise(small):
    from spdepy_amd import _native as nat
    c, m = small
    par = c["par"].copy()
    par[3] = np.nan
    step("NaN parameters")
    with pytest.raises(ValueError):
        m.logLike(par)
    eng = m._engine()
    qval = eng.assemble(m._coeffs(c["par"]))
    e = eng.plan.export()
    k = 40
    row = np.arange(e["indptr"][k], e["indptr"][k + 1])
    qn = qval.clone()
    qn[int(row[e["indices"][row] == k][0])] = float("nan")   # the diagonal: read whatever the ordering
    step("NaN values reach the factorisation")
    with pytest.raises(nat.NotPositiveDefinite):
        eng.plan.factor(1, qn)


def test_adam_trajectory_matches_oracle(small):
    from spdepy_amd.optim import Optimize
    c, m = small
    g = c["grid"]
    S = g.getS(c["idx"])

    def oracle_fun(x):
        Q, dQ = c["model"].makeQ(x, grad=True)
        like, jac, _ = Mo.loglike(Q, dQ, S, c["y"], x[-1], trace="exact")
        return like, jac
    x0 = c["par"] + 0.1
    step("12 Adam steps over the GPU logLike")
    opt = Optimize(m.logLike)
    res = opt.fit(x0=x0.copy(), stepType="adam", lr=0.02, max_steps=12, pol=4)
    xs_o, fs_o, _, xp_o = O.adam_fit(oracle_fun, x0, 0.02, 12, pol=4)
    xs = np.array(opt.histX)
    scale = np.abs(xs_o).max()
    print("  max iterate difference / scale %.2e, objective %.2e" % (np.abs(xs - xs_o).max() / scale,
                                                                   np.abs(np.array(opt.histF) - fs_o).max()))
    assert np.allclose(np.array(opt.histF), fs_o, rtol=1e-9, atol=0)
    assert np.allclose(xs, xs_o, rtol=0, atol=1e-8 * scale)
    assert np.allclose(res["x"], xp_o, rtol=0, atol=1e-8 * scale)
