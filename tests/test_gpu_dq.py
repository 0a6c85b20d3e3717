"""makeQ(par, grad=True): the explicit dQ_i list (spde_assemble_tangent
through the C ABI) against the oracle's dQ list -- the reference's block
formula per parameter, pinned by central differences of Q in
tests/test_oracle_models.py -- for the B-spline model and the stationary
models, with and without initial-state parameters, bc 1, 2, 3."""
import numpy as np
import pytest

from oracle import fixtures as F
from tests.gpu_util import product_case, step

pytestmark = pytest.mark.gpu


def _compare(m, c, par, label):
    Q_o, dQ_o = c["model"].makeQ(par, grad=True)
    step("dQ %s" % label)
    Q, _, dQ = m.makeQ(par, grad=True)
    assert len(dQ) == len(dQ_o) == par.size - 1
    worst = 0.0
    for i in range(len(dQ)):
        a = dQ[i].toarray()
        b = dQ_o[i].toarray()
        scale = np.abs(b).max()
        err = np.abs(a - b).max() / scale
        worst = max(worst, err)
        assert err <= 1e-12, (label, i, err)
    print("  %s: %d matrices, max error / scale %.2e" % (label, len(dQ), worst))


@pytest.mark.parametrize("bc", [1, 3])
def test_var_model_dq(bc):
    c = F.small(8, 7, 4, Nbs=3, bc=bc, m=30)
    m = product_case(c)
    _compare(m, c, c["par"], "var bc=%d" % bc)


@pytest.mark.parametrize("kind,bc", [("iso", 1), ("ani", 2), ("ha", 3), ("ani", 1)])
def test_stationary_model_dq(kind, bc):
    M, N = (9, 9) if bc == 2 else (9, 8)
    c = F.stationary(M=M, N=N, T=4, bc=bc, kind=kind, m=30)
    m = product_case(c, kind=kind)
    _compare(m, c, c["par"], "%s bc=%d" % (kind, bc))


def test_dq_with_initial_state_parameters():
    c = F.stationary(M=9, N=8, T=4, bc=1, kind="ani", m=30)
    m = product_case(c, kind="ani")
    par = np.hstack([c["par"][:-1], c["mod0"].par[:-1], c["par"][-1:]])
    _compare(m, c, par, "ani with mod0")
