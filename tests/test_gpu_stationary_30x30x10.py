"""Parity at the CPU-runnable configuration (stationary isotropic
advection-diffusion, 30 x 30 x 10, n = 9,000, bc = 1, 200 observations):
the drop-in model's Q, logLike and exact gradient against the oracle's
values in tests/golden/stationary_30x30x10.npz (written by tests/golden/make_stationary_30x30x10.py;
the oracle's exact trace there uses dense inverses)."""
import os

import numpy as np
import pytest

from oracle import fixtures as F
from tests.gpu_util import product_case, step

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stationary_30x30x10.npz"))


def test_q_loglike_gradient_30x30x10():
    c = F.stationary()
    assert np.array_equal(c["idx"], GOLD["idx"]) and np.array_equal(c["y"], GOLD["y"])
    par = GOLD["par"]
    m = product_case(c)
    step("30x30x10 makeQ (n = 9,000)")
    Q, _, _ = m.makeQ(par, grad=False)
    Q = Q.tocsr()
    Q.sort_indices()
    assert Q.nnz == int(GOLD["q_nnz"])
    # pattern bit-exact: the reference's scipy pattern (exact zeros dropped)
    assert np.array_equal(Q.indptr, GOLD["q_indptr"])
    assert np.array_equal(Q.indices, GOLD["q_indices"])
    assert np.allclose(Q.diagonal(), GOLD["q_diag"], rtol=1e-12, atol=0)
    qs = float(GOLD["q_abs_sum"])
    assert abs(np.abs(Q.data).sum() - qs) <= 1e-12 * qs
    m.initFit(GOLD["y"], idx=GOLD["idx"], fitQ0=False)
    step("30x30x10 logLike exact")
    like, jac = m.logLike(par)
    like_o, jac_o = float(GOLD["like"]), GOLD["jac"]
    assert abs(like - like_o) <= 1e-9 * abs(like_o)
    assert jac.shape == jac_o.shape
    scale = np.abs(jac_o).max()
    assert np.allclose(jac, jac_o, rtol=1e-9, atol=1e-9 * scale), np.abs(jac - jac_o).max() / scale
