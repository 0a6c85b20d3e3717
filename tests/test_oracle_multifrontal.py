"""The oracle's CPU multifrontal Cholesky (the CPU baseline's CHOLMOD
stand-in) against dense log-determinants and the closed form."""
import numpy as np
import pytest

from oracle import fixtures as F, models as Mo, multifrontal as MF


@pytest.mark.parametrize("M,N,T,leaf", [(9, 8, 4, 16), (12, 10, 5, 40), (7, 7, 3, 1000)])
def test_multifrontal_logdet(M, N, T, leaf):
    c = F.small(M, N, T, bc=1)
    Q, _ = c["model"].makeQ(c["par"], grad=False)
    mf = MF.Multifrontal(Q, M, N, T, leaf=leaf)
    ld = mf.factor(Q)
    ref = np.linalg.slogdet(Q.toarray())[1]
    assert abs(ld - ref) <= 1e-11 * abs(ref)


def test_multifrontal_matches_closed_form_30x30x10():
    c = F.var_model(30, 30, 10, Nbs=8, m=50)
    Q, _ = c["model"].makeQ(c["par"], grad=False)
    ld = MF.Multifrontal(Q, 30, 30, 10).factor(Q)
    A, q, s = c["model"].spatial_operator(c["par"])
    A0 = c["mod0"].A0(c["mod0"].par)[0]
    cf = Mo.logdet_closed_form(A, q, A0, c["grid"].V, 10, s, c["grid"].dt)
    assert abs(ld - cf) <= 1e-11 * abs(cf)
