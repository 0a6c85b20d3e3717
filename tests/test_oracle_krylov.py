"""The iterative-path oracle (oracle/krylov.py) against what mathematics fixes:
the IC(0) identity, the independent-set property of the colouring, the
published outputs of splitmix64, exact Gauss quadrature with full Lanczos,
and the Woodbury form of the exact log-likelihood against the dense one."""
import numpy as np
import pytest
from scipy import sparse

from oracle import fixtures as F, krylov as K, models as Mo


def _qc(M=9, N=8, T=4, m=30):
    c = F.small(M, N, T, bc=1, m=m)
    g = c["grid"]
    Q, _ = c["model"].makeQ(c["par"], grad=False)
    S = g.getS(c["idx"])
    return (Q + np.exp(c["par"][-1]) * (S.T @ S)).tocsr(), c


def test_splitmix64_published_outputs():
    # first outputs of Vigna's splitmix64 generator seeded with 0
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        out = K.splitmix64(np.array([0, g], dtype=np.uint64))
    assert int(out[0]) == 0xE220A8397B1DCDAF
    assert int(out[1]) == 0x6E789E6AA1B965F4


def test_rademacher_signs_balanced_and_seeded():
    v = K.rademacher(4000, 8, seed=7)
    assert set(np.unique(v)) == {-1.0, 1.0}
    assert abs(v.mean()) < 0.03
    assert np.array_equal(v, K.rademacher(4000, 8, seed=7))
    assert not np.array_equal(v, K.rademacher(4000, 8, seed=8))


def test_gaussian_draws():
    """Box-Muller on the splitmix64 uniforms: the uniforms lie in (0, 1) and
    are uniform (first two moments), the draws have the normal moments and
    the Box-Muller radius identity z^2 + z'^2 = -2 log U1 holds with the
    sine partner."""
    u1, u2 = K.uniforms(20000, 4, seed=11)
    assert u1.min() > 0.0 and u1.max() < 1.0 and u2.min() >= 0.0 and u2.max() < 1.0
    assert abs(u1.mean() - 0.5) < 0.01 and abs(u1.var() - 1 / 12) < 0.003
    z = K.gaussian(20000, 4, seed=11)
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.03
    assert abs((z ** 4).mean() - 3.0) < 0.15
    zs = np.sqrt(-2.0 * np.log(u1)) * np.sin(2.0 * np.pi * u2)
    assert np.allclose(z ** 2 + zs ** 2, -2.0 * np.log(u1), rtol=1e-12)
    assert np.array_equal(z, K.gaussian(20000, 4, seed=11))
    assert not np.array_equal(z, K.gaussian(20000, 4, seed=12))


@pytest.mark.parametrize("bc", [1, 3])
def test_fform_square_root_reproduces_q(bc):
    """R = F^T D^(1/2) (applied to the identity) gives R R^T = Q, the
    precision makeQ assembles by the reference's block formula."""
    c = F.small(7, 6, 4, bc=bc, m=20)
    g, model, par = c["grid"], c["model"], c["par"]
    Q, _ = model.makeQ(par, grad=False)
    A, q, sigma = model.spatial_operator(par)
    A0 = c["mod0"].A0(c["mod0"].par)[0]
    M, N, T = g.shape
    R = K.fform_t(A, q, A0, g.V, T, sigma, g.dt, np.eye(g.n))
    Qd = Q.toarray()
    assert np.allclose(R @ R.T, Qd, rtol=1e-12, atol=1e-12 * np.abs(Qd).max())


def test_colouring_is_an_independent_set_partition():
    A, c = _qc()
    M, N, T = c["grid"].shape
    col = K.colours(M, N, T)
    Ac = A.tocoo()
    off = Ac.row != Ac.col
    assert not np.any(col[Ac.row[off]] == col[Ac.col[off]])
    assert set(np.unique(col)) == set(range(18))


@pytest.mark.parametrize("shift", [0.1, 0.3])
def test_ic0_identity_on_the_pattern(shift):
    """(F F^T)_ij = A_ij on the pattern of A off the diagonal and
    (1 + shift) A_ii on it: the defining property of IC(0)."""
    A, c = _qc()
    M, N, T = c["grid"].shape
    col = K.colours(M, N, T)
    L, d = K.ic0(A, col, shift)
    Fm = K.full_factor(L, d)
    # F is lower triangular in the colour ordering
    Fc = Fm.tocoo()
    assert np.all((col[Fc.row] > col[Fc.col]) | (Fc.row == Fc.col))
    P = (Fm @ Fm.T).tocsr()
    Ac = A.tocoo()
    prod = np.asarray(P[Ac.row, Ac.col]).ravel()
    want = np.where(Ac.row == Ac.col, (1 + shift) * Ac.data, Ac.data)
    assert np.allclose(prod, want, rtol=1e-12, atol=1e-12 * np.abs(A.data).max())


def test_ic_preconditioned_cg_solves():
    A, c = _qc()
    M, N, T = c["grid"].shape
    col = K.colours(M, N, T)
    L, d = K.ic0(A, col, 0.1)
    b = np.random.default_rng(0).normal(size=A.shape[0])
    x, it = K.pcg(A, b, K.ic_solver(L, d, col), tol=1e-10)
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)
    _, it_j = K.pcg(A, b, lambda r: r / A.diagonal(), tol=1e-10)
    assert it < it_j


def test_slq_full_lanczos_is_exact_per_probe():
    """With as many steps as the dimension, Gauss quadrature gives
    v^T log(A) v exactly for each probe."""
    rng = np.random.default_rng(1)
    n = 40
    B = rng.normal(size=(n, n))
    A = B @ B.T + n * np.eye(n)
    w, U = np.linalg.eigh(A)
    logA = (U * np.log(w)) @ U.T
    V = K.rademacher(n, 4, seed=3)
    al, be = [], []
    for r in range(4):
        a, b = K.lanczos(lambda x: A @ x, V[:, r] / np.sqrt(n), n)
        al.append(a)
        be.append(b)
    mean, se, vals = K.slq_logdet(al, be, n)
    exact = [V[:, r] @ logA @ V[:, r] for r in range(4)]
    assert np.allclose(vals, exact, rtol=1e-8)
    assert abs(mean - np.mean(exact)) <= 1e-8 * abs(np.mean(exact))


def test_slq_estimate_of_a_preconditioned_logdet():
    """log|Q_c| = log|F F^T| + tr log(F^-1 Q_c F^-T): the estimate is within
    four standard errors of the exact value."""
    A, c = _qc()
    M, N, T = c["grid"].shape
    col = K.colours(M, N, T)
    L, d = K.ic0(A, col, 0.1)
    n = A.shape[0]
    order = np.lexsort((np.arange(n), col))
    Pm = sparse.csr_matrix((np.ones(n), (np.arange(n), order)), shape=(n, n))
    Fm = (Pm @ K.full_factor(L, d) @ Pm.T).toarray()
    Ap = (Pm @ A @ Pm.T).toarray()
    Bm = np.linalg.solve(Fm, np.linalg.solve(Fm, Ap).T)
    V = K.rademacher(n, 32, seed=11)
    al, be = zip(*[K.lanczos(lambda x: Bm @ x, V[:, r] / np.sqrt(n), 30) for r in range(32)])
    est, se, _ = K.slq_logdet(al, be, n)
    exact = np.linalg.slogdet(Ap)[1] - 2 * np.log(d).sum()
    assert abs(est - exact) <= 4 * se + 1e-6 * abs(exact), (est, exact, se)


@pytest.mark.parametrize("M,N,T,bc", [(8, 7, 4, 1), (9, 8, 5, 3)])
def test_woodbury_loglike_equals_dense_exact(M, N, T, bc):
    c = F.small(M, N, T, bc=bc, m=40)
    g, model, par = c["grid"], c["model"], c["par"]
    Q, dQ = model.makeQ(par, grad=True)
    l1, j1, _ = Mo.loglike(Q, dQ, g.getS(c["idx"]), c["y"], par[-1], "exact")
    A, q, s = model.spatial_operator(par)
    A0 = c["mod0"].A0(c["mod0"].par)[0]
    l2, j2, _ = Mo.loglike_woodbury(Q, model.dQ_iter(par), A, q, A0, g.V, T, s, g.dt, M, N, g.n2e(c["idx"]),
                                    c["y"], par[-1])
    assert abs(l1 - l2) <= 1e-12 * abs(l1)
    assert np.allclose(j1, j2, rtol=1e-10, atol=1e-11 * np.abs(j1).max())
