"""The learned-preconditioner oracle (oracle/learned.py) against what
mathematics fixes: bfloat16 round-to-nearest-even, the network's backward
pass against central differences of its float64 forward pass, the KL loss
gradient against central differences, and the pair channel layout."""
import numpy as np
from scipy import sparse

from oracle import fixtures as F, krylov as K, learned as Le


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e-3, 1 + 2 ** -7 + 2 ** -9], np.float32)
    got = Le.bf16(vals)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, -2.5, None, 1 + 2 ** -7], dtype=object)
    for g, w in zip(got, want):
        if w is not None:
            assert g == np.float32(w)
    # exactly representable: unchanged; mantissa beyond 7 bits cleared
    assert Le.bf16(np.float32(0.0078125)) == np.float32(0.0078125)
    assert (Le.bf16(np.random.default_rng(0).normal(size=100).astype(np.float32)).view(np.uint32) & 0xFFFF == 0).all()


def _net(shape, seed):
    rng = np.random.default_rng(seed)
    th = []
    for cin, _ in Le.LAYERS:
        th.append(rng.normal(scale=np.sqrt(2 / (9 * cin)), size=9 * cin * Le.NCH))
        th.append(rng.normal(scale=0.1, size=Le.NCH))
    P = int(np.prod(shape))
    return np.concatenate(th), rng.normal(size=(P, 8))


def test_network_backward_matches_central_differences():
    shape = (5, 4, 2)
    th, feat = _net(shape, 1)
    dout = np.random.default_rng(2).normal(size=(feat.shape[0], Le.NPAIR))

    def loss(t):
        return float((Le.forward(t, feat, shape, rounding=False)[-1][:, :Le.NPAIR] * dout).sum())
    acts = Le.forward(th, feat, shape, rounding=False)
    g = Le.backward(th, acts, dout, shape, rounding=False)
    rng = np.random.default_rng(3)
    for e in rng.choice(th.size, 40, replace=False):
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[e] += h
        tm[e] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * max(1.0, abs(g[e])), (e, fd, g[e])


def test_conv_data_gradient_matches_central_differences():
    shape = (4, 3, 2)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(24, 32))
    W = rng.normal(size=(9 * 32, 32))
    G = rng.normal(size=(24, 32))
    Y = Le.conv(X, W, None, shape, True, rounding=False)
    dX, dW, db = Le.conv_backward(X, W, G, Y, shape, rounding=False)
    for (p, c) in [(0, 0), (5, 7), (23, 31), (11, 16)]:
        h = 1e-6
        Xp, Xm = X.copy(), X.copy()
        Xp[p, c] += h
        Xm[p, c] -= h
        fd = ((Le.conv(Xp, W, None, shape, True, False) - Le.conv(Xm, W, None, shape, True, False)) * G).sum() / (2 * h)
        assert abs(fd - dX[p, c]) <= 1e-6 * max(1.0, abs(fd))


def _factor_case():
    c = F.small(6, 5, 4, bc=1, m=20)
    g = c["grid"]
    Q, _ = c["model"].makeQ(c["par"], grad=False)
    S = g.getS(c["idx"])
    A = (Q + np.exp(c["par"][-1]) * (S.T @ S)).tocsr()
    M, N, T = g.shape
    col = K.colours(M, N, T)
    L, d = K.ic0(A, col, 0.1)
    L = L.tocsr()
    L.sort_indices()
    lptr, lcol, lower = L.indptr.astype(np.int64), L.indices, L.data
    rows = np.repeat(np.arange(A.shape[0]), np.diff(lptr))
    lch = Le.pair_channels(M, N, T, rows, lcol)
    return A, lower, d, lptr, lcol, lch, (M, N, T)


def test_pair_channels_cover_each_pair_once():
    A, lower, d, lptr, lcol, lch, (M, N, T) = _factor_case()
    ok = lch[lch >= 0]
    assert np.unique(ok).size == ok.size                 # one entry per (cell, channel)
    assert np.all(lch >= 0)                              # every pair of the pattern has a channel
    assert np.all(ok % Le.NPAIR >= 1)                    # channel 0 is the diagonal's


def test_kl_gradient_matches_central_differences():
    A, lower, d, lptr, lcol, lch, shape = _factor_case()
    n = A.shape[0]
    x = np.random.default_rng(5).normal(size=(n, 8)) * 0.05
    adiag = A.diagonal()
    corr = np.random.default_rng(6).normal(scale=0.01, size=(n, Le.NPAIR))

    def loss(c):
        lo, dg = Le.corrected_factor(lower, d, lptr, lcol, lch, adiag, c)
        return Le.kl_grad(lo, dg, lptr, lcol, lch, adiag, x)[0]
    lo, dg = Le.corrected_factor(lower, d, lptr, lcol, lch, adiag, corr)
    _, dout = Le.kl_grad(lo, dg, lptr, lcol, lch, adiag, x)
    rng = np.random.default_rng(7)
    used = np.concatenate([lch[lch >= 0], np.arange(n) * Le.NPAIR])
    for e in rng.choice(used, 30, replace=False):
        i, ch = divmod(int(e), Le.NPAIR)
        h = 1e-6
        cp, cm = corr.copy(), corr.copy()
        cp[i, ch] += h
        cm[i, ch] -= h
        fd = (loss(cp) - loss(cm)) / (2 * h)
        assert abs(fd - dout[i, ch]) <= 1e-5 * max(1.0, abs(fd)), (i, ch, fd, dout[i, ch])


def test_kl_loss_attains_its_bound_at_the_exact_factor():
    """With samples whose second moment is exactly Sigma = A^-1, the loss
    1/2 tr(F^T Sigma F) - sum log F_ii is >= n/2 + 1/2 log|Sigma| for every
    lower-triangular F, with equality at the exact Cholesky factor of A (in
    the colour ordering, full pattern); the IC(0) factor lies above it."""
    A, lower, d, lptr, lcol, lch, (M, N, T) = _factor_case()
    n = A.shape[0]
    Ad = A.toarray()
    Sig = np.linalg.inv(Ad)
    x = np.sqrt(n) * np.linalg.cholesky(Sig)   # (1/n) x x^T = Sigma
    bound = n / 2 + 0.5 * np.linalg.slogdet(Sig)[1]
    col = K.colours(M, N, T)
    order = np.lexsort((np.arange(n), col))
    rank = np.empty(n, int)
    rank[order] = np.arange(n)
    Lo = np.linalg.cholesky(Ad[np.ix_(order, order)])
    rows, cols, vals = [], [], []
    for a in range(n):
        for b in range(n):
            if rank[b] < rank[a] and Lo[rank[a], rank[b]] != 0.0:
                rows.append(a)
                cols.append(b)
                vals.append(Lo[rank[a], rank[b]])
    Lf = sparse.csr_matrix((vals, (rows, cols)), shape=(n, n))
    Lf.sort_indices()
    full_diag = np.diag(Lo)[rank]
    none = np.full(Lf.nnz, -1, np.int64)
    loss_exact = Le.kl_grad(Lf.data, full_diag, Lf.indptr.astype(np.int64), Lf.indices, none, Ad.diagonal(), x)[0]
    assert abs(loss_exact - bound) <= 1e-9 * abs(bound)
    loss_ic = Le.kl_grad(lower, d, lptr, lcol, lch, A.diagonal(), x)[0]
    assert loss_ic > bound


def test_jeffreys_gradient_matches_central_differences():
    A, lower, d, lptr, lcol, lch, (M, N, T) = _factor_case()
    n = A.shape[0]
    col = K.colours(M, N, T)
    x = np.random.default_rng(8).normal(size=(n, 8)) * 0.05
    z = K.rademacher(n, 8, 3)
    adiag = A.diagonal()
    corr = np.random.default_rng(9).normal(scale=0.01, size=(n, Le.NPAIR))

    def loss(c):
        lo, dg = Le.corrected_factor(lower, d, lptr, lcol, lch, adiag, c)
        return Le.jeffreys_grad(lo, dg, lptr, lcol, lch, A, col, x, z)[0]
    lo, dg = Le.corrected_factor(lower, d, lptr, lcol, lch, adiag, corr)
    _, dout = Le.jeffreys_grad(lo, dg, lptr, lcol, lch, A, col, x, z)
    rng = np.random.default_rng(10)
    used = np.concatenate([lch[lch >= 0], np.arange(n) * Le.NPAIR])
    for e in rng.choice(used, 30, replace=False):
        i, ch = divmod(int(e), Le.NPAIR)
        h = 1e-6
        cp, cm = corr.copy(), corr.copy()
        cp[i, ch] += h
        cm[i, ch] -= h
        fd = (loss(cp) - loss(cm)) / (2 * h)
        assert abs(fd - dout[i, ch]) <= 1e-5 * max(1.0, abs(fd)), (i, ch, fd, dout[i, ch])


def test_jeffreys_vanishes_at_the_exact_factor():
    """With exact moments (x x^T / K = A^-1, z z^T / K = I) J = sum(lambda +
    1/lambda)/2 - n over the spectrum of L'^-1 A L'^-T: 0 at the exact
    Cholesky factor, positive for IC(0)."""
    A, lower, d, lptr, lcol, lch, (M, N, T) = _factor_case()
    n = A.shape[0]
    col = K.colours(M, N, T)
    Ad = A.toarray()
    x = np.sqrt(n) * np.linalg.cholesky(np.linalg.inv(Ad))
    z = np.sqrt(n) * np.eye(n)
    loss_ic = Le.jeffreys_grad(lower, d, lptr, lcol, lch, A, col, x, z)[0]
    order = np.lexsort((np.arange(n), col))
    Fo = K.full_factor(sparse.csr_matrix((lower, lcol, lptr), shape=(n, n)), d).toarray()[np.ix_(order, order)]
    Bm = np.linalg.solve(Fo, np.linalg.solve(Fo, Ad[np.ix_(order, order)]).T)
    lam = np.linalg.eigvalsh(0.5 * (Bm + Bm.T))
    assert abs(loss_ic - (0.5 * (lam + 1 / lam).sum() - n)) <= 1e-8 * n
    assert loss_ic > 0


def test_kaporin_gradient_matches_central_differences():
    """Kaporin's condition number (scale-invariant) of the corrected factor:
    the composed gradient against central differences of the loss, and the
    loss is unchanged by a global rescaling of L' (corr[:, 0] += c)."""
    A, lower, d, lptr, lcol, lch, (M, N, T) = _factor_case()
    n = A.shape[0]
    col = K.colours(M, N, T)
    z = K.rademacher(n, 8, 3)
    adiag = A.diagonal()
    corr = np.random.default_rng(11).normal(scale=0.01, size=(n, Le.NPAIR))

    def loss(c):
        lo, dg = Le.corrected_factor(lower, d, lptr, lcol, lch, adiag, c)
        return Le.kaporin_grad(lo, dg, lptr, lcol, lch, A, col, z, 2 * np.log(dg).sum())[0]
    lo, dg = Le.corrected_factor(lower, d, lptr, lcol, lch, adiag, corr)
    _, dout = Le.kaporin_grad(lo, dg, lptr, lcol, lch, A, col, z, 2 * np.log(dg).sum())
    rng = np.random.default_rng(12)
    used = np.concatenate([lch[lch >= 0], np.arange(n) * Le.NPAIR])
    for e in rng.choice(used, 30, replace=False):
        i, ch = divmod(int(e), Le.NPAIR)
        h = 1e-6
        cp, cm = corr.copy(), corr.copy()
        cp[i, ch] += h
        cm[i, ch] -= h
        fd = (loss(cp) - loss(cm)) / (2 * h)
        assert abs(fd - dout[i, ch]) <= 1e-5 * max(1.0, abs(fd)), (i, ch, fd, dout[i, ch])
    # a global scale of L' (with the off-diagonal entries scaled alike) leaves it unchanged
    lo2, dg2 = 1.7 * lo, 1.7 * dg
    l1 = Le.kaporin_grad(lo, dg, lptr, lcol, lch, A, col, z, 2 * np.log(dg).sum())[0]
    l2 = Le.kaporin_grad(lo2, dg2, lptr, lcol, lch, A, col, z, 2 * np.log(dg2).sum())[0]
    assert abs(l1 - l2) <= 1e-10 * abs(l1)

