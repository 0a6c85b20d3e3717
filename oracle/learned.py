"""CPU restatement of the learned preconditioner's network and loss (test
infrastructure only; no reference counterpart -- the reference has no
iterative solver).

  * ``bf16``          -- round-to-nearest-even to bfloat16 (as float32 values);
  * ``conv`` / ``conv_backward`` -- the 3x3 same-padded convolution over the
                         time slices of an M x N x T lattice, pixel-major,
                         with the operands rounded to bfloat16 and exact
                         (float64) accumulation, and its data, weight and bias
                         gradients; ``rounding=False`` gives the float64
                         network whose gradients central differences pin;
  * ``forward`` / ``backward`` -- the 4-layer network of spdepy_amd.learned;
  * ``kl_grad``       -- KL(N(0, A^-1) || N(0, (L' L'^T)^-1)) for a factor L'
                         and samples x, and its gradient for the correction
                         channels (pinned by central differences).
"""
import numpy as np

LAYERS = ((8, True), (32, True), (32, True), (32, False))
NCH, NPAIR = 32, 22


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _shift(X, M, N, T, dx, dy):
    """Y[p] = X[p + (dx, dy)] within each slice, 0 outside."""
    Xi = X.reshape(T, N, M, -1)
    Y = np.zeros_like(Xi)
    ys, ye = max(0, -dy), min(N, N - dy)
    xs, xe = max(0, -dx), min(M, M - dx)
    Y[:, ys:ye, xs:xe] = Xi[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return Y.reshape(M * N * T, -1)


def conv(X, W, b, shape, relu, rounding=True):
    """out[p, co] = act(sum_{tap, ci} W[tap cin + ci, co] X[p + off(tap), ci] + b[co]),
    tap = 3 (dy + 1) + (dx + 1)."""
    M, N, T = shape
    cin = X.shape[1]
    r = bf16 if rounding else (lambda a: a)
    Xr = r(X).astype(np.float64)
    Wr = r(W).astype(np.float64)
    out = np.zeros((X.shape[0], W.shape[1]))
    for tap in range(9):
        dx, dy = tap % 3 - 1, tap // 3 - 1
        out += _shift(Xr, M, N, T, dx, dy) @ Wr[tap * cin:(tap + 1) * cin]
    if b is not None:
        out += b
    return np.maximum(out, 0.0) if relu else out


def conv_backward(X, W, G, Y, shape, rounding=True):
    """Gradients of a layer (output Y = act(conv(X))) from G = dL/dY:
    (dX, dW, db) with dZ = G (Y > 0) when the layer has a ReLU (Y given)."""
    M, N, T = shape
    cin = X.shape[1]
    dZ = G * (Y > 0) if Y is not None else G
    r = bf16 if rounding else (lambda a: a)
    Zr = r(dZ).astype(np.float64)
    Xr = r(X).astype(np.float64)
    Wr = r(W).astype(np.float64)
    dW = np.zeros(W.shape)
    dX = np.zeros(X.shape)
    for tap in range(9):
        dx, dy = tap % 3 - 1, tap // 3 - 1
        dW[tap * cin:(tap + 1) * cin] = _shift(Xr, M, N, T, dx, dy).T @ Zr
        dX += _shift(Zr @ Wr[tap * cin:(tap + 1) * cin].T, M, N, T, -dx, -dy)
    db = dZ.sum(0)
    return dX, dW, db


def unpack(theta):
    out, off = [], 0
    for cin, _ in LAYERS:
        w = 9 * cin * NCH
        out.append((theta[off:off + w].reshape(9 * cin, NCH), theta[off + w:off + w + NCH]))
        off += w + NCH
    return out


def forward(theta, feat, shape, rounding=True):
    acts = [np.asarray(feat, np.float64)]
    for (cin, relu), (W, b) in zip(LAYERS, unpack(theta)):
        acts.append(conv(acts[-1], W, b, shape, relu, rounding))
    return acts


def backward(theta, acts, dout, shape, rounding=True):
    """Flat parameter gradient from dL/d(output channels 0..21)."""
    G = np.zeros((acts[0].shape[0], NCH))
    G[:, :NPAIR] = dout
    params = unpack(theta)
    grads = [None] * len(LAYERS)
    for li in range(len(LAYERS) - 1, -1, -1):
        W, b = params[li]
        relu = LAYERS[li][1]
        dX, dW, db = conv_backward(acts[li], W, G, acts[li + 1] if relu else None, shape, rounding)
        grads[li] = (dW, db)
        G = dX
    return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def corrected_factor(lower, diag, lptr, lcol, lch, adiag, corr):
    """Apply a correction (n x 22) to an IC factor in the Jacobi-scaled space:
    L_ij += sqrt(A_ii) corr[lch], L_ii *= exp(corr[i, 0])."""
    lower = np.array(lower, dtype=np.float64)
    n = diag.size
    rows = np.repeat(np.arange(n), np.diff(lptr))
    ok = lch >= 0
    flat = np.asarray(corr, np.float64).reshape(-1)
    lower[ok] += np.sqrt(adiag[rows[ok]]) * flat[lch[ok]]
    return lower, diag * np.exp(np.asarray(corr, np.float64)[:, 0])


def kl_grad(lower, diag, lptr, lcol, lch, adiag, x):
    """(loss, dout): loss = 1/2 mean_s ||L'^T x_s||^2 - sum log L'_ii for the
    factor L' (strictly lower entries `lower` of row i at columns lcol, and
    diag), dout[c, ch] = d loss / d corr for the correction channels."""
    n, K = x.shape
    rows = np.repeat(np.arange(n), np.diff(lptr))
    # w = L'^T x: w_j = L'_jj x_j + sum_{i: (i, j) lower} L'_ij x_i
    w = diag[:, None] * x
    np.add.at(w, lcol, lower[:, None] * x[rows])
    loss = 0.5 * (w * w).sum() / K - np.log(diag).sum()
    dout = np.zeros(n * NPAIR)
    ok = lch >= 0
    g = (x[rows[ok]] * w[lcol[ok]]).sum(1) / K
    dout[lch[ok]] = np.sqrt(adiag[rows[ok]]) * g
    dd = (x * w).sum(1) / K
    dout[np.arange(n) * NPAIR] = diag * dd - 1.0
    return loss, dout.reshape(n, NPAIR)


def forward_pairs():
    """Lattice offsets (dx, dy, dt) of correction channels 1..21: the forward
    half of the 5x5 window of the slice ((dy, dx) lexicographic with dy > 0
    or dy = 0, dx > 0), then the 3x3 window of the next slice."""
    same = [(dx, dy, 0) for dy in range(0, 3) for dx in range(-2, 3) if dy > 0 or dx > 0]
    nxt = [(dx, dy, 1) for dy in range(-1, 2) for dx in range(-1, 2)]
    return same + nxt


def pair_channels(M, N, T, rows, cols):
    """cell * 22 + channel of each factor entry (rows[e], cols[e]) on a
    non-periodic lattice: the pair's forward offset from one end picks the
    channel, predicted at that end."""
    ch = {o: c + 1 for c, o in enumerate(forward_pairs())}
    Ns = M * N
    rows, cols = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    out = np.full(rows.size, -1, np.int64)
    for e, (a, b) in enumerate(zip(rows, cols)):
        d = (int(b % M - a % M), int((b // M) % N - (a // M) % N), int(b // Ns - a // Ns))
        if d in ch:
            out[e] = a * NPAIR + ch[d]
        else:
            nd = (-d[0], -d[1], -d[2])
            if nd in ch:
                out[e] = b * NPAIR + ch[nd]
    return out


def jeffreys_grad(lower, diag, lptr, lcol, lch, A, col, x, z):
    """(loss, dout) of J = 1/2 mean ||L'^T x||^2 + 1/2 mean y^T A y - n with
    y = L'^-T z (L' lower-triangular in the colour ordering `col`): the
    symmetrised KL divergence between N(0, A^-1) (samples x) and
    N(0, (L' L'^T)^-1), and dJ/dcorr for the correction channels."""
    from scipy import sparse
    from scipy.sparse import linalg as spla
    n, K = x.shape
    rows = np.repeat(np.arange(n), np.diff(lptr))
    w = diag[:, None] * x
    np.add.at(w, lcol, lower[:, None] * x[rows])
    Lf = sparse.csr_matrix((np.concatenate([lower, diag]), (np.concatenate([rows, np.arange(n)]),
                                                            np.concatenate([lcol, np.arange(n)]))), shape=(n, n))
    order = np.lexsort((np.arange(n), col))
    P = sparse.csr_matrix((np.ones(n), (np.arange(n), order)), shape=(n, n))
    Fo = (P @ Lf @ P.T).tocsr()
    y = P.T @ spla.spsolve_triangular(Fo.T.tocsr(), P @ z, lower=False)
    v = A @ y
    u = P.T @ spla.spsolve_triangular(Fo, P @ v, lower=True)
    loss = 0.5 * (w * w).sum() / K + 0.5 * (y * v).sum() / K - n
    dout = np.zeros(n * NPAIR)
    adiag = A.diagonal()
    ok = lch >= 0
    g = ((x[rows[ok]] * w[lcol[ok]]).sum(1) - (y[rows[ok]] * u[lcol[ok]]).sum(1)) / K
    dout[lch[ok]] = np.sqrt(adiag[rows[ok]]) * g
    dd = ((x * w).sum(1) - (y * u).sum(1)) / K
    dout[np.arange(n) * NPAIR] = diag * dd
    return loss, dout.reshape(n, NPAIR)


def kaporin_grad(lower, diag, lptr, lcol, lch, A, col, z, ld_m):
    """(loss, dout) of Kaporin's condition number per row,
        K/n = log(t / n) + (1/n) log|L' L'^T|  (+ the constant -(1/n) log|A|),
    t = mean_s z_s^T L'^-1 A L'^-T z_s, from the Jeffreys gradient with zero
    samples (its first term vanishes): dt = 2 dJ0, and d log L'_ii / d corr[i, 0] = 1.
    ld_m = log|L' L'^T| = 2 sum log L'_ii."""
    n = diag.size
    j0, g0 = jeffreys_grad(lower, diag, lptr, lcol, lch, A, col, np.zeros((n, z.shape[1])), z)
    t = 2.0 * (j0 + n)
    loss = np.log(t / n) + ld_m / n
    dout = (2.0 / t) * g0
    dout[:, 0] += 2.0 / n
    return loss, dout

