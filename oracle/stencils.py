"""Finite-volume diffusion / upwind-advection stencils as slot triplets.

Oracle restatement (test infrastructure) of the reference's native stencils
``src/spdepy/spdes/ccode/{AcH,AH,Acw,Aw}_2D_b{1,2,3}.cpp``.  Each builder
returns ``(row, col, val)`` with one slot per (cell, stencil point) in the
reference's slot order, ``row == M*N`` marking a dropped slot, exactly the
layout the reference's ``*_Row/_Col/_Val`` accessors expose.  Arithmetic is
written in the reference's evaluation order so the golden vectors produced
by the compiled reference match bit for bit.

Slot order: diffusion [C, E, W, N, S, NE, SW, NW, SE]; advection [C, E, W, N, S].
Face order: diffusion H[k][W,E,S,N][2][2]; advection G[k][E,N,W,S].
"""
import numpy as np
from scipy import sparse


def _cells(M, N):
    k = np.arange(M * N)
    return k, k % M, k // M


def _nbr_clamped(i, j, M, N):
    """bc=1 neighbour indices: the out-of-domain neighbour is the cell itself
    (``if (i==0) i_n=i; else if (i==M-1) i_p=i``)."""
    iN = np.where(i == 0, i, i - 1)
    iP = np.where((i == M - 1) & (i != 0), i, i + 1)
    jN = np.where(j == 0, j, j - 1)
    jP = np.where((j == N - 1) & (j != 0), j, j + 1)
    return iN, iP, jN, jP


def _nbr_wrap(i, j, M, N):
    iN = np.where(i == 0, M - 1, i - 1)
    iP = np.where((i == M - 1) & (i != 0), 0, i + 1)
    jN = np.where(j == 0, N - 1, j - 1)
    jP = np.where((j == N - 1) & (j != 0), 0, j + 1)
    return iN, iP, jN, jP


def _diff_cols(M, i, j, iN, iP, jN, jP):
    k = i + j * M
    return np.stack([k, j * M + iP, j * M + iN, jP * M + i, jN * M + i,
                     jP * M + iP, jN * M + iN, jP * M + iN, jN * M + iP], axis=1)


def _diff_vals(hx, hy, W00, E00, W10, E10, S01, N01, S11, N11):
    """Per-cell 9-point values (before any zero-flux fold), reference order."""
    rx, ry = hy / hx, hx / hy
    cE = rx * E00 + 0.25 * (N01 - S01)
    cW = rx * W00 - 0.25 * (N01 - S01)
    cN = ry * N11 + 0.25 * (E10 - W10)
    cS = ry * S11 - 0.25 * (E10 - W10)
    cNE = 0.25 * (N01 + E10)
    cSW = 0.25 * (S01 + W10)
    cNW = -0.25 * (N01 + W10)
    cSE = -0.25 * (S01 + E10)
    cC = (-hy / hx) * (E00 + W00) - ry * (N11 + S11)
    return cC, [cE, cW, cN, cS, cNE, cSW, cNW, cSE]


def _fold_b1(M, N, cols, hx, hy, W00, E00, W10, E10, S01, N01, S11, N11):
    """bc=1 zero-flux: a slot whose column is the cell itself is dropped and
    its value folded into the centre (``AH_2D_b1.cpp:54-101``, same order)."""
    k = cols[:, 0]
    rx, ry = hy / hx, hx / hy
    a = [rx * E00, rx * W00, ry * N11, ry * S11]
    b = [0.25 * (N01 - S01), 0.25 * (N01 - S01), 0.25 * (E10 - W10), 0.25 * (E10 - W10)]
    sign = [1.0, -1.0, 1.0, -1.0]
    corner = [0.25 * (N01 + E10), 0.25 * (S01 + W10), 0.25 * (N01 + W10), 0.25 * (S01 + E10)]
    csign = [1.0, 1.0, -1.0, -1.0]
    rem = np.zeros(k.size)
    self_mask = cols[:, 1:] == k[:, None]
    for s in range(4):
        m = self_mask[:, s]
        upd = rem + a[s]
        upd = upd + b[s] if sign[s] > 0 else upd - b[s]
        rem = np.where(m, upd, rem)
    for s in range(4):
        m = self_mask[:, 4 + s]
        upd = rem + corner[s] if csign[s] > 0 else rem - corner[s]
        rem = np.where(m, upd, rem)
    return rem, self_mask


def _pack(M, N, cols, vals, drop):
    n = M * N
    nslot = cols.shape[1]
    row = np.repeat(np.arange(n), nslot).reshape(n, nslot)
    row = np.where(drop, n, row)
    return row.ravel().astype(np.int32), cols.ravel().astype(np.int32), vals.ravel()


def _dirichlet_drop9(i, j, M, N):
    drop = np.zeros((i.size, 9), dtype=bool)
    lo, hi = i == 0, (i == M - 1) & (i != 0)
    for s in (2, 6, 7):
        drop[:, s] |= lo
    for s in (1, 5, 8):
        drop[:, s] |= hi
    lo, hi = j == 0, (j == N - 1) & (j != 0)
    for s in (4, 6, 8):
        drop[:, s] |= lo
    for s in (3, 5, 7):
        drop[:, s] |= hi
    return drop


def ah_var(M, N, H, hx, hy, bc):
    """Variable-H diffusion stencil, H of shape (M*N, 4, 2, 2).

    bc=1 reproduces the stale-index zeroing of ``AH_2D_b1.cpp:25-52``: the
    boundary-face zeroing lands on the previously visited cell, so only cell
    0 has its W and S face coefficients zeroed."""
    H = np.array(H, dtype=np.float64).reshape(M * N, 4, 2, 2).copy()
    k, i, j = _cells(M, N)
    if bc == 1:
        H[0, 0, 0, 0] = 0.0
        H[0, 0, 1, 0] = 0.0
        H[0, 2, 1, 1] = 0.0
        H[0, 2, 0, 1] = 0.0
    faces = dict(W00=H[:, 0, 0, 0], E00=H[:, 1, 0, 0], W10=H[:, 0, 1, 0], E10=H[:, 1, 1, 0],
                 S01=H[:, 2, 0, 1], N01=H[:, 3, 0, 1], S11=H[:, 2, 1, 1], N11=H[:, 3, 1, 1])
    return _diffusion(M, N, k, i, j, hx, hy, bc, faces)


def _diffusion(M, N, k, i, j, hx, hy, bc, faces):
    if bc == 1:
        iN, iP, jN, jP = _nbr_clamped(i, j, M, N)
    elif bc == 2:
        iN, iP, jN, jP = _nbr_wrap(i, j, M, N)
    else:
        iN, iP, jN, jP = i - 1, i + 1, j - 1, j + 1
    cols = _diff_cols(M, i, j, iN, iP, jN, jP)
    cC, off = _diff_vals(hx, hy, **faces)
    if bc == 1:
        rem, drop8 = _fold_b1(M, N, cols, hx, hy, **faces)
        cC = cC + rem
        drop = np.hstack([np.zeros((k.size, 1), bool), drop8])
    elif bc == 2:
        cC = cC + 0.0
        drop = np.zeros((k.size, 9), bool)
    else:
        cC = cC + 0.0
        drop = _dirichlet_drop9(i, j, M, N)
    vals = np.stack([cC] + off, axis=1)
    return _pack(M, N, cols, vals, drop)


def ah_const(M, N, H, hx, hy, bc):
    """Constant-H diffusion stencil (``AcH_2D_b{1,3}.cpp``).

    bc=1 zeroes the boundary-face copies of the current cell correctly
    (``AcH_2D_b1.cpp:51-68``); the off-diagonal face terms all read H[1][0].
    bc=2: the reference takes a scalar where its caller passes a pointer, so
    its output is undefined; this restatement uses the periodic variable
    stencil with every face equal to H (parity unpinned)."""
    H = np.asarray(H, dtype=np.float64).reshape(2, 2)
    k, i, j = _cells(M, N)
    n = M * N
    if bc == 3:
        cols = _diff_cols(M, i, j, i - 1, i + 1, j - 1, j + 1)
        rx, ry = hy / hx, hx / hy
        cC = np.full(n, (-2.0 * hy / hx) * H[0, 0] - (2.0 * hx / hy) * H[1, 1] + 0.0)
        e = np.full(n, rx * H[0, 0])
        nn = np.full(n, ry * H[1, 1])
        c = np.full(n, 0.25 * (H[0, 1] + H[1, 0]))
        vals = np.stack([cC, e, e, nn, nn, c, c, -c, -c], axis=1)
        return _pack(M, N, cols, vals, _dirichlet_drop9(i, j, M, N))
    if bc == 2:
        return ah_var(M, N, np.broadcast_to(H, (n, 4, 2, 2)), hx, hy, 2)
    one = np.ones(n)
    W00, E00 = H[0, 0] * one, H[0, 0] * one
    W10, E10 = H[1, 0] * one, H[1, 0] * one
    S01, N01 = H[1, 0] * one, H[1, 0] * one
    S11, N11 = H[1, 1] * one, H[1, 1] * one
    lo_i, hi_i = i == 0, (i == M - 1) & (i != 0)
    lo_j, hi_j = j == 0, (j == N - 1) & (j != 0)
    W00 = np.where(lo_i, 0.0, W00)
    W10 = np.where(lo_i, 0.0, W10)
    E00 = np.where(hi_i, 0.0, E00)
    E10 = np.where(hi_i, 0.0, E10)
    S11 = np.where(lo_j, 0.0, S11)
    S01 = np.where(lo_j, 0.0, S01)
    N11 = np.where(hi_j, 0.0, N11)
    N01 = np.where(hi_j, 0.0, N01)
    faces = dict(W00=W00, E00=E00, W10=W10, E10=E10, S01=S01, N01=N01, S11=S11, N11=N11)
    return _diffusion(M, N, k, i, j, hx, hy, 1, faces)


def _adv_cols(M, i, j, iN, iP, jN, jP):
    k = i + j * M
    return np.stack([k, j * M + iP, j * M + iN, jP * M + i, jN * M + i], axis=1)


def _dirichlet_drop5(i, j, M, N):
    drop = np.zeros((i.size, 5), dtype=bool)
    drop[:, 1] = i == M - 1
    drop[:, 2] = i == 0
    drop[:, 3] = j == N - 1
    drop[:, 4] = j == 0
    return drop


def aw_const(M, N, G, hx, hy, diff, bc):
    """Constant-velocity upwind stencil (``Acw_2D_b{1,2,3}.cpp``).

    diff=3 is the operator; diff=1/2 are the reference's d/dwx, d/dwy
    stencils (NaN at G=0, not filtered).  bc=1 subtracts |G|h/2 from the
    centre at every boundary for all diff modes (``Acw_2D_b1.cpp:63-80``)."""
    G = np.asarray(G, dtype=np.float64).ravel()
    k, i, j = _cells(M, N)
    n = M * N
    with np.errstate(invalid="ignore", divide="ignore"):
        s0, s1 = G[0] / abs(G[0]), G[1] / abs(G[1])
    one = np.ones(n)
    zero = np.zeros(n)
    if diff == 1:
        cC = s0 * hy * one
        vals = [cC, (-(s0 - 1.0) * hy / 2) * one, (-(s0 + 1.0) * hy / 2) * one, zero, zero]
    elif diff == 2:
        cC = s1 * hx * one
        vals = [cC, zero, zero, (-(s1 - 1.0) * hx / 2) * one, (-(s1 + 1.0) * hx / 2) * one]
    else:
        cC = (abs(G[0]) * hy + abs(G[1]) * hx) * one
        vals = [cC, (-(abs(G[0]) - G[0]) * hy / 2) * one, (-(abs(G[0]) + G[0]) * hy / 2) * one,
                (-(abs(G[1]) - G[1]) * hx / 2) * one, (-(abs(G[1]) + G[1]) * hx / 2) * one]
    if bc == 1:
        iN, iP, jN, jP = _nbr_clamped(i, j, M, N)
        lo_i, hi_i = i == 0, (i == M - 1) & (i != 0)
        lo_j, hi_j = j == 0, (j == N - 1) & (j != 0)
        c = vals[0]
        c = np.where(lo_i | hi_i, c - abs(G[0]) * hy / 2, c)
        c = np.where(lo_j | hi_j, c - abs(G[1]) * hx / 2, c)
        vals[0] = c
        drop = np.zeros((n, 5), bool)
        drop[:, 1], drop[:, 2], drop[:, 3], drop[:, 4] = hi_i, lo_i, hi_j, lo_j
    elif bc == 2:
        iN, iP, jN, jP = _nbr_wrap(i, j, M, N)
        vals[0] = vals[0] + 0.0
        drop = np.zeros((n, 5), bool)
    else:
        iN, iP, jN, jP = i - 1, i + 1, j - 1, j + 1
        vals[0] = vals[0] + 0.0
        drop = _dirichlet_drop5(i, j, M, N)
    cols = _adv_cols(M, i, j, iN, iP, jN, jP)
    return _pack(M, N, cols, np.stack(vals, axis=1), drop)


def aw_var(M, N, G, hx, hy, diff, bc, dG=None):
    """Variable-velocity upwind stencil, G and dG of shape (M*N, 4) in face
    order E, N, W, S (``Aw_2D_b{1,2,3}.cpp``).  bc=1 zeroes G and dG on the
    boundary faces of the current cell before evaluating."""
    G = np.array(G, dtype=np.float64).reshape(M * N, 4).copy()
    dG = np.zeros((M * N, 4)) if dG is None else np.array(dG, dtype=np.float64).reshape(M * N, 4).copy()
    k, i, j = _cells(M, N)
    n = M * N
    if bc == 1:
        iN, iP, jN, jP = _nbr_clamped(i, j, M, N)
        e_b, w_b = iP == i, iN == i
        n_b, s_b = jP == j, jN == j
        for f, m in ((0, e_b), (2, w_b), (1, n_b), (3, s_b)):
            G[m, f] = 0.0
            dG[m, f] = 0.0
        drop = np.stack([np.zeros(n, bool), e_b, w_b, n_b, s_b], axis=1)
    elif bc == 2:
        iN, iP, jN, jP = _nbr_wrap(i, j, M, N)
        drop = np.zeros((n, 5), bool)
    else:
        iN, iP, jN, jP = i - 1, i + 1, j - 1, j + 1
        drop = _dirichlet_drop5(i, j, M, N)
    g0, g1, g2, g3 = G[:, 0], G[:, 1], G[:, 2], G[:, 3]
    d0, d1, d2, d3 = dG[:, 0], dG[:, 1], dG[:, 2], dG[:, 3]
    zero = np.zeros(n)
    with np.errstate(invalid="ignore", divide="ignore"):
        if diff == 1:
            u0, u2 = g0 / np.abs(g0) * d0, g2 / np.abs(g2) * d2
            vals = [(u0 + d0 + u2 - d2) * hy / 2, (-(u0 - d0)) * hy / 2, (-(u2 + d2)) * hy / 2,
                    zero, zero]
        elif diff == 2:
            u1, u3 = g1 / np.abs(g1) * d1, g3 / np.abs(g3) * d3
            vals = [(u1 + d1 + u3 - d3) * hx / 2, zero, zero, (-(u1 - d1)) * hx / 2,
                    (-(u3 + d3)) * hx / 2]
        else:
            a0, a1, a2, a3 = np.abs(g0), np.abs(g1), np.abs(g2), np.abs(g3)
            vals = [(a0 + g0 + a2 - g2) * hy / 2 + (a1 + g1 + a3 - g3) * hx / 2,
                    (-(a0 - g0)) * hy / 2, (-(a2 + g2)) * hy / 2,
                    (-(a1 - g1)) * hx / 2, (-(a3 + g3)) * hx / 2]
    cols = _adv_cols(M, i, j, iN, iP, jN, jP)
    return _pack(M, N, cols, np.stack(vals, axis=1), drop)


def triplets_to_csc(row, col, val, n, nan_to_zero=False):
    """The reference wrappers' conversion (``advection_idiffusion2D.py:210-244``,
    ``var_advection_var_diffusion2D.py:239-276``): drop sentinel rows, optionally
    zero NaN (variable advection only), COO->CSC summing duplicates and keeping
    explicit zeros."""
    keep = row != n
    r, c, v = row[keep], col[keep], val[keep].copy()
    if nan_to_zero:
        v[np.isnan(v)] = 0.0
    return sparse.csc_matrix((v, (r, c)), shape=(n, n))
