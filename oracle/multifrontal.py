"""CPU multifrontal Cholesky of Q (test infrastructure and CPU baseline).

Stand-in for the CHOLMOD factorisation the reference calls
(``advection_idiffusion2D.py:116,173``; scikit-sparse 0.4.12 / SuiteSparse,
not installed here): a supernodal (multifrontal) Cholesky with a geometric
nested-dissection ordering of the M x N x T lattice, dense LAPACK/BLAS on each
front (numpy's multithreaded BLAS).  Written independently of the GPU
library: its own box bisection (separators two cells wide in space, one
slice in time -- the reach of Q's stencil), its own symbolic analysis from
Q's sparsity pattern.
"""
import numpy as np
from scipy import linalg as sla


def box_nd(M, N, T, leaf=256):
    """Postorder list of (own lattice indices, child ids) for a recursive
    bisection of the box [0,M) x [0,N) x [0,T) (non-periodic lattices)."""
    nodes = []

    def idx(x0, x1, y0, y1, t0, t1):
        i, j, t = np.meshgrid(np.arange(x0, x1), np.arange(y0, y1), np.arange(t0, t1), indexing="ij")
        return (i + j * M + t * M * N).ravel()

    def rec(x0, x1, y0, y1, t0, t1):
        nx, ny, nt = x1 - x0, y1 - y0, t1 - t0
        if nx <= 0 or ny <= 0 or nt <= 0:
            return None
        if nx * ny * nt <= leaf:
            nodes.append((np.sort(idx(x0, x1, y0, y1, t0, t1)), []))
            return len(nodes) - 1
        cand = []
        if nx >= 5:
            cand.append((2 * ny * nt, "x"))
        if ny >= 5:
            cand.append((2 * nx * nt, "y"))
        if nt >= 3:
            cand.append((nx * ny, "t"))
        if not cand:
            nodes.append((np.sort(idx(x0, x1, y0, y1, t0, t1)), []))
            return len(nodes) - 1
        _, ax = min(cand)
        if ax == "x":
            c = x0 + (nx - 2) // 2
            a = rec(x0, c, y0, y1, t0, t1)
            b = rec(c + 2, x1, y0, y1, t0, t1)
            sep = idx(c, c + 2, y0, y1, t0, t1)
        elif ax == "y":
            c = y0 + (ny - 2) // 2
            a = rec(x0, x1, y0, c, t0, t1)
            b = rec(x0, x1, c + 2, y1, t0, t1)
            sep = idx(x0, x1, c, c + 2, t0, t1)
        else:
            c = t0 + (nt - 1) // 2
            a = rec(x0, x1, y0, y1, t0, c)
            b = rec(x0, x1, y0, y1, c + 1, t1)
            sep = idx(x0, x1, y0, y1, c, c + 1)
        nodes.append((np.sort(sep), [k for k in (a, b) if k is not None]))
        return len(nodes) - 1

    rec(0, M, 0, N, 0, T)
    return nodes


class Multifrontal:
    """Symbolic analysis once, then numeric factorisations of matrices with
    the same pattern.  logdet() after factor()."""

    def __init__(self, Q, M, N, T, leaf=256):
        nodes = box_nd(M, N, T, leaf)
        perm = np.concatenate([own for own, _ in nodes])
        n = Q.shape[0]
        assert perm.size == n and np.array_equal(np.sort(perm), np.arange(n))
        iperm = np.empty(n, np.int64)
        iperm[perm] = np.arange(n)
        P = Q.tocsr()[perm][:, perm].tocsr()
        P.sort_indices()
        self.P_pattern = P
        self.perm, self.iperm = perm, iperm
        starts = np.cumsum([0] + [len(own) for own, _ in nodes])
        self.fronts = []
        Bs = []
        for k, (own, ch) in enumerate(nodes):
            s0, s1 = starts[k], starts[k + 1]
            cols = P.indices[P.indptr[s0]:P.indptr[s1]]
            cand = [cols[cols >= s1]] + [Bs[c][Bs[c] >= s1] for c in ch]
            B = np.unique(np.concatenate(cand)) if cand else np.zeros(0, np.int64)
            Bs.append(B)
            self.fronts.append(dict(s0=s0, s1=s1, B=B, ch=ch))
        for f in self.fronts:
            ind = np.concatenate([np.arange(f["s0"], f["s1"]), f["B"]])
            f["ind"] = ind
        for f in self.fronts:
            for c in f["ch"]:
                cf = self.fronts[c]
                cf["rel"] = np.searchsorted(f["ind"], cf["B"])
        self.flops = sum((f["s1"] - f["s0"]) ** 3 / 3 + len(f["B"]) * (f["s1"] - f["s0"]) ** 2
                         + len(f["B"]) ** 2 * (f["s1"] - f["s0"]) for f in self.fronts)

    def factor(self, Q):
        P = Q.tocsr()[self.perm][:, self.perm].tocsr()
        P.sort_indices()
        logdet = 0.0
        upd = {}
        for k, f in enumerate(self.fronts):
            s0, s1, ind = f["s0"], f["s1"], f["ind"]
            p, nf = s1 - s0, len(ind)
            F = np.zeros((nf, nf))
            sub = P[s0:s1]
            rows = np.repeat(np.arange(p), np.diff(sub.indptr))
            cols = sub.indices
            keep = cols >= s0
            pos = np.searchsorted(ind, cols[keep])
            F[rows[keep], pos] = sub.data[keep]
            for c in f["ch"]:
                rel = self.fronts[c]["rel"]
                F[np.ix_(rel, rel)] += upd.pop(c)
            L11 = sla.cholesky(F[:p, :p], lower=True, check_finite=False)
            logdet += 2.0 * np.log(np.diag(L11)).sum()
            if nf > p:
                L21 = sla.solve_triangular(L11, F[:p, p:], lower=True, check_finite=False).T
                U = F[p:, p:] - L21 @ L21.T
                upd[k] = U
        return logdet
