"""CPU restatement of the iterative path (test infrastructure only).

The reference factorises Q_c with CHOLMOD and estimates tr(Q_c^-1 dQ) with
Rademacher probes (``advection_idiffusion2D.py:173-189``,
``var_advection_var_diffusion2D.py:205-221``).  The GPU path replaces the
factorisation with preconditioned conjugate gradients; this module restates
its pieces in numpy so the kernels can be checked on the same inputs:

  * ``colours``       -- the 18-colour ordering (i mod 3) + 3 (j mod 3) + 9 (t mod 2);
  * ``ic0``           -- IC(0) of A on its own pattern in that ordering, with the
                         diagonal scaled by (1 + shift); pinned by the IC(0)
                         identity (L L^T)_ij = A_ij on the pattern;
  * ``splitmix64``    -- the probe hash (pinned by the published first outputs
                         of Vigna's splitmix64 generator);
  * ``rademacher``    -- the probe matrix the GPU draws for a seed;
  * ``gaussian``      -- the normal draws of the conditional samplers (Box-Muller
                         on splitmix64 uniforms), pinned by moments and the
                         Box-Muller identity;
  * ``fform_t``       -- F^T D^(1/2) z for Q = F^T D F (the right-hand side of a
                         factor-free conditional sample), pinned by
                         (F^T D^(1/2)) (F^T D^(1/2))^T = Q against makeQ;
  * ``slq_logdet``    -- Gauss quadrature of log over Lanczos tridiagonals
                         (Ubaru, Chen, Saad 2017), pinned by exact log-determinants;
  * ``pcg``           -- textbook preconditioned conjugate gradients.
"""
import numpy as np
from scipy import sparse
from scipy.sparse import linalg as spla


def colours(M, N, T):
    k = np.arange(M * N * T)
    i, j, t = k % M, (k // M) % N, k // (M * N)
    return ((i % 3) + 3 * (j % 3) + 9 * (t % 2)).astype(np.int64)


def ic0(A, col, shift=0.0):
    """IC(0) of A (CSR, symmetric) in the colour ordering ``col``: returns
    (L, d) with L strictly lower in that ordering (CSR in lattice indices,
    L[i, j] for colour(j) < colour(i)) and d = diag(L).  Row i's entries are
    computed in (colour, index) order of their columns:
        L_ij = (A_ij - sum_{k < j} L_ik L_jk) / L_jj,  L_ii = sqrt((1 + shift) A_ii - sum_j L_ij^2)."""
    A = A.tocsr()
    n = A.shape[0]
    order = np.lexsort((np.arange(n), col))
    low = {}
    d = np.zeros(n)
    for i in order:
        cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
        vals = A.data[A.indptr[i]:A.indptr[i + 1]]
        lo = [(col[j], j, v) for j, v in zip(cols, vals) if j != i and col[j] < col[i]]
        lo.sort()
        row = {}
        for _, j, a in lo:
            lj = low.get(j, {})
            for kk, ljk in lj.items():
                lik = row.get(kk)
                if lik is not None:
                    a -= lik * ljk
            row[j] = a / d[j]
        aii = (1.0 + shift) * A[i, i]
        s = aii - sum(v * v for v in row.values())
        if s <= 0:
            raise np.linalg.LinAlgError("IC(0) breakdown at %d" % i)
        d[i] = np.sqrt(s)
        low[i] = row
    r, c, v = [], [], []
    for i, row in low.items():
        for j, x in row.items():
            r.append(i)
            c.append(j)
            v.append(x)
    L = sparse.csr_matrix((v, (r, c)), shape=(n, n))
    return L, d


def full_factor(L, d):
    return (L + sparse.diags(d)).tocsr()


def splitmix64(x):
    """Vigna's splitmix64 output function of the state x (uint64 arrays)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rademacher(n, k, seed, scale=1.0):
    """The probe matrix V (n x k) of ``spde_rademacher``: V[i, r] = +-scale
    with the sign from the top bit of splitmix64(seed ^ splitmix64(i k + r))."""
    u = np.arange(n * k, dtype=np.uint64)
    h = splitmix64(np.uint64(seed) ^ splitmix64(u))
    s = np.where((h >> np.uint64(63)) != 0, -scale, scale)
    return s.reshape(n, k)


def uniforms(n, k, seed):
    """The two 53-bit uniforms of unit u = i k + r: U1 = ((h1 >> 11) + 0.5) 2^-53,
    U2 = (h2 >> 11) 2^-53 with h1, h2 = splitmix64(seed ^ splitmix64(2u)), (2u + 1)."""
    u = np.arange(n * k, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h1 = splitmix64(np.uint64(seed) ^ splitmix64(np.uint64(2) * u))
        h2 = splitmix64(np.uint64(seed) ^ splitmix64(np.uint64(2) * u + np.uint64(1)))
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return u1.reshape(n, k), u2.reshape(n, k)


def gaussian(n, k, seed):
    """n x k standard normals: Box-Muller sqrt(-2 log U1) cos(2 pi U2)."""
    u1, u2 = uniforms(n, k, seed)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def fform_t(A, q, A0, V, T, sigma, dt, z):
    """F^T D^(1/2) z for Q = F^T D F with F_00 = V^(-1/2) A0, F_tt = A / V
    (t >= 1), F_(t+1,t) = -I and D = blockdiag(I, q / (dt sigma), ...)
    (the factorisation ``models.sample_fform`` inverts): with w_0 = z_0 and
    w_t = sqrt(q / (dt sigma)) z_t,  (F^T w)_t = F_tt^T w_t - w_(t+1)."""
    Ns = A.shape[0]
    z = np.asarray(z, dtype=np.float64).reshape(T, Ns, -1)
    s = np.sqrt(q / (dt * sigma))[:, None]
    w = z.copy()
    w[1:] *= s
    out = np.empty_like(w)
    out[0] = (A0.T @ w[0]) / np.sqrt(V)
    for t in range(1, T):
        out[t] = (A.T @ w[t]) / V
    out[:-1] -= w[1:]
    return out.reshape(T * Ns, -1)


def slq_logdet(alpha, beta, n):
    """Stochastic Lanczos quadrature: for each probe's tridiagonal T_r (alpha
    on the diagonal, beta[:-1] beside it), n * sum_i (e_1^T y_i)^2 log theta_i;
    returns (mean over probes, standard error, per-probe values).  Steps
    after a zero beta are dropped (invariant subspace)."""
    vals = []
    for a, b in zip(alpha, beta):
        m = len(a)
        for j in range(m - 1):
            if not b[j] > 0:
                m = j + 1
                break
        T = np.diag(a[:m]) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
        th, Y = np.linalg.eigh(T)
        th = np.maximum(th, 1e-300)
        vals.append(n * float(np.sum(Y[0] ** 2 * np.log(th))))
    vals = np.asarray(vals)
    se = vals.std(ddof=1) / np.sqrt(len(vals)) if len(vals) > 1 else np.inf
    return float(vals.mean()), float(se), vals


def lanczos(matvec, q0, steps):
    """Plain Lanczos (no reorthogonalisation) from unit q0: (alpha, beta)."""
    q, qp, bp = q0.copy(), np.zeros_like(q0), 0.0
    a, b = [], []
    for _ in range(steps):
        w = matvec(q)
        al = float(q @ w)
        w = w - al * q - bp * qp
        be = float(np.linalg.norm(w))
        a.append(al)
        b.append(be)
        qp, q, bp = q, (w / be if be > 0 else np.zeros_like(w)), be
    return np.array(a), np.array(b)


def pcg(A, b, apply_m, tol=1e-8, maxit=10000):
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for it in range(maxit):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        z = apply_m(r)
        if np.linalg.norm(r) <= tol * nb:
            return x, it + 1
        rz2 = r @ z
        p = z + rz2 / rz * p
        rz = rz2
    return x, maxit


def ic_solver(L, d, col):
    """apply_m(r) = (F F^T)^-1 r with F = L + diag(d) in the colour ordering."""
    n = L.shape[0]
    order = np.lexsort((np.arange(n), col))
    P = sparse.csr_matrix((np.ones(n), (np.arange(n), order)), shape=(n, n))
    F = (P @ full_factor(L, d) @ P.T).tocsr()
    Ft = F.T.tocsr()

    def apply(r):
        y = spla.spsolve_triangular(F, P @ r, lower=True)
        return P.T @ spla.spsolve_triangular(Ft, y, lower=False)
    return apply
