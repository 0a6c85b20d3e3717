"""Host-side symbolic analysis (no GPU): structural pattern of Q, nested
dissection, front index sets and maps.  A numpy multifrontal factorisation
and selected inverse driven by the exported plan must reproduce the dense
Cholesky log-determinant and the dense inverse on the pattern."""
import numpy as np
import pytest
from scipy import sparse

from oracle import fixtures as F
from spdepy_amd import _native as nat
from spdepy_amd.plan import Plan


def _mf_factor(e, A):
    """Multifrontal Cholesky + selected inverse from the exported plan (dense
    numpy per front).  A: scipy CSR in lattice order on the structural pattern."""
    nodes, foff, bnd, relpos, smap = e["nodes"], e["foff"], e["bnd"], e["relpos"], e["map"]
    nn = nodes.shape[0]
    total = int(foff[-1] + nodes[-1, 2] ** 2)
    arena = np.zeros(total)
    coo = A.tocoo()
    indptr, indices = e["indptr"], e["indices"]
    # values in structural order
    Ad = A.tocsr()
    vals = np.zeros(len(indices))
    for r in range(len(indptr) - 1):
        row = Ad.getrow(r).toarray().ravel()
        vals[indptr[r]:indptr[r + 1]] = row[indices[indptr[r]:indptr[r + 1]]]
    m = smap >= 0
    arena[smap[m]] = vals[m]
    fronts = []
    for k in range(nn):
        n = nodes[k, 2]
        fronts.append(arena[foff[k]:foff[k] + n * n].reshape(n, n, order="F"))
    Ls = [None] * nn
    logdet = 0.0
    for k in range(nn):
        start, p, n, parent, c0, c1 = nodes[k, :6]
        Fm = fronts[k]
        for c in (c0, c1):
            if c < 0:
                continue
            pc, nc = nodes[c, 1], nodes[c, 2]
            U = np.tril(fronts[c][pc:, pc:])
            rp = relpos[c]
            Fm[np.ix_(rp, rp)] += U
        Fl = np.tril(Fm)
        Fs = Fl + np.tril(Fl, -1).T
        L11 = np.linalg.cholesky(Fs[:p, :p])
        L21 = np.linalg.solve(L11, Fs[:p, p:]).T if n > p else np.zeros((0, p))
        logdet += 2 * np.log(np.diag(L11)).sum()
        Fm[:, :] = 0
        Fm[:p, :p] = L11
        Fm[p:, :p] = L21
        Fm[p:, p:] = np.tril(Fs[p:, p:] - L21 @ L21.T)
        Ls[k] = (L11, L21)
    # selected inverse, top-down
    S = [None] * nn
    for k in range(nn - 1, -1, -1):
        start, p, n, parent = nodes[k, :4]
        L11, L21 = Ls[k]
        b = n - p
        if parent >= 0 and b > 0:
            rp = relpos[k]
            S22 = S[parent][np.ix_(rp, rp)]
        else:
            S22 = np.zeros((0, 0))
        Z = np.linalg.inv(L11)
        X = L21 @ Z
        S21 = -S22 @ X
        S11 = Z.T @ Z - X.T @ S21
        Sf = np.zeros((n, n))
        Sf[:p, :p], Sf[p:, :p], Sf[:p, p:], Sf[p:, p:] = S11, S21, S21.T, S22
        S[k] = Sf
    return logdet, S


CASES = [
    (7, 6, 3, 1, 16),
    (6, 6, 4, 2, 20),
    (6, 7, 3, 3, 12),
    (9, 8, 5, 1, 40),
    (5, 5, 2, 1, 1000),
]


@pytest.mark.parametrize("M,N,T,bc,leaf", CASES)
def test_plan_reproduces_dense_factor_and_inverse(M, N, T, bc, leaf):
    if bc == 2:
        c = F.small(M, N, T, bc=3)
    else:
        c = F.small(M, N, T, bc=bc)
    Q, _ = c["model"].makeQ(c["par"], grad=False)
    pl = Plan(M, N, T, bc, leaf_max=leaf, nslots=0, host_only=True)
    e = pl.export()
    n = M * N * T
    S = sparse.csr_matrix((np.ones(len(e["indices"])), e["indices"], e["indptr"]), shape=(n, n))
    if bc != 2:
        # the structural pattern contains the reference pattern of Q
        assert (abs(Q).astype(bool) > S.astype(bool)).nnz == 0
        A = Q.tocsr()
    else:
        rng = np.random.default_rng(0)
        R = sparse.random(n, n, density=1.0, random_state=1).multiply(S)
        A = (R @ R.T).multiply(S) + sparse.eye(n) * 50.0
        A = sparse.csr_matrix(A)
    perm = e["perm"]
    assert np.array_equal(np.sort(perm), np.arange(n))
    logdet, Sig = _mf_factor(e, A)
    dense = A.toarray()
    sign, ref = np.linalg.slogdet(dense)
    assert sign > 0
    assert abs(logdet - ref) <= 1e-10 * abs(ref)
    # selected inverse on the pattern (lower entries via the map)
    inv = np.linalg.inv(dense)
    nodes, foff = e["nodes"], e["foff"]
    flat = np.concatenate([Sig[k].ravel(order="F") for k in range(nodes.shape[0])])
    smap = e["map"]
    rows = np.repeat(np.arange(n), np.diff(e["indptr"]))
    m = smap >= 0
    got = flat[smap[m]]
    want = inv[rows[m], e["indices"][m]]
    assert np.allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(inv).max())


@pytest.mark.parametrize("M,N,bc,leaf", [(7, 6, 1, 16), (9, 8, 3, 20), (8, 8, 2, 12), (12, 10, 1, 40)])
def test_one_slice_plan_reproduces_dense_factor_and_inverse(M, N, bc, leaf):
    """One-slice lattices (the spatial matrices of the closed-form
    log-determinant): 25-point pattern, random SPD values on it."""
    pl = Plan(M, N, 1, bc, leaf_max=leaf, nslots=0, host_only=True)
    e = pl.export()
    n = M * N
    assert np.all(e["code"] < 25)
    S = sparse.csr_matrix((np.ones(len(e["indices"])), e["indices"], e["indptr"]), shape=(n, n))
    R = sparse.random(n, n, density=1.0, random_state=3).multiply(S)
    A = sparse.csr_matrix((R @ R.T).multiply(S) + sparse.eye(n) * 50.0)
    logdet, Sig = _mf_factor(e, A)
    dense = A.toarray()
    sign, ref = np.linalg.slogdet(dense)
    assert sign > 0 and abs(logdet - ref) <= 1e-10 * abs(ref)
    inv = np.linalg.inv(dense)
    nodes = e["nodes"]
    flat = np.concatenate([Sig[k].ravel(order="F") for k in range(nodes.shape[0])])
    rows = np.repeat(np.arange(n), np.diff(e["indptr"]))
    m = e["map"] >= 0
    assert np.allclose(flat[e["map"][m]], inv[rows[m], e["indices"][m]], rtol=1e-9, atol=1e-12 * np.abs(inv).max())


@pytest.mark.parametrize("M,N,T,bc,leaf", [(7, 6, 4, 1, 16), (6, 6, 3, 2, 20), (6, 7, 5, 3, 1000)])
def test_wide_time_plan_reproduces_dense_factor_and_inverse(M, N, T, bc, leaf):
    """Wide-time lattices (separable models, Qt (x) Qs): every row repeats the
    one-slice row in slices t-1, t, t+1 and the plan factors and inverts a
    Kronecker matrix on that pattern."""
    pl = Plan(M, N, T, bc, leaf_max=leaf, nslots=0, host_only=True, wide=True)
    sp = Plan(M, N, 1, bc, leaf_max=leaf, nslots=0, host_only=True)
    e, es = pl.export(), sp.export()
    n, Ns = M * N * T, M * N
    Ss = sparse.csr_matrix((np.ones(len(es["indices"])), es["indices"], es["indptr"]), shape=(Ns, Ns))
    Qt = sparse.diags([np.full(T - 1, -0.4), np.full(T, 1.5), np.full(T - 1, -0.4)], [-1, 0, 1])
    S = sparse.csr_matrix((np.ones(len(e["indices"])), e["indices"], e["indptr"]), shape=(n, n))
    assert abs(sparse.kron(Qt, Ss).astype(bool).astype(float) - S).sum() == 0
    R = sparse.random(Ns, Ns, density=1.0, random_state=5).multiply(Ss)
    Qs = sparse.csr_matrix((R @ R.T).multiply(Ss) + sparse.eye(Ns) * 30.0)
    A = sparse.csr_matrix(sparse.kron(Qt, Qs))
    # Q's entry ((t,i),(t+d,j)) at indptr + (t>0) cnt_i + d cnt_i + q (q: position in the one-slice row)
    for r in range(0, n, 7):
        t, i = divmod(r, Ns)
        cnt = es["indptr"][i + 1] - es["indptr"][i]
        for d in (-1, 0, 1):
            if 0 <= t + d < T:
                b = e["indptr"][r] + (cnt if t > 0 else 0) + d * cnt
                assert np.array_equal(e["indices"][b:b + cnt] - (t + d) * Ns,
                                      es["indices"][es["indptr"][i]:es["indptr"][i + 1]])
    logdet, Sig = _mf_factor(e, A)
    dense = A.toarray()
    sign, ref = np.linalg.slogdet(dense)
    assert sign > 0 and abs(logdet - ref) <= 1e-10 * abs(ref)
    inv = np.linalg.inv(dense)
    flat = np.concatenate([Sig[k].ravel(order="F") for k in range(e["nodes"].shape[0])])
    rows = np.repeat(np.arange(n), np.diff(e["indptr"]))
    m = e["map"] >= 0
    assert np.allclose(flat[e["map"][m]], inv[rows[m], e["indices"][m]], rtol=1e-9, atol=1e-12 * np.abs(inv).max())


def test_plan_statistics_100x100x20():
    """100x100x20 lattice: the analysis the GPU factorisation is sized from."""
    pl = Plan(100, 100, 20, 1, leaf_max=128, nslots=0, host_only=True)
    info = pl.info
    assert info["n"] == 200000
    assert info["nnz_struct"] == 8256548 or info["nnz_struct"] > 8.0e6
    assert info["flops_factor"] < 1.5e12
    assert info["max_front"] <= 8000


def test_library_exports_every_declared_symbol():
    import re
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "spdepy_amd.h")).read()
    declared = set(re.findall(r"\b(spde_[a-z0-9_]+)\s*\(", hdr))
    L = nat.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
