"""The oracle's stencils against golden vectors written by the reference's own
C++ (compiled from its sources, tests/golden/make_golden.py): bit-exact."""
import os
import numpy as np
import pytest

from oracle import stencils as S

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "stencils.npz"))
NAMES = sorted({k.split("/")[0] for k in GOLD.files if "/" in k})


def _shape(tag):
    dims, b = tag.split("_b")
    M, N = dims.split("x")
    return int(M), int(N), int(b[0])


@pytest.mark.parametrize("name", NAMES)
def test_stencil_matches_reference_bitwise(name):
    kind, rest = name.split("_", 1)
    M, N, bc = _shape(rest)
    hx, hy = float(GOLD["hx"]), float(GOLD["hy"])
    if kind in ("AHv", "AHiso"):
        got = S.ah_var(M, N, GOLD[name + "/in"], hx, hy, bc)
    elif kind in ("AHc", "AHciso"):
        got = S.ah_const(M, N, GOLD[name + "/in"], hx, hy, bc)
    elif kind in ("Awc", "Awc0"):
        diff = int(rest.split("_d")[1])
        got = S.aw_const(M, N, GOLD[name + "/in"], hx, hy, diff, bc)
    else:
        diff = int(rest.split("_d")[1])
        got = S.aw_var(M, N, GOLD[name + "/G"], hx, hy, diff, bc, GOLD[name + "/dG"])
    row, col, val = GOLD[name + "/row"], GOLD[name + "/col"], GOLD[name + "/val"]
    assert np.array_equal(got[0], row)
    keep = row != M * N
    assert np.array_equal(got[1][keep], col[keep])
    assert np.array_equal(got[2][keep], val[keep], equal_nan=True)


def test_stale_index_quirk_only_touches_cell0():
    """AH_2D_b1.cpp:25-52 zeroes the boundary faces of the previously visited
    cell, so only cell 0 gets zero-flux face coefficients.  A stencil with the
    faces zeroed at every boundary cell agrees at cell 0 and differs elsewhere."""
    M, N, hx, hy = 5, 4, 0.9, 1.1
    rng = np.random.default_rng(0)
    H = rng.normal(size=(M * N, 4, 2, 2))
    H = H + H.transpose(0, 1, 3, 2)
    ref = S.ah_var(M, N, H, hx, hy, 1)[2].reshape(M * N, 9)
    Hz = H.copy()
    k = np.arange(M * N)
    i, j = k % M, k // M
    Hz[i == 0, 0, :, 0] = 0.0
    Hz[(i == M - 1), 1, :, 0] = 0.0
    Hz[j == 0, 2, :, 1] = 0.0
    Hz[(j == N - 1), 3, :, 1] = 0.0
    fixed = S.ah_var(M, N, Hz, hx, hy, 1)[2].reshape(M * N, 9)
    assert np.array_equal(ref[0], fixed[0])
    boundary = (i == 0) | (i == M - 1) | (j == 0) | (j == N - 1)
    differs = np.any(ref != fixed, axis=1)
    assert differs[boundary & (k != 0)].any()
    assert not differs[~boundary].any()
