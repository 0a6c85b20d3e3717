"""Unit checks of the dense FP64 kernels (MFMA layout, batched GEMM)."""
import numpy as np
import pytest

from tests.gpu_util import step

pytestmark = pytest.mark.gpu


def _lib():
    from spdepy_amd import _native as nat
    return nat


def test_mfma_f64_lane_layout():
    import torch
    nat = _lib()
    out = torch.zeros(256, dtype=torch.float64, device="cuda")
    res = []
    for which in (0, 1):
        step("mfma probe %d" % which)
        nat.check(nat.lib().spde_debug_mfma_probe(which, out.data_ptr(), None))
        torch.cuda.synchronize()
        res.append(out.cpu().numpy().reshape(64, 4).copy())
    lane = np.arange(64)[:, None]
    v = np.arange(4)[None, :]
    # D[i][j] = (j+1) (probe 0) and (i+1) (probe 1): lane l, register v holds
    # D[l//16 + 4v][l%16]
    assert np.array_equal(res[0], np.broadcast_to(lane % 16 + 1, (64, 4)).astype(float))
    assert np.array_equal(res[1], (lane // 16 + 4 * v + 1).astype(float))


@pytest.mark.parametrize("tile", [64, 128])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 37, 70), (5, 130, 3), (129, 65, 200), (300, 260, 33),
                                   (1100, 1060, 40)])
def test_gemm_matches_numpy(ta, tb, M, N, K, tile):
    """The last shape has more than 64 tiles and a partial last group of 64
    (306 tiles of 64, 81 of 128): the XCD-grouped block-to-tile mapping must
    cover every tile once."""
    import torch
    nat = _lib()
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    C = rng.normal(size=(M, N))
    want = -1.5 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.5 * C
    At = torch.as_tensor(np.asfortranarray(A).ravel(order="F")).cuda()
    Bt = torch.as_tensor(np.asfortranarray(B).ravel(order="F")).cuda()
    Ct = torch.as_tensor(np.asfortranarray(C).ravel(order="F")).cuda()
    step("gemm %d%d %dx%dx%d tile %d" % (ta, tb, M, N, K, tile))
    nat.check(nat.lib().spde_debug_gemm(ta, tb, M, N, K, -1.5, At.data_ptr(), A.shape[0], Bt.data_ptr(),
                                        B.shape[0], 0.5, Ct.data_ptr(), M, tile, None))
    got = Ct.cpu().numpy().reshape(N, M).T
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
