"""spde_fields_contract_noalloc (static device scratch, no allocation) against
spde_fields_contract (caller workspace) through the C ABI: the same kernels in
the same order, so the outputs are bitwise equal.  Two calls in a row, the
second on a smaller lattice and basis (the scratch keeps the first call's
contents)."""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import step

pytestmark = pytest.mark.gpu


def _case(torch, Ns, nb, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)

    def dev(*shape):
        return torch.randn(*shape, generator=g, dtype=torch.float64).cuda()
    return dict(bs=dev(Ns, nb), bsH=dev(Ns, 4, nb), bsA=dev(Ns, 4, nb), cells=dev(Ns, 40),
                kappa=dev(Ns).abs() + 0.5, fgv=dev(Ns, 4, 3))


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_workspace_entry_point(order):
    import torch
    from spdepy_amd import _native as nat
    L = nat.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for Ns, nb, seed in [(10000, 64, 1), (900, 9, 2)]:
        step("fields_contract noalloc vs workspace: Ns=%d nb=%d" % (Ns, nb))
        c = _case(torch, Ns, nb, seed)
        wd = ctypes.c_int64()
        nat.check(L.spde_fields_contract_workspace(Ns, nb, ctypes.byref(wd)))
        work = torch.empty(wd.value, dtype=torch.float64, device="cuda")
        ref = torch.full((6 * nb + 40,), np.nan, dtype=torch.float64, device="cuda")
        out = torch.full((6 * nb + 40,), np.nan, dtype=torch.float64, device="cuda")
        args = [_p(c[k]) for k in ("bs", "bsH", "bsA", "cells", "kappa", "fgv")]
        nat.check(L.spde_fields_contract(Ns, nb, *args, _p(work), _p(ref), stream))
        nat.check(L.spde_fields_contract_noalloc(Ns, nb, *args, _p(out), stream))
        torch.cuda.synchronize()
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        assert np.isfinite(r).all()
        assert np.array_equal(r, o), np.abs(r - o).max()
        # and against a float64 host contraction of the same maps
        h = {k: v.cpu().numpy() for k, v in c.items()}
        ck = h["cells"]
        g, vx, vy = h["fgv"][:, :, 0], h["fgv"][:, :, 1], h["fgv"][:, :, 2]
        h00, h01, h10, h11 = ck[:, 1:17:4], ck[:, 2:17:4], ck[:, 3:17:4], ck[:, 4:17:4]
        exp = np.concatenate([
            h["bs"].T @ (ck[:, 0] * h["kappa"]),
            np.einsum("kfj,kf->j", h["bsH"], (h00 + h11) * g),
            np.einsum("kfj,kf->j", h["bsH"], 2 * vx * h00 + vy * (h01 + h10)),
            np.einsum("kfj,kf->j", h["bsH"], 2 * vy * h11 + vx * (h01 + h10)),
            h["bsA"][:, 0].T @ ck[:, 17] + h["bsA"][:, 2].T @ ck[:, 19],
            h["bsA"][:, 1].T @ ck[:, 22] + h["bsA"][:, 3].T @ ck[:, 24],
            ck.sum(0)])
        assert np.allclose(o, exp, rtol=1e-11, atol=1e-11 * np.abs(exp).max())
