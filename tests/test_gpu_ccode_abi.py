"""The stencil libraries with the reference's C interface
(spdepy_amd/ccode/lib_{AcH,AH,Acw,Aw}_2D_b{1,2,3}.so), bound exactly as the
reference's setClib binds its own (advection_idiffusion2D.py:246-300,
var_advection_var_diffusion2D.py:290-330: the same argtypes and restypes),
give the reference's golden triplets bit for bit."""
import ctypes
import os

import numpy as np
import pytest

from tests.gpu_util import step

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "stencils.npz"))
CCODE = os.path.join(os.path.dirname(HERE), "spdepy_amd", "ccode")


def _bind(kind, bc, M, N):
    """The reference's ctypes binding of one stencil library (setClib)."""
    lib = ctypes.cdll.LoadLibrary(os.path.join(CCODE, "lib_%s_2D_b%d.so" % (kind, bc)))
    if kind in ("AcH", "AH"):
        new = lib.AH_new
        shape = (2, 2) if kind == "AcH" else (M * N, 4, 2, 2)
        new.argtypes = [ctypes.c_int, ctypes.c_int, np.ctypeslib.ndpointer(dtype=np.float64, ndim=len(shape),
                                                                          shape=shape), ctypes.c_double,
                        ctypes.c_double]
        names, slots = ("AH_Row", "AH_Col", "AH_Val", "AH_delete"), 9
    else:
        new = lib.Aw_new
        if kind == "Acw":
            new.argtypes = [ctypes.c_int, ctypes.c_int, np.ctypeslib.ndpointer(dtype=np.float64, ndim=1, shape=(2,)),
                            ctypes.c_double, ctypes.c_double, ctypes.c_int]
        else:
            arr = np.ctypeslib.ndpointer(dtype=np.float64, ndim=2, shape=(M * N, 4))
            new.argtypes = [ctypes.c_int, ctypes.c_int, arr, ctypes.c_double, ctypes.c_double, ctypes.c_int, arr]
        names, slots = ("Aw_Row", "Aw_Col", "Aw_Val", "Aw_delete"), 5
    new.restype = ctypes.c_void_p
    row, col, val, dele = (getattr(lib, nm) for nm in names)
    for f, dt in ((row, ctypes.c_int), (col, ctypes.c_int), (val, ctypes.c_double)):
        f.argtypes = [ctypes.c_void_p]
        f.restype = np.ctypeslib.ndpointer(dtype=dt, shape=(M * N * slots,))
    dele.argtypes = [ctypes.c_void_p]
    dele.restype = None
    return new, row, col, val, dele


def test_reference_bindings_give_the_golden_triplets():
    # a process that also uses torch initialises torch's HIP runtime before
    # the stencil libraries bring in theirs (the same order the main library
    # keeps in spdepy_amd._native.lib)
    import torch
    assert torch.cuda.is_available()
    torch.cuda.init()
    hx, hy = float(GOLD["hx"]), float(GOLD["hy"])
    names = sorted({k.split("/")[0] for k in GOLD.files if "/" in k})
    bad = []
    for name in names:
        kind, rest = name.split("_", 1)
        dims, b = rest.split("_b")
        M, N = map(int, dims.split("x"))
        bc = int(b[0])
        lib_kind = {"AHv": "AH", "AHiso": "AH", "AHc": "AcH", "AHciso": "AcH", "Awc": "Acw", "Awc0": "Acw",
                    "Awv": "Aw"}[kind]
        step("reference binding " + name)
        new, row, col, val, dele = _bind(lib_kind, bc, M, N)
        if lib_kind in ("AcH", "AH"):
            H = np.ascontiguousarray(GOLD[name + "/in"], dtype=np.float64)
            H = H.reshape((2, 2) if lib_kind == "AcH" else (M * N, 4, 2, 2))
            obj = new(M, N, H, hx, hy)
        elif lib_kind == "Acw":
            obj = new(M, N, np.ascontiguousarray(GOLD[name + "/in"], dtype=np.float64).reshape(2), hx, hy,
                      int(rest.split("_d")[1]))
        else:
            G = np.ascontiguousarray(GOLD[name + "/G"], dtype=np.float64).reshape(M * N, 4)
            dG = np.ascontiguousarray(GOLD[name + "/dG"], dtype=np.float64).reshape(M * N, 4)
            obj = new(M, N, G, hx, hy, int(rest.split("_d")[1]), dG)
        assert obj, name
        r, c, v = row(obj).copy(), col(obj).copy(), val(obj).copy()
        dele(obj)
        g_row = GOLD[name + "/row"]
        keep = g_row != M * N
        ok = (np.array_equal(r, g_row) and np.array_equal(c[keep], GOLD[name + "/col"][keep])
              and np.array_equal(v[keep], GOLD[name + "/val"][keep], equal_nan=True))
        if not ok:
            bad.append(name)
    assert not bad, bad
