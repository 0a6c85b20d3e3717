"""ctypes access to the reference's own stencils compiled by ``oracle/Makefile``
into ``oracle/_ref`` (test infrastructure; used to write and re-check the golden
stencil vectors).  Signatures follow the reference C ABI
(``AcH_2D_b1.cpp:170-185``, ``AH_2D_b1.cpp:154-169``, ``Acw_2D_b1.cpp:117-132``,
``Aw_2D_b1.cpp:136-151``)."""
import ctypes
import os
import numpy as np

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref")
_libs = {}


def available():
    return os.path.exists(os.path.join(_DIR, "lib_AH_2D_b1.so"))


def _lib(name):
    if name not in _libs:
        _libs[name] = ctypes.cdll.LoadLibrary(os.path.join(_DIR, "lib_%s.so" % name))
    return _libs[name]


def _read(lib, prefix, obj, n):
    for fn, ty in (("_Row", ctypes.c_int), ("_Col", ctypes.c_int), ("_Val", ctypes.c_double)):
        getattr(lib, prefix + fn).restype = ctypes.POINTER(ty)
        getattr(lib, prefix + fn).argtypes = [ctypes.c_void_p]
    row = np.ctypeslib.as_array(getattr(lib, prefix + "_Row")(obj), shape=(n,)).copy()
    col = np.ctypeslib.as_array(getattr(lib, prefix + "_Col")(obj), shape=(n,)).copy()
    val = np.ctypeslib.as_array(getattr(lib, prefix + "_Val")(obj), shape=(n,)).copy()
    getattr(lib, prefix + "_delete").argtypes = [ctypes.c_void_p]
    getattr(lib, prefix + "_delete")(obj)
    return row, col, val


def ah(M, N, H, hx, hy, bc, variable):
    lib = _lib(("AH_2D_b%d" if variable else "AcH_2D_b%d") % bc)
    Hc = np.ascontiguousarray(np.array(H, dtype=np.float64))
    lib.AH_new.restype = ctypes.c_void_p
    lib.AH_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    obj = lib.AH_new(M, N, Hc.ctypes.data, hx, hy)
    return _read(lib, "AH", obj, M * N * 9)


def aw(M, N, G, hx, hy, diff, bc, variable, dG=None):
    lib = _lib(("Aw_2D_b%d" if variable else "Acw_2D_b%d") % bc)
    Gc = np.ascontiguousarray(np.array(G, dtype=np.float64))
    lib.Aw_new.restype = ctypes.c_void_p
    if variable:
        dGc = np.zeros((M * N, 4)) if dG is None else np.ascontiguousarray(np.array(dG, dtype=np.float64))
        lib.Aw_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        obj = lib.Aw_new(M, N, Gc.ctypes.data, hx, hy, diff, dGc.ctypes.data)
    else:
        lib.Aw_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int]
        obj = lib.Aw_new(M, N, Gc.ctypes.data, hx, hy, diff)
    return _read(lib, "Aw", obj, M * N * 5)
