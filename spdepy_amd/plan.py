"""Python handle of a native plan: structural pattern, nested-dissection
ordering and multifrontal schedule of Q for one (M x N x T, bc) lattice, plus
the device arenas of its factorisations.  Buffers are torch tensors on the
GPU; only their data pointers cross the C ABI."""
import ctypes

import numpy as np

from . import _native as nat


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Plan:
    def __init__(self, M, N, T, bc, leaf_max=128, nslots=2, host_only=False, wide=False):
        """wide=True: neighbouring time slices coupled through the 5x5 window
        (separable models, values from assemble_kron)."""
        self.shape = (M, N, T)
        self.bc = bc
        self.wide = wide
        L = nat.lib()
        h = ctypes.c_void_p()
        flags = (nat.SPDE_PLAN_HOST_ONLY if host_only else 0) | (nat.SPDE_PLAN_WIDE_TIME if wide else 0)
        nat.check(L.spde_plan_create(M, N, T, bc, leaf_max, nslots, flags, ctypes.byref(h)))
        self._h = h
        self.host_only = host_only
        info = nat.PlanInfo()
        nat.check(L.spde_plan_get_info(h, ctypes.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in nat.PlanInfo._fields_}
        self.n = int(info.n)
        self.nnz = int(info.nnz_struct)
        self._struct = None
        self.slot_owner = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is None or not h.value:
            return
        self._h = None
        try:
            nat.lib().spde_plan_destroy(h)
        except (AttributeError, TypeError):  # at interpreter exit the modules are gone; the process frees the device
            pass

    # ------------------------------------------------------------ host data
    def export(self):
        """Structure and symbolic analysis as numpy arrays (host)."""
        n, nnz, nn = self.n, self.nnz, int(self.info["nnodes"])
        indptr = np.zeros(n + 1, np.int64)
        indices = np.zeros(nnz, np.int32)
        code = np.zeros(nnz, np.uint8)
        perm = np.zeros(n, np.int32)
        nodes = np.zeros((nn, 8), np.int32)
        foff = np.zeros(nn, np.int64)
        nb_total = None
        # boundary totals: first call with nodes only
        nat.check(nat.lib().spde_plan_export(self._h, indptr.ctypes.data, indices.ctypes.data, code.ctypes.data,
                                             perm.ctypes.data, nodes.ctypes.data, foff.ctypes.data, None, None,
                                             None))
        nb_total = int(nodes[:, 7].sum())
        bnd = np.zeros(max(nb_total, 1), np.int32)
        relpos = np.zeros(max(nb_total, 1), np.int32)
        smap = np.zeros(nnz, np.int64)
        nat.check(nat.lib().spde_plan_export(self._h, None, None, None, None, None, None, bnd.ctypes.data,
                                             relpos.ctypes.data, smap.ctypes.data))
        off = np.concatenate([[0], np.cumsum(nodes[:, 7])]).astype(np.int64)
        return dict(indptr=indptr, indices=indices, code=code, perm=perm, nodes=nodes, foff=foff,
                    bnd=[bnd[off[k]:off[k + 1]] for k in range(nn)],
                    relpos=[relpos[off[k]:off[k + 1]] for k in range(nn)], map=smap)

    def structure(self):
        if self._struct is None:
            e = self.export()
            self._struct = (e["indptr"], e["indices"])
        return self._struct

    # ------------------------------------------------------------ device ops
    def assemble(self, coeffs, qval):
        import torch
        nat.check(nat.lib().spde_assemble(self._h, ctypes.byref(coeffs), _ptr(qval), _stream(torch)))

    def prune(self, qval):
        import torch
        dev = qval.device
        indptr = torch.empty(self.n + 1, dtype=torch.int64, device=dev)
        indices = torch.empty(self.nnz, dtype=torch.int32, device=dev)
        data = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        nnz = ctypes.c_int64()
        nat.check(nat.lib().spde_prune(self._h, _ptr(qval), _ptr(indptr), _ptr(indices), _ptr(data),
                                       ctypes.byref(nnz), _stream(torch)))
        k = int(nnz.value)
        return indptr, indices[:k], data[:k]

    def spmv(self, qval, x):
        import torch
        y = torch.empty_like(x)
        nrhs = 1 if x.dim() == 1 else x.shape[1]
        xt = x if x.dim() == 1 else x.t().contiguous()
        yt = torch.empty_like(xt)
        nat.check(nat.lib().spde_spmv(self._h, _ptr(qval), _ptr(xt), _ptr(yt), nrhs, _stream(torch)))
        return yt if x.dim() == 1 else yt.t()

    def factor(self, slot, qval, diag_add=None, owner=None):
        import torch
        ld = ctypes.c_double()
        self.slot_owner.pop(slot, None)   # the slot's contents change whatever the outcome
        nat.check(nat.lib().spde_factor(self._h, slot, _ptr(qval), _ptr(diag_add), ctypes.byref(ld),
                                        _stream(torch)))
        self.slot_owner[slot] = owner     # only a successful factorisation has an owner
        return ld.value

    def assemble_spatial(self, st_plan, which, out=None):
        """One-slice plan: Q0 (which=0) or B^T B (which=1) of the last
        assembly on the space-time plan `st_plan`."""
        import torch
        if out is None:
            out = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        nat.check(nat.lib().spde_assemble_spatial(st_plan._h, self._h, which, _ptr(out), _stream(torch)))
        return out

    def logq_sum(self, c):
        """sum_k log(q_k c) for the last assembly (synchronises the stream)."""
        import torch
        v = ctypes.c_double()
        nat.check(nat.lib().spde_logq_sum(self._h, float(c), ctypes.byref(v), _stream(torch)))
        return v.value

    def factor_async(self, slot, qval, diag_add=None, owner=None):
        """Enqueues the factorisation on the current stream; factor_finish
        waits for it and returns the log-determinant."""
        import torch
        self.slot_owner.pop(slot, None)
        nat.check(nat.lib().spde_factor_async(self._h, slot, _ptr(qval), _ptr(diag_add), _stream(torch)))
        self._pending_owner = getattr(self, "_pending_owner", {})
        self._pending_owner[slot] = owner

    def factor_finish(self, slot):
        ld = ctypes.c_double()
        owner = getattr(self, "_pending_owner", {}).pop(slot, None)
        nat.check(nat.lib().spde_factor_finish(self._h, slot, ctypes.byref(ld)))
        self.slot_owner[slot] = owner
        return ld.value

    def solve(self, slot, B, mode=nat.SOLVE_A):
        """B: (n,) or (n, k) float64 cuda tensor; returns a new tensor."""
        import torch
        one = B.dim() == 1
        X = (B.reshape(-1, 1) if one else B).t().contiguous()  # (k, n) row-major = column-major (n, k)
        nat.check(nat.lib().spde_solve(self._h, slot, mode, _ptr(X), X.shape[0], self.n, _stream(torch)))
        return X.reshape(-1) if one else X.t()

    def selinv(self, slot, out=None):
        import torch
        if out is None:
            out = torch.empty(self.nnz, dtype=torch.float64, device="cuda")
        nat.check(nat.lib().spde_selinv(self._h, slot, _ptr(out), _stream(torch)))
        return out
