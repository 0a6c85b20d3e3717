"""ctypes binding of libspdepy_amd.so (the C ABI declared in include/spdepy_amd.h).

The library is built in place by ``spdepy_amd.build``; importing a compute
entry point without it raises -- there is no CPU fallback."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspdepy_amd.so")

SPDE_OK, SPDE_ERR_ARG, SPDE_ERR_HIP, SPDE_ERR_NOT_SPD, SPDE_ERR_ALLOC, SPDE_ERR_STATE = range(6)
SPDE_PLAN_HOST_ONLY = 1
SPDE_PLAN_WIDE_TIME = 2
SOLVE_A, SOLVE_L, SOLVE_LT, SOLVE_P, SOLVE_PT, SOLVE_L_NOPERM, SOLVE_LT_NOPERM = 0, 1, 2, 3, 4, 5, 6

c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_dp = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


class SpdeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class NotPositiveDefinite(SpdeError):
    pass


class PlanInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz_struct", ctypes.c_int64), ("nnz_L", ctypes.c_int64),
                ("front_doubles", ctypes.c_int64), ("z_doubles", ctypes.c_int64),
                ("nnodes", ctypes.c_int64), ("nlevels", ctypes.c_int64), ("max_front", ctypes.c_int64),
                ("max_p", ctypes.c_int64), ("flops_factor", ctypes.c_double),
                ("flops_selinv", ctypes.c_double), ("split_doubles", ctypes.c_int64)]


class Coeffs(ctypes.Structure):
    _fields_ = [("variable", ctypes.c_int), ("hx", ctypes.c_double), ("hy", ctypes.c_double),
                ("dt", ctypes.c_double), ("V", ctypes.c_double), ("sigma", ctypes.c_double),
                ("kappa_c", ctypes.c_double), ("H_c", ctypes.c_double * 4), ("G_c", ctypes.c_double * 2),
                ("kappa", vp), ("H", vp), ("G", vp), ("kappa0", ctypes.c_double),
                ("H0", ctypes.c_double * 4), ("a0_variable", ctypes.c_int),
                ("mix", ctypes.c_int)]


class Direction(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dkappa", vp), ("dkappa_c", ctypes.c_double), ("dH", vp),
                ("dH_c", ctypes.c_double * 4), ("has_dH", ctypes.c_int), ("dG", vp), ("adv_diff", ctypes.c_int),
                ("dkappa0", ctypes.c_double), ("dH0", ctypes.c_double * 4)]


class LoglikOut(ctypes.Structure):
    _fields_ = [("like", ctypes.c_double), ("like_raw", ctypes.c_double), ("logdet_q", ctypes.c_double),
                ("logdet_qc", ctypes.c_double), ("g_sigma", ctypes.c_double), ("g_tau", ctypes.c_double)]


class GradArgs(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("r", ctypes.c_int), ("nh1", ctypes.c_int), ("qval", vp),
                ("sigma_q", vp), ("sigma_qc", vp), ("trq", vp), ("trqc", vp), ("probes", vp),
                ("mu", vp), ("diag_add", vp), ("coeffs", Coeffs),
                ("sigma_gram", vp), ("sigma_q0", vp), ("spatial", vp), ("weights", vp), ("adv_dir", vp)]


_SIGS = {
    "spde_last_error": (ctypes.c_char_p, []),
    "spde_stencil_diffusion": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, vp, vp, vp, vp]),
    "spde_stencil_advection": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]),
    "spde_plan_create": (ctypes.c_int, [ctypes.c_int] * 7 + [ctypes.POINTER(vp)]),
    "spde_plan_destroy": (None, [vp]),
    "spde_plan_get_info": (ctypes.c_int, [vp, ctypes.POINTER(PlanInfo)]),
    "spde_plan_export": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "spde_assemble": (ctypes.c_int, [vp, ctypes.POINTER(Coeffs), vp, vp]),
    "spde_assemble_spatial": (ctypes.c_int, [vp, vp, ctypes.c_int, vp, vp]),
    "spde_assemble_tangent": (ctypes.c_int, [vp, ctypes.POINTER(Coeffs), ctypes.POINTER(Direction), vp, vp]),
    "spde_logq_sum": (ctypes.c_int, [vp, ctypes.c_double, c_dp, vp]),
    "spde_assemble_kron": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "spde_loglik_grad": (ctypes.c_int, [vp, vp, ctypes.POINTER(Coeffs), vp, vp, ctypes.c_int64, ctypes.c_int,
                                        ctypes.c_double, vp, ctypes.POINTER(LoglikOut), vp]),
    "spde_loglik_enqueue": (ctypes.c_int, [vp, vp, ctypes.POINTER(Coeffs), vp, vp, ctypes.c_int64, ctypes.c_int,
                                           ctypes.c_double, vp, vp]),
    "spde_loglik_finish": (ctypes.c_int, [vp, ctypes.POINTER(LoglikOut)]),
    "spde_loglik_mu": (ctypes.c_int, [vp, vp, ctypes.c_int, vp]),
    "spde_loglik_batch": (ctypes.c_int, [ctypes.c_int, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp, vp,
                                         ctypes.c_int, vp]),
    "spde_kron_reduce": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_double, vp, vp, ctypes.c_int, vp, vp, vp,
                                        vp]),
    "spde_fields_eval": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int] + [vp] * 8 + [vp]),
    "spde_fields_contract_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, c_i64p]),
    "spde_fields_contract": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int] + [vp] * 8 + [vp]),
    "spde_fields_contract_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int] + [vp] * 7 + [vp]),
    "spde_prune": (ctypes.c_int, [vp, vp, vp, vp, vp, c_i64p, vp]),
    "spde_spmv": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int, vp]),
    "spde_factor": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, c_dp, vp]),
    "spde_factor_async": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp]),
    "spde_factor_finish": (ctypes.c_int, [vp, ctypes.c_int, c_dp]),
    "spde_solve": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int64, vp]),
    "spde_selinv": (ctypes.c_int, [vp, ctypes.c_int, vp, vp]),
    "spde_prof_begin": (ctypes.c_int, [vp]),
    "spde_prof_begin_kinds": (ctypes.c_int, [vp, ctypes.c_int]),
    "spde_prof_end": (ctypes.c_int, [vp, c_dp, c_dp, c_i64p, c_dp, c_dp]),
    "spde_prof_dump": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, ctypes.c_int64, c_i64p]),
    "spde_debug_mfma_probe": (ctypes.c_int, [ctypes.c_int, vp, vp]),
    "spde_debug_gemm": (ctypes.c_int, [ctypes.c_int] * 5 + [ctypes.c_double, vp, ctypes.c_int, vp, ctypes.c_int,
                                                            ctypes.c_double, vp, ctypes.c_int, ctypes.c_int, vp]),
    "spde_grad_cells": (ctypes.c_int, [vp, ctypes.POINTER(GradArgs), vp, vp]),
    "spde_ic_create": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "spde_ic_view": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    "spde_ic_destroy": (None, [vp]),
    "spde_ic_info": (ctypes.c_int, [vp, c_i64p, c_i32p]),
    "spde_ic_export": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "spde_ic_factor": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, c_dp, vp]),
    "spde_ic_correct": (ctypes.c_int, [vp, vp, c_dp, vp]),
    "spde_ic_values": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "spde_ic_apply": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, vp]),
    "spde_ic_sweep": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]),
    "spde_ic_spmm": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, vp]),
    "spde_pcg_solve": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, vp, vp]),
    "spde_pcg_refine": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       c_i32p, c_i32p, c_dp, vp]),
    "spde_pcg_replace": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, c_i32p, c_i32p,
                                        c_dp, vp]),
    "spde_pcg_replace_block": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_int, vp, c_i32p, vp, vp]),
    "spde_pcg_solve_coeffs": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, vp, vp, vp,
                                             ctypes.c_int, vp]),
    "spde_ic_lmul": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, vp]),
    "spde_rademacher": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, ctypes.c_float, vp, vp]),
    "spde_pcg_refine_block": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]),
    "spde_gaussian": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, vp, vp]),
    "spde_fform_apply_t": (ctypes.c_int, [vp, ctypes.POINTER(Coeffs), vp, vp, ctypes.c_int, vp]),
    "spde_add_rows": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int64, vp, vp, vp, vp]),
    "spde_slq": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, vp, vp, vp]),
    "spde_hutch_pattern": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp]),
    "spde_ic_features": (ctypes.c_int, [vp, vp, vp]),
    "spde_ic_kl_grad": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_double, vp, c_dp, vp]),
    "spde_ic_jeffreys_grad": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_double, vp, c_dp, vp]),
    "spde_conv_packed_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, c_i64p]),
    "spde_conv_pack": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, vp, vp, vp]),
    "spde_conv_forward": (ctypes.c_int, [ctypes.c_int] * 4 + [vp, vp, vp, vp, ctypes.c_int, vp, vp]),
    "spde_conv_wgrad_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, c_i64p]),
    "spde_conv_backward": (ctypes.c_int, [ctypes.c_int] * 4 + [vp] * 7),
    "spde_adam": (ctypes.c_int, [ctypes.c_int64, vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_int, vp, vp, vp, vp]),
    "spde_debug_bf16_probe": (ctypes.c_int, [vp, vp]),
}

_lib = None


def lib():
    """Load the native library; raises if it is not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libspdepy_amd.so is not built; run `python -m spdepy_amd.build`")
        # torch brings its own HIP runtime: it is loaded and initialised before
        # this library makes its first HIP call, so that torch's device
        # initialisation (which then shares the device with the library's
        # allocations and streams) always comes first in the process.
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return [n for n in _SIGS if hasattr(L, n)]


def check(rc):
    if rc != SPDE_OK:
        msg = lib().spde_last_error().decode()
        if rc == SPDE_ERR_NOT_SPD:
            raise NotPositiveDefinite(rc, msg)
        raise SpdeError(rc, msg)
