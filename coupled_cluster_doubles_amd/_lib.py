"""ctypes binding of libccd_amd.so (include/ccd_amd.h).

The library is built in-tree by coupled_cluster_doubles_amd.build; loading
fails loudly if it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libccd_amd.so")
_lib = None


class CCDError(RuntimeError):
    pass


class Problem(C.Structure):
    _fields_ = [
        ("n_sp", C.c_int32), ("n_holes", C.c_int32), ("eps", C.c_void_p),
        ("n_chan", C.c_int32), ("pp_dim", C.c_void_p), ("hh_dim", C.c_void_p),
        ("pp_pairs", C.c_void_p), ("hh_pairs", C.c_void_p),
        ("n_mod", C.c_int32), ("ph_dim", C.c_void_p), ("hp_dim", C.c_void_p),
        ("ph_pairs", C.c_void_p), ("hp_pairs", C.c_void_p),
        ("v_hhhh", C.c_void_p), ("v_hhpp", C.c_void_p), ("v_pppp", C.c_void_p),
        ("v_modA", C.c_void_p), ("v_modB", C.c_void_p),
        ("gen_kind", C.c_int32), ("qn", C.c_void_p), ("gen_g", C.c_double),
        ("gen_nmax", C.c_int32), ("gen_table", C.c_void_p),
    ]


class LadderPlan(C.Structure):
    """ccd_ladder_plan: per size class of the pppp ladder, the kernel kind (per
    arithmetic), persistent blocks per compute unit and tile assignment, and
    the launch partition; the tuner's timings."""
    _fields_ = [
        ("kind", (C.c_int32 * 8) * 2), ("blocks_per_cu", C.c_int32 * 8), ("assign", C.c_int32 * 8),
        ("partition", C.c_int32), ("us", (C.c_double * 8) * 2), ("us_partition", C.c_double * 2),
        ("ring_fp32", C.c_int32),
    ]


# (name, restype, argtypes) of every exported function in include/ccd_amd.h
vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_void_p
i32p, i64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
SIGNATURES = [
    ("ccd_last_error", C.c_char_p, []),
    ("ccd_ctx_create", C.c_int, [vp, C.c_int, C.POINTER(vp)]),
    ("ccd_ctx_destroy", C.c_int, [vp]),
    ("ccd_ctx_info", C.c_int, [vp, i64p, i64p]),
    ("ccd_set_precision", C.c_int, [vp, C.c_int]),
    ("ccd_set_schedule", C.c_int, [vp, C.c_int]),
    ("ccd_get_schedule", C.c_int, [vp, C.POINTER(C.c_int), ip]),
    ("ccd_get_ladder_plan", C.c_int, [vp, C.POINTER(LadderPlan)]),
    ("ccd_set_ladder_plan", C.c_int, [vp, C.POINTER(LadderPlan)]),
    ("ccd_tune_ladder", C.c_int, [vp, C.c_int]),
    ("ccd_mbpt2", C.c_int, [vp, dp]),
    ("ccd_mbpt2_noalloc", C.c_int, [vp, vp, vp]),
    ("ccd_reset", C.c_int, [vp]),
    ("ccd_iterate", C.c_int, [vp, C.c_double, dp]),
    ("ccd_iterate_n", C.c_int, [vp, C.c_int, C.c_double, ip]),
    ("ccd_solve", C.c_int, [vp, C.c_double, C.c_int, C.c_double, ip, C.c_int,
                            C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("ccd_solve_mixed", C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_double, ip, C.c_int,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("ccd_get_t2", C.c_int, [vp, ip]),
    ("ccd_set_t2", C.c_int, [vp, ip]),
    ("ccd_ladder", C.c_int, [vp, ip, ip]),
    ("ccd_ring", C.c_int, [vp, ip, ip]),
    ("ccd_remap_ph", C.c_int, [vp, ip, ip]),
    ("ccd_onebody", C.c_int, [vp, ip, ip, ip]),
    ("ccd_step", C.c_int, [vp, ip, ip, C.c_double, dp]),
    ("ccd_time_phases", C.c_int, [vp, C.c_int, C.c_double, ip]),
    ("ccd_time_kernels", C.c_int, [vp, C.c_int, C.c_int, C.c_double, ip, C.c_int]),
    ("ccd_work", C.c_int, [vp, dp, dp, dp]),
    ("ccd_kernel_work", C.c_int, [vp, ip, ip, ip, C.c_int]),
    ("ccd_batch_create", C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp)]),
    ("ccd_batch_info", C.c_int, [vp, C.POINTER(C.c_int), ip]),
    ("ccd_batch_mbpt2", C.c_int, [vp, ip]),
    ("ccd_batch_solve", C.c_int, [vp, C.c_double, C.c_int, C.c_double, ip, C.c_int, ip, ip]),
    ("ccd_batch_solve_mixed", C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_double, ip, C.c_int,
                                        ip, ip, ip]),
    ("ccd_batch_iterate_n", C.c_int, [vp, C.c_int, C.c_double, ip]),
    ("ccd_model_pairing", C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(vp)]),
    ("ccd_model_electron_gas", C.c_int, [C.c_double, C.c_int, C.c_int, C.POINTER(vp)]),
    ("ccd_model_inf_matter", C.c_int, [C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    ("ccd_model_destroy", C.c_int, [vp]),
    ("ccd_model_eref", C.c_int, [vp, dp]),
    ("ccd_model_sizes", C.c_int, [vp, i32p, i32p, i32p, i64p]),
    ("ccd_model_problem", C.c_int, [vp, C.POINTER(C.POINTER(Problem))]),
    ("ccd_model_problem_host_v", C.c_int, [vp, C.POINTER(C.POINTER(Problem))]),
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CCDError(f"{LIB_PATH} is not built: run `python -m coupled_cluster_doubles_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().ccd_last_error().decode(errors="replace")
        raise CCDError(f"{what} failed ({rc}): {msg}")
