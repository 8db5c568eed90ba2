"""ctypes binding of libqlgpu.so (the C ABI declared in include/qlgpu.h).

The library is built in-tree by ``python -m ultradmrg_amd.build``; importing
this module without it raises immediately -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqlgpu.so")

QLG_OK = 0
QLG_F64 = 0
QLG_C128 = 1
QLG_PREC_F64 = 0
QLG_PREC_F32 = 1
QLG_PREC_BF16X3 = 2
QLG_PREC_BF16X6 = 3

ERRORS = {1: "QLG_ERR_INVALID", 2: "QLG_ERR_HIP", 3: "QLG_ERR_ALLOC", 4: "QLG_ERR_UNSUPPORTED",
          5: "QLG_ERR_NUMERIC"}

# every entry point declared in include/qlgpu.h
EXPORTS = [
    "qlg_ctx_create", "qlg_ctx_destroy", "qlg_last_error", "qlg_ctx_synchronize",
    "qlg_ctx_release_cached", "qlg_ctx_memory_stats",
    "qlg_device_alloc", "qlg_device_free", "qlg_memcpy_h2d", "qlg_memcpy_d2h", "qlg_memcpy_d2d",
    "qlg_plan_two_site", "qlg_plan_one_site", "qlg_plan_super_block", "qlg_plan_destroy", "qlg_plan_set_precision",
    "qlg_plan_flops", "qlg_plan_vector_layout", "qlg_plan_workspace", "qlg_plan_enable_timing",
    "qlg_plan_stage_ms", "qlg_plan_stage_flops", "qlg_plan_stage_work", "qlg_matvec", "qlg_plan_contract", "qlg_plan_env",
    "qlg_plan_transpose", "qlg_execute", "qlg_dot", "qlg_nrm2", "qlg_scal", "qlg_axpy",
    "qlg_block_svd", "qlg_svd_result", "qlg_svd_bond", "qlg_svd_tensor", "qlg_svd_destroy",
    "qlg_linear_combine", "qlg_linear_combine_noalloc", "qlg_lanczos_gs", "qlg_lanczos_expmv", "qlg_eigh_batched", "qlg_plan_batch", "qlg_plan_batch_offsets",
    "qlg_plan_use_graphs", "qlg_plan_set_reorth", "qlg_lanczos_gs_batch", "qlg_tridiag_gs",
    "qlg_tridiag_expme1",
]


class QlgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Bsten(C.Structure):
    _fields_ = [
        ("rank", C.c_int32), ("qn_len", C.c_int32),
        ("nsect", C.POINTER(C.c_int32)), ("sect_qn", C.POINTER(C.c_int32)),
        ("sect_deg", C.POINTER(C.c_int64)), ("dir", C.POINTER(C.c_int8)),
        ("nblocks", C.c_int64), ("blk_coords", C.POINTER(C.c_int32)),
        ("blk_offset", C.POINTER(C.c_int64)), ("dtype", C.c_int32), ("data", C.c_void_p),
    ]


class Term(C.Structure):
    _fields_ = [("op0", C.POINTER(Bsten)), ("op1", C.POINTER(Bsten))]


class EighGroup(C.Structure):
    _fields_ = [("n", C.c_int64), ("batch", C.c_int64), ("a", C.c_void_p), ("w", C.c_void_p)]


class LanczosParams(C.Structure):
    _fields_ = [("error", C.c_double), ("max_iterations", C.c_int64)]


class LanczosResult(C.Structure):
    _fields_ = [("iters", C.c_int64), ("gs_eng", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ultradmrg_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "qlg_ctx_create": (C.c_int, [C.c_int, vp, C.POINTER(vp)]),
        "qlg_ctx_destroy": (C.c_int, [vp]),
        "qlg_last_error": (C.c_char_p, [vp]),
        "qlg_ctx_synchronize": (C.c_int, [vp]),
        "qlg_ctx_release_cached": (C.c_int, [vp]),
        "qlg_ctx_memory_stats": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "qlg_device_alloc": (C.c_int, [vp, i64, C.POINTER(vp)]),
        "qlg_device_free": (C.c_int, [vp, vp]),
        "qlg_memcpy_h2d": (C.c_int, [vp, vp, vp, i64]),
        "qlg_memcpy_d2h": (C.c_int, [vp, vp, vp, i64]),
        "qlg_memcpy_d2d": (C.c_int, [vp, vp, vp, i64]),
        "qlg_plan_two_site": (C.c_int, [vp] + [C.POINTER(Bsten)] * 5 + [C.c_int, C.POINTER(vp)]),
        "qlg_plan_one_site": (C.c_int, [vp] + [C.POINTER(Bsten)] * 4 + [C.c_int, C.POINTER(vp)]),
        "qlg_plan_super_block": (C.c_int, [vp, i64, C.POINTER(i64), C.POINTER(Term), C.POINTER(i64),
                                           C.POINTER(Term), C.POINTER(Bsten), C.c_int, C.POINTER(vp)]),
        "qlg_plan_destroy": (C.c_int, [vp]),
        "qlg_plan_set_precision": (C.c_int, [vp, C.c_int]),
        "qlg_plan_flops": (C.c_int, [vp, dp, dp]),
        "qlg_plan_vector_layout": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(C.c_int32),
                                             C.POINTER(i64)]),
        "qlg_plan_workspace": (C.c_int, [vp, C.POINTER(i64)]),
        "qlg_plan_enable_timing": (C.c_int, [vp, C.c_int]),
        "qlg_plan_stage_ms": (C.c_int, [vp, dp]),
        "qlg_plan_stage_flops": (C.c_int, [vp, dp]),
        "qlg_plan_stage_work": (C.c_int, [vp, dp, dp]),
        "qlg_matvec": (C.c_int, [vp, vp, vp]),
        "qlg_plan_contract": (C.c_int, [vp, C.POINTER(Bsten), C.POINTER(Bsten), C.c_int, C.c_int, C.c_int,
                                        C.POINTER(vp)]),
        "qlg_plan_env": (C.c_int, [vp, C.POINTER(Bsten), C.POINTER(Bsten), C.POINTER(Bsten), C.c_char,
                                   C.POINTER(vp)]),
        "qlg_plan_transpose": (C.c_int, [vp, C.POINTER(Bsten), C.POINTER(C.c_int32), C.POINTER(vp)]),
        "qlg_execute": (C.c_int, [vp, vp, vp, vp]),
        "qlg_dot": (C.c_int, [vp, C.c_int, i64, vp, vp, dp]),
        "qlg_nrm2": (C.c_int, [vp, C.c_int, i64, vp, dp]),
        "qlg_scal": (C.c_int, [vp, C.c_int, i64, dp, vp]),
        "qlg_axpy": (C.c_int, [vp, C.c_int, i64, dp, vp, vp]),
        "qlg_linear_combine": (C.c_int, [vp, C.c_int, i64, i64, dp, C.POINTER(vp), dp, vp]),
        "qlg_linear_combine_noalloc": (C.c_int, [vp, C.c_int, i64, i64, dp, C.POINTER(vp), dp, vp]),
        "qlg_block_svd": (C.c_int, [vp, vp, C.c_int32, C.POINTER(C.c_int32), C.c_double, i64, i64, C.POINTER(vp)]),
        "qlg_svd_result": (C.c_int, [vp, dp, C.POINTER(i64), C.POINTER(i64)]),
        "qlg_svd_bond": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(i64)]),
        "qlg_svd_tensor": (C.c_int, [vp, C.c_int, C.POINTER(i64), C.POINTER(i64), C.POINTER(C.c_int32),
                                     C.POINTER(i64), C.POINTER(vp)]),
        "qlg_svd_destroy": (C.c_int, [vp]),
        "qlg_lanczos_gs": (C.c_int, [vp, vp, C.POINTER(LanczosParams), C.POINTER(LanczosResult)]),
        "qlg_lanczos_expmv": (C.c_int, [vp, vp, C.c_double, C.POINTER(LanczosParams), C.POINTER(i64)]),
        "qlg_eigh_batched": (C.c_int, [vp, C.c_int, i64, C.POINTER(EighGroup)]),
        "qlg_plan_batch": (C.c_int, [vp, i64, C.POINTER(vp), C.POINTER(vp)]),
        "qlg_plan_batch_offsets": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64)]),
        "qlg_plan_use_graphs": (C.c_int, [vp, C.c_int]),
        "qlg_plan_set_reorth": (C.c_int, [vp, C.c_int]),
        "qlg_lanczos_gs_batch": (C.c_int, [vp, vp, C.POINTER(LanczosParams), C.POINTER(LanczosResult)]),
        "qlg_tridiag_gs": (C.c_int, [dp, dp, i64, C.c_char, dp, dp]),
        "qlg_tridiag_expme1": (C.c_int, [dp, dp, i64, C.c_double, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc != QLG_OK:
        msg = lib().qlg_last_error(ctx)
        raise QlgError(rc, msg.decode() if msg else "")
