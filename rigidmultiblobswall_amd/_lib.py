"""ctypes binding of librmw_hip.so (the C ABI declared in include/rmw_hip.h).

The library is built in-tree by ``rigidmultiblobswall_amd.build``.  There is
no fallback: if the shared object is missing or a symbol is absent, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmw_hip.so")

RMW_OK = 0
RMW_UF, RMW_UT, RMW_WF, RMW_WT = 1, 2, 4, 8
RMW_PREC_F32_F64ACC, RMW_PREC_F32_KAHAN, RMW_PREC_F64, RMW_PREC_F32_BLOCKED = 0, 1, 2, 3
RMW_E_ARG, RMW_E_HIP, RMW_E_WORKSPACE, RMW_E_NOTFINITE, RMW_E_LIMIT = -1, -2, -3, -4, -5
RMW_FLAG_CHECK_FINITE = 1
RMW_NOALLOC_SCRATCH_BYTES = 64 << 20     # include/rmw_hip.h

PRECISIONS = {
    "fp32": RMW_PREC_F32_F64ACC,
    "fp32_f64acc": RMW_PREC_F32_F64ACC,
    "fp32_kahan": RMW_PREC_F32_KAHAN,
    "fp32_blocked": RMW_PREC_F32_BLOCKED,
    "fp64": RMW_PREC_F64,
}


class RmwError(RuntimeError):
    """Error reported by librmw_hip (status code + message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"librmw_hip error {code}: {msg}")
        self.code = code


class Tuning(ctypes.Structure):
    _fields_ = [("chunk_len", ctypes.c_int64), ("max_wgs", ctypes.c_int32), ("rhs_group", ctypes.c_int32),
                ("symmetric", ctypes.c_int32), ("flags", ctypes.c_int32)]


class Segments(ctypes.Structure):
    _fields_ = [("offsets", ctypes.c_void_p), ("nsys", ctypes.c_int32), ("comp", ctypes.c_int32),
                ("nparts", ctypes.c_int32), ("reserved", ctypes.c_int32), ("nblobs", ctypes.c_int64),
                ("part_stride", ctypes.c_int64)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("symmetric", ctypes.c_int32), ("rhs_group", ctypes.c_int32), ("chunk_len", ctypes.c_int64),
                ("work_items", ctypes.c_int64), ("workspace_bytes", ctypes.c_int64)]


_vp = ctypes.c_void_p
_dp = ctypes.c_void_p  # device pointers are passed as integers
_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_sz = ctypes.c_size_t
_szp = ctypes.POINTER(ctypes.c_size_t)
_tp = ctypes.POINTER(Tuning)
_hdp = ctypes.POINTER(ctypes.c_double)
_sgp = ctypes.POINTER(Segments)

# name -> (restype, argtypes); every symbol declared in include/rmw_hip.h.
PROTOTYPES = {
    "rmw_last_error": (ctypes.c_char_p, []),
    "rmw_version": (ctypes.c_char_p, []),
    "rmw_wall_rpy_workspace_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _tp, _szp]),
    "rmw_radii_workspace_size": (ctypes.c_int, [ctypes.c_int, _i64, _i64, _tp, _szp]),
    "rmw_radii_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _i64, _i64,
                                       ctypes.c_double, _vp, _tp, _vp, _sz, _vp]),
    "rmw_radii_self_workspace_size": (ctypes.c_int, [ctypes.c_int, _i64, _vp, _tp, _szp]),
    "rmw_radii_self_describe": (ctypes.c_int, [ctypes.c_int, _i64, _vp, _tp, ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_int64), _szp]),
    "rmw_radii_self_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _i64, ctypes.c_double,
                                            _vp, _tp, _vp, _sz, _vp]),
    "rmw_blob_forces_workspace_size": (ctypes.c_int, [ctypes.c_int, _i64, _tp, _szp]),
    "rmw_blob_forces_describe": (ctypes.c_int, [ctypes.c_int, _i64, _tp, ctypes.POINTER(PlanInfo)]),
    "rmw_blob_forces": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _i64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, _vp, _tp, _vp, _sz, _vp]),
    "rmw_rollers_one_blob_forces": (ctypes.c_int, [_dp, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _vp]),
    "rmw_rollers_uncorrelated_velocity": (ctypes.c_int, [_dp, _dp, _i64, _hdp, ctypes.c_int, ctypes.c_double,
                                                         ctypes.c_double, _dp, _dp, _vp]),
    "rmw_rollers_uncorrelated_noise": (ctypes.c_int, [_dp, _dp, _i64, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double, _dp, _vp]),
    "rmw_rollers_advance_workspace_size": (ctypes.c_int, [_i64, _szp]),
    "rmw_rollers_advance": (ctypes.c_int, [_dp, _dp, _dp, _dp, _i64, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, _dp, _dp, _vp, _sz, _vp]),
    "rmw_profile_collect_kinds": (ctypes.c_int, [_i64p, _hdp]),
    "rmw_profile_pending": (ctypes.c_int, [_i64p]),
    "rmw_profile_collect_launches": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32), _hdp, _i64, _i64p]),
    "rmw_mobility_workspace_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int,
                                                   _tp, _szp]),
    "rmw_mobility_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _i64,
                                          ctypes.c_int, _i64, ctypes.c_double, ctypes.c_double, _hdp, _tp, _vp, _sz,
                                          _vp]),
    "rmw_rigid_force_torque": (ctypes.c_int, [_dp, _i64p, ctypes.c_int, _dp, _dp, _vp, _sz, _vp]),
    "rmw_bodies_advance": (ctypes.c_int, [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, _dp, _dp, _vp]),
    "rmw_count_below_workspace_size": (ctypes.c_int, [_i64, _szp]),
    "rmw_count_below": (ctypes.c_int, [_dp, _i64, ctypes.c_double, _dp, _vp, _sz, _vp]),
    "rmw_wall_rpy_describe": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, ctypes.c_int, _tp, ctypes.c_int,
                                             ctypes.POINTER(PlanInfo)]),
    "rmw_wall_rpy_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _i64, ctypes.c_int,
                                          _i64, ctypes.c_double, ctypes.c_double, _hdp, _tp, _vp, _sz, _vp]),
    "rmw_wall_rpy_apply_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _i64,
                                                  ctypes.c_int, _i64, ctypes.c_double, ctypes.c_double, _hdp, _tp,
                                                  _vp]),
    "rmw_wall_rpy_ragged_workspace_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int,
                                                          ctypes.c_int, _tp, _szp]),
    "rmw_wall_rpy_ragged_apply": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _i64p,
                                                 ctypes.c_int, ctypes.c_int, _i64, ctypes.c_double, ctypes.c_double,
                                                 _hdp, _tp, _vp, _sz, _vp]),
    "rmw_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "rmw_profile_collect": (ctypes.c_int, [_i64p, _hdp, _hdp]),
    "rmw_device_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]),
    # rigid bodies
    "rmw_rigid_workspace_size": (ctypes.c_int, [_i64p, ctypes.c_int, _szp, _szp, _szp]),
    "rmw_rigid_setup": (ctypes.c_int, [_i64p, ctypes.c_int, _vp, _sz, _vp]),
    "rmw_body_blobs": (ctypes.c_int, [_dp, _dp, _dp, _i64p, ctypes.c_int, _dp, _dp, _vp, _sz, _vp]),
    "rmw_rigid_operator_post": (ctypes.c_int, [_dp, _i64p, ctypes.c_int, _dp, _dp, _dp, _vp, _sz, _vp]),
    "rmw_dense_wall_mobility": (ctypes.c_int, [_dp, _i64p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp,
                                               _vp, _sz, _vp]),
    "rmw_rigid_pc_build": (ctypes.c_int, [_dp, _dp, _i64p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp,
                                          _dp, _dp, _vp, _dp, _sz, _vp, _sz, _vp]),
    "rmw_rigid_pc_build_ex": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _i64p, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _vp, _dp, _sz, _vp, _sz,
                                             _vp]),
    "rmw_dense_mobility": (ctypes.c_int, [ctypes.c_int, _dp, _i64p, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, _dp, _vp, _sz, _vp]),
    "rmw_rigid_block_apply": (ctypes.c_int, [_i64p, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                             ctypes.c_int, _i64, _vp, _sz, _vp]),
    "rmw_rigid_point_field": (ctypes.c_int, [_dp, _dp, _dp, _i64, _i64p, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _vp, _sz, _vp]),
    "rmw_rigid_pc_apply": (ctypes.c_int, [_i64p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _vp, _sz, _vp]),
    # Krylov
    "rmw_krylov_workspace_size": (ctypes.c_int, [_i64, ctypes.c_int, _szp]),
    "rmw_dot_many": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _i64, _dp, ctypes.c_int, _vp, _sz, _vp]),
    "rmw_sub_combination": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _dp, _i64, _vp]),
    "rmw_combination": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _dp, _i64, ctypes.c_int, _vp]),
    "rmw_axpby": (ctypes.c_int, [ctypes.c_double, _dp, ctypes.c_double, _dp, _i64, _dp, _vp]),
    "rmw_arnoldi_cgs2": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _i64, _dp, _dp, _vp, _sz, _vp]),
    # segmented (ragged-batch) Krylov kernels
    "rmw_seg_workspace_size": (ctypes.c_int, [_sgp, ctypes.c_int, _szp]),
    "rmw_seg_dot_many": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _sgp, _dp, _i64, ctypes.c_int, _vp, _sz, _vp]),
    "rmw_seg_combination": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _i64, _dp, _sgp, ctypes.c_int, _vp]),
    "rmw_seg_scale": (ctypes.c_int, [_dp, _dp, _dp, _i64, _sgp, _vp]),
    "rmw_seg_arnoldi_cgs2": (ctypes.c_int, [_dp, _i64, ctypes.c_int, _dp, _sgp, _dp, _i64, _dp, _vp, _sz, _vp]),
    # learned warm start
    "rmw_gemm_bf16": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _i64,
                                     _dp, _i64, _dp, _i64, ctypes.c_float, _dp, _i64, _dp, ctypes.c_int, _vp]),
    "rmw_gemm_bf16_splitk_workspace": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _szp]),
    "rmw_gemm_bf16_splitk": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                            _i64, _dp, _i64, _dp, _i64, ctypes.c_float, ctypes.c_int, _dp, _sz, _vp]),
    "rmw_f32_to_bf16": (ctypes.c_int, [_dp, _dp, _i64, _vp]),
    "rmw_silu_backward": (ctypes.c_int, [_dp, _dp, _dp, _i64, _vp]),
    "rmw_colsum_bf16": (ctypes.c_int, [_dp, _i64, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_float, _vp]),
    "rmw_gnn_gather": (ctypes.c_int, [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                      ctypes.c_int, _vp]),
    "rmw_gnn_scatter": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       _dp, _vp]),
    "rmw_mse": (ctypes.c_int, [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _vp]),
    "rmw_adam": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, ctypes.c_int, ctypes.c_float, _vp]),
    "rmw_gnn_knn": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_int, _dp, _vp]),
    "rmw_gnn_features": (ctypes.c_int, [_dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                        _vp]),
    "rmw_gnn_motion_features": (ctypes.c_int, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _vp]),
    "rmw_gnn_body_frame": (ctypes.c_int, [_dp, _dp, ctypes.c_int, _i64, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                          ctypes.c_double, ctypes.c_int, _vp]),
    # Lanczos
    "rmw_lanczos_scratch_doubles": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _szp]),
    "rmw_lanczos_sqrt_coeffs": (ctypes.c_int, [_dp, _i64, _i64, ctypes.c_int, ctypes.c_int, _dp, _dp, _i64, _dp,
                                               _sz, _vp, _vp]),
}

_lock = threading.Lock()
_lib = None


def library_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load librmw_hip.so; raise if it is not built or incomplete."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built. Build it with `python -m rigidmultiblobswall_amd.build` "
                "(there is no CPU fallback).")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)  # AttributeError if the symbol is missing
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int) -> None:
    if rc != RMW_OK:
        msg = load().rmw_last_error()
        raise RmwError(rc, msg.decode() if msg else "")


def precision_code(precision) -> int:
    if isinstance(precision, int):
        return precision
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"unknown precision {precision!r}; use one of {sorted(PRECISIONS)}") from None


def tuning_ptr(tuning):
    if tuning is None:
        return None
    if isinstance(tuning, Tuning):
        return ctypes.byref(tuning)
    t = Tuning(int(tuning.get("chunk_len", 0)), int(tuning.get("max_wgs", 0)), int(tuning.get("rhs_group", 0)),
               int(tuning.get("symmetric", 0)), int(tuning.get("flags", 0)))
    return ctypes.byref(t)


def with_flags(tuning, flags: int):
    """A Tuning equal to `tuning` (None, a dict or a Tuning; None = the
    built-in launch shape) with `flags` set in addition."""
    if tuning is None:
        return Tuning(0, 0, 0, 0, int(flags))
    if isinstance(tuning, Tuning):
        return Tuning(tuning.chunk_len, tuning.max_wgs, tuning.rhs_group, tuning.symmetric, tuning.flags | int(flags))
    return Tuning(int(tuning.get("chunk_len", 0)), int(tuning.get("max_wgs", 0)), int(tuning.get("rhs_group", 0)),
                  int(tuning.get("symmetric", 0)), int(tuning.get("flags", 0)) | int(flags))
