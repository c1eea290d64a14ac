"""ctypes binding of libmaddening_lbm.so (include/lbm_c_api.h).

The library is the only compute path: if it is missing this module raises,
there is no Python fallback.
"""
from __future__ import annotations

import ctypes
import os

# torch first: its HIP runtime (soname libamdhip64.so.7) must be the one the
# library binds to, so both share devices, streams and allocations.
import torch  # noqa: F401

from .build import library_path

LBM_F32, LBM_F64, LBM_F16_DEV, LBM_BF16_DEV, LBM_F32_DEV = 0, 1, 2, 3, 4
LBM_BB_NONE, LBM_BB_REFERENCE, LBM_BB_HALFWAY = 0, 1, 2
LBM_SCHEME_AUTO, LBM_SCHEME_PULL, LBM_SCHEME_PUSH = 0, 1, 2
LBM_FLAG_XCD_REMAP = 1
LBM_FLAG_ADJ_IN_SWAPPED = 2
LBM_FLAG_ADJ_OUT_SWAPPED = 4
LBM_FLAG_ADJ_OUT_FACES = 8
LBM_FLAG_PAIR_OFF = 16
LBM_FLAG_PAIR_FORCE = 32
LBM_FLAG_TRI_OFF = 64

STORAGE_NAMES = {"f32": LBM_F32, "f64": LBM_F64, "f16dev": LBM_F16_DEV,
                 "bf16dev": LBM_BF16_DEV, "f32dev": LBM_F32_DEV}
BB_NAMES = {"none": LBM_BB_NONE, "reference": LBM_BB_REFERENCE,
            "halfway": LBM_BB_HALFWAY}
SCHEME_NAMES = {"auto": LBM_SCHEME_AUTO, "pull": LBM_SCHEME_PULL,
                "push": LBM_SCHEME_PUSH}


class LbmDesc(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("bb_mode", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 3),
        ("tau", ctypes.c_double),
        ("force", ctypes.c_double * 3),
        ("lid_velocity", ctypes.c_double * 3),
        ("block_x", ctypes.c_int32),
        ("row_vec", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("xcd_group", ctypes.c_int32),
        ("inlet_face", ctypes.c_int32),
        ("outlet_face", ctypes.c_int32),
        ("inlet_rho", ctypes.c_double),
        ("outlet_rho", ctypes.c_double),
    ]


class LbmPipeDesc(ctypes.Structure):
    """Mirror of lbm_pipe_desc."""
    _fields_ = [
        ("dims", ctypes.c_int64 * 3),
        ("storage", ctypes.c_int32),
        ("multiphase", ctypes.c_int32),
        ("tau", ctypes.c_double),
        ("tau_tracer", ctypes.c_double),
        ("propeller_strength", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("G", ctypes.c_double),
        ("rho_0", ctypes.c_double),
        ("rho_wall", ctypes.c_double),
        ("rho_liquid", ctypes.c_double),
        ("rho_gas", ctypes.c_double),
        ("flags", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class LbmBatchItem(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_int32),
        ("bb_mode", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 3),
        ("tau", ctypes.c_double),
        ("force", ctypes.c_double * 3),
        ("lid_velocity", ctypes.c_double * 3),
        ("f", ctypes.c_void_p * 2),
        ("mask", ctypes.c_void_p),
        ("rho", ctypes.c_void_p),
        ("u", ctypes.c_void_p),
        ("p", ctypes.c_void_p),
        ("pitch", ctypes.c_int64),
        ("links", ctypes.c_void_p),
    ]


_vp = ctypes.c_void_p
_PROTOS = {
    "lbm_version": (ctypes.c_int, []),
    "lbm_last_error": (ctypes.c_char_p, []),
    "lbm_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(LbmDesc), ctypes.c_int]),
    "lbm_state_bytes": (ctypes.c_size_t, [ctypes.POINTER(LbmDesc)]),
    "lbm_step": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, ctypes.c_int]),
    "lbm_moments": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp, _vp, _vp, _vp,
                                   _vp]),
    "lbm_equilibrium": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp, _vp]),
    "lbm_convert": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), ctypes.c_int, ctypes.c_int,
                                   _vp, _vp]),
    "lbm_reduce_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(LbmDesc)]),
    "lbm_batch_plan_size": (ctypes.c_size_t, [ctypes.c_int]),
    "lbm_batch_plan_create": (ctypes.c_int, [_vp, ctypes.POINTER(LbmBatchItem), ctypes.c_int,
                                             ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_void_p),
                                             ctypes.POINTER(ctypes.c_longlong)]),
    "lbm_batch_plan_destroy": (ctypes.c_int, [_vp]),
    "lbm_batch_plan_pairs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lbm_batch_run": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int]),
    "lbm_batch_graph_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "lbm_batch_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "lbm_batch_graph_destroy": (ctypes.c_int, [_vp]),
    "lbm_adjoint_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(LbmDesc)]),
    "lbm_adjoint_begin": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp]),
    "lbm_adjoint_step": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, ctypes.c_int, _vp,
                                        _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lbm_pull_phase": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), ctypes.c_int, _vp, _vp, _vp,
                                      _vp, _vp, _vp, _vp, _vp, _vp]),
    "lbm_step_schedule": (ctypes.c_int, [ctypes.POINTER(LbmDesc), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]),
    "lbm_links_size": (ctypes.c_size_t, [ctypes.POINTER(LbmDesc)]),
    "lbm_links_build": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp]),
    "lbm_moments_adjoint": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp]),
    "lbm_adjoint_end": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp]),
    "lbm_conservation": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp, _vp]),
    "lbm_conservation_static": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp]),
    "lbm_adjoint_face_end": (ctypes.c_int, [_vp, ctypes.POINTER(LbmDesc), _vp, _vp]),
    "lbm_pipe_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(LbmPipeDesc)]),
    "lbm_pipe_step": (ctypes.c_int, [_vp, ctypes.POINTER(LbmPipeDesc), _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _vp, ctypes.c_int]),
}

_LIB = None


class LBMError(RuntimeError):
    """A C-ABI call returned a negative lbm_status."""


def lib() -> ctypes.CDLL:
    """Load the library once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it with `python -m maddening_amd.build` "
                "(there is no non-HIP fallback)")
        h = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().lbm_last_error().decode(errors="replace")
        raise LBMError(f"{what} failed ({rc}): {msg}")


def exported_symbols() -> list[str]:
    return list(_PROTOS)
