"""ctypes binding of libhor.so (the C ABI declared in include/hor.h)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libhor.so")

_D, _I, _P, _S = C.c_double, C.c_int, C.c_void_p, C.c_size_t
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

HOR_OK = 0
SL_ON, SL_OFF, DL_ON, DL_OFF, P_OFF, SLDL_OFF = 0, 1, 2, 3, 4, 5
ALL_NEAR, ON_AT = 0x10, 0x20   # kind modifier bits (hor.h)
DIRECT, TREE = 0, 1
FP64, MIXED = 0, 1
FAR_VALU, FAR_MFMA = 0, 1   # far-field kernel of a mixed single-layer treecode plan


class TreeParams(C.Structure):
    _fields_ = [("P", _I), ("N0", _I), ("theta", _D)]


class Body(C.Structure):
    _fields_ = [("abc", _D * 3), ("ctr", _D * 3), ("R", _D * 9), ("rotated", _I)]


class SurfaceDesc(C.Structure):
    _fields_ = [("n", _I), ("x", _P), ("nrml", _P), ("area", _P), ("ctr", _D * 3),
                ("nrst", _P), ("x0", _P), ("n0", _P), ("b", _P), ("i123", _P)]


class PrecondPolicy(C.Structure):
    _fields_ = [("force_plain", _I), ("gate", _I), ("gate_ratio", _D)]


class FgmresInfo(C.Structure):
    _fields_ = [("iterations", _I), ("learned_steps", _I), ("plain_steps", _I), ("nonfinite_outputs", _I),
                ("cycles_gated_on", _I), ("cycles_gated_off", _I), ("monitor_stops", _I), ("gate_misfit_identity", _D),
                ("gate_misfit_network", _D), ("relres", _D)]


# name -> (restype, argtypes)
_SIGS = {
    "hor_last_error": (C.c_char_p, []),
    "hor_version": (C.c_char_p, []),
    "hor_device_count": (_I, [_IP]),
    "hor_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "hor_ctx_destroy": (_I, [_P]),
    "hor_ctx_last_timing": (_I, [_P, _DP, _DP]),
    "hor_plan_create": (_I, [_P, _I, _I, _I, _D, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P,
                             C.POINTER(TreeParams), C.POINTER(_P)]),
    "hor_plan_destroy": (_I, [_P]),
    "hor_plan_eval": (_I, [_P, _P, _P, _P]),
    "hor_plan_eval_dev": (_I, [_P, _P, _P, _P]),
    "hor_plan_stats": (_I, [_P, _DP, _DP, _DP, _DP]),
    "hor_plan_set_far_kernel": (_I, [_P, _I]),
    "hor_plan_phase_timing": (_I, [_P, _DP, _DP, _DP]),
    "hor_sl_offsurf": (_I, [_P, _D, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    "hor_sl_onsurf": (_I, [_P, _D, _I, _P, _P, _P, _P, _I, _P, _P]),
    "hor_dl_offsurf": (_I, [_P, _D, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P]),
    "hor_dl_onsurf": (_I, [_P, _D, _I, _P, _P, _P, _P, _I, _P, _P]),
    "hor_dl_onsurf_dev_noalloc": (_I, [_P, _D, _I, _P, _P, _P, _P, _I, _P, _P]),
    "hor_sl_onsurf_tree": (_I, [_P, _P, C.POINTER(TreeParams), _D, _I, _P, _P, _P, _P, _I, _P, _P]),
    "hor_sl_offsurf_tree": (_I, [_P, _P, C.POINTER(TreeParams), _D, _I, _P, _P, _P, _P, _I, _P,
                                 _P, _P, _P, _P]),
    "hor_dl_onsurf_tree": (_I, [_P, _P, C.POINTER(TreeParams), _D, _I, _P, _P, _P, _P, _I, _P, _P]),
    "hor_dl_offsurf_tree": (_I, [_P, _P, C.POINTER(TreeParams), _D, _I, _P, _P, _P, _P, _I, _P,
                                 _P, _P, _P, _P, _P]),
    "hor_spheroid_exact_pressure": (_I, [C.POINTER(Body), _I, _P, _P]),
    "hor_level_set": (_I, [C.POINTER(Body), _I, _P, _P]),
    "hor_dst_poisson3d": (_I, [_I, _P]),
    "hor_grid_solve_poisson": (_I, [_I, _D, _P, _P, _P, _P, _I]),
    "hor_pressure_7ord": (_I, [_P, _D, _I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P]),
    "hor_pressure_away": (_I, [_P, _I, _P, _P, _P, _P, _I, _P, _P]),
    "hor_solver_create": (_I, [_P, _D, _D, _I, _I, C.POINTER(TreeParams), C.POINTER(TreeParams),
                               C.POINTER(SurfaceDesc), C.POINTER(SurfaceDesc), C.POINTER(_P)]),
    "hor_solver_destroy": (_I, [_P]),
    "hor_solver_size": (_I, [_P, _IP]),
    "hor_solver_stats": (_I, [_P, _DP, _DP, _DP, _DP]),
    "hor_solver_phase_timing": (_I, [_P, _DP, _DP, _DP]),
    "hor_solver_plan_stats": (_I, [_P, _I, _DP, _DP, _DP, _DP]),
    "hor_solver_matvec": (_I, [_P, _P, _P]),
    "hor_solver_matvec_dev": (_I, [_P, _P, _P]),
    "hor_solver_eval_sl": (_I, [_P, _P, _P, _P, _P, _P]),
    "hor_solver_interp_g0": (_I, [_P, _I, _P, _P]),
    "hor_solver_gmres": (_I, [_P, _P, _P, _I, _D, _I, _IP, _DP, _P]),
    "hor_solver_fgmres": (_I, [_P, _P, _P, _I, _D, _I, _P, _IP, _DP, _P, _P, _P, _I, _IP]),
    "hor_solver_set_precond_policy": (_I, [_P, C.POINTER(PrecondPolicy)]),
    "hor_solver_fgmres_info": (_I, [_P, C.POINTER(FgmresInfo)]),
    "hor_mpnn_param_count": (_I, []),
    "hor_mpnn_create": (_I, [_P, _I, C.POINTER(_P)]),
    "hor_mpnn_destroy": (_I, [_P]),
    "hor_mpnn_set_params": (_I, [_P, _P]),
    "hor_mpnn_get_params": (_I, [_P, _P]),
    "hor_mpnn_get_grad": (_I, [_P, _P]),
    "hor_mpnn_set_geometry": (_I, [_P, _I, _I, _P, _P, _P, _P, _I, _P]),
    "hor_mpnn_check_size": (_I, [_P, _I]),
    "hor_mpnn_forward_dev": (_I, [_P, _I, _P, _P]),
    "hor_mpnn_apply_dev": (_I, [_P, _P, _P, _P]),
    "hor_mpnn_loss_grad_dev": (_I, [_P, _I, _P, _P, _DP]),
    "hor_mpnn_adam_step": (_I, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "hor_knn_graph": (_I, [_I, _P, _I, _P]),
    "hor_dev_alloc": (_I, [_P, _S, C.POINTER(_P)]),
    "hor_dev_free": (_I, [_P, _P]),
    "hor_dev_upload": (_I, [_P, _P, _P, _S]),
    "hor_dev_download": (_I, [_P, _P, _P, _S]),
    "hor_body_init": (_I, [C.POINTER(Body), _P, _P, _P]),
    "hor_surface_generate": (_I, [C.POINTER(Body), _D, _D, _I, _IP, _P, _P, _P, _P]),
    "hor_nearest_data": (_I, [C.POINTER(Body), _D, _D, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "hor_surface_force": (_I, [C.POINTER(Body), _I, _P, _P, _P]),
    "hor_band_targets": (_I, [C.POINTER(Body), _P, _D, _I, _D, _I, _P, _I, _IP, _P, _P, _P, _P, _P]),
    "hor_surface_generate_gpu": (_I, [_P, C.POINTER(Body), _D, _D, _I, _IP, _P, _P, _P, _P]),
    "hor_nearest_data_gpu": (_I, [_P, C.POINTER(Body), _D, _D, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "hor_stencil_build": (_I, [_P, _P, _D, _I, _P, _I, _P, _P, _P, _P, _P]),
    "hor_traction_translating": (_I, [C.POINTER(Body), _P, _I, _P, _P]),
    "hor_spheroid_exact_velocity": (_I, [C.POINTER(Body), _I, _P, _P]),
    "hor_translating_exact_velocity": (_I, [C.POINTER(Body), _P, _I, _P, _P]),
}

EXPORTED = sorted(_SIGS)


class HorError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libhor.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, where):
    if rc != HOR_OK:
        msg = lib().hor_last_error()
        raise HorError(rc, where, msg.decode() if msg else "")


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(_P)


def f64(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def i32(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.int32).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a
