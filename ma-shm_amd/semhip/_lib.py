"""ctypes binding of libsemhip.so (the C-ABI declared in include/sem_hip.h)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libsemhip.so"

SEM_OK, SEM_E_ARG, SEM_E_HIP, SEM_E_STATE, SEM_E_DIVERGED, SEM_E_NOMEM, SEM_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
SEM_SHELL5, SEM_SOLID3 = 1, 2
SEM_FP64, SEM_FP32_SPLIT = 0, 1
SEM_Q_DISP, SEM_Q_VEL, SEM_Q_DISP_POST, SEM_Q_ACC = 0, 1, 2, 3

_p = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_pi32 = C.POINTER(C.c_int32)
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)
_pint = C.POINTER(C.c_int)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "sem_create": (_i, [C.POINTER(_p), _i]),
    "sem_destroy": (_i, [_p]),
    "sem_last_error": (C.c_char_p, [_p]),
    "sem_version": (C.c_char_p, []),
    "sem_mesh_change_turn_quad": (_i, [_i64, _pi32, _i64, _pd, _pi64]),
    "sem_mesh_quad2spectral": (_i, [_i64, _pi32, _i64, _pd, _i, _pi32, _pd, _pi64]),
    "sem_mesh_extrude": (_i, [_i64, _i, _pi32, _i64, _pd, _i, _i, _pd, _pi32, _pd]),
    "sem_mesh_baskets": (_i, [_i64, _i, _pi32, _i64, _pi32, _pi32]),
    "sem_laminate_shell": (_i, [_i, _pd, _pd, _pd, _pd, _pd]),
    "sem_laminate_shell_delam": (_i, [_i, _pd, _pd, _pd, _pd, _i, _i, _pd]),
    "sem_grid_create": (_i, [C.POINTER(C.c_void_p), _i, _i64, _i, _pi32, _i64, _pd, _i64, _i, _i]),
    "sem_grid_destroy": (_i, [C.c_void_p]),
    "sem_grid_plan": (_i, [_i64, _i, _pi32, _i64, _pd, _i64, _i, _i, _pd, _pd, _pi32, _pd, _pd]),
    "sem_grid_last_error": (C.c_char_p, [C.c_void_p]),
    "sem_grid_points": (_i, [C.c_void_p, _pd, _pd, _pi32, _pd, _pd]),
    "sem_grid_eval": (_i, [C.c_void_p, _i64, _i64, _pd, _i, _i, C.c_double, _i64, _pi32, _pi32, _pd]),
    "sem_mesh_boundary_nodes": (_i, [_i64, _i, _pi32, _pi32, _pi64]),
    "sem_mesh_split_delam": (_i, [_i64, _i, _pi32, _i64, _pd, _i64, _pi32, _i64, _pi32, _pi32, _pd, _pi64,
                                  _pi32, _pi32]),
    "sem_laminate_solid_smeared": (_i, [_i, _pd, _pd, _pd, _pd, _pd]),
    "sem_laminate_solid_layered": (_i, [_i, _pd, _pd, _pd, _pd]),
    "sem_upload_mesh": (_i, [_p, _i, _i, _i, _i64, _pi32, _i64, _pd]),
    "sem_upload_batch": (_i, [_p, _i, _i, _i, _i, _pi64, _pi32, _pi64, _pd]),
    "sem_basket_rows": (_i, [_p, _pi64]),
    "sem_get_baskets": (_i, [_p, _pi32, _pi32]),
    "sem_mesh_is_affine": (_i, [_p, _pint]),
    "sem_set_kernel": (_i, [_p, _i, _i]),
    "sem_get_kernel": (_i, [_p, _pint, _pint]),
    "sem_set_material": (_i, [_p, _i, _i, _pd, _pi32, _pd, _pd]),
    "sem_set_damping": (_i, [_p, _pd]),
    "sem_get_mass": (_i, [_p, _pd]),
    "sem_set_excitation": (_i, [_p, _i64, _pi64, _pd, _i64, _pd]),
    "sem_set_excitation_multi": (_i, [_p, _i64, _pi64, _pd, _pi32, _i, _i64, _pd]),
    "sem_set_outputs": (_i, [_p, _i, _pi64, _pi64, _pd, _i]),
    "sem_set_frames": (_i, [_p, _i, _pi64, _i]),
    "sem_set_guard": (_i, [_p, _d, _i, _i]),
    "sem_set_state": (_i, [_p, _pd, _pd]),
    "sem_get_state": (_i, [_p, _pd, _pd]),
    "sem_run": (_i, [_p, _d, _i64, _i64, _i, _pd, _i64, _pd, _pd, _pi64, _pd]),
    "sem_profile": (_i, [_p, _d, _i64, _i64, _i, _pd]),
    "sem_apply_K": (_i, [_p, _pd, _pd, _i]),
    "sem_apply_K_noalloc": (_i, [_p, _p, _p, _i]),
    "sem_gll": (_i, [_i, _pd, _pd]),
    "sem_pzt_const": (_i, [_i, _pd, _pd, _pd, _d, _pd, _pd, _pd, _pd]),
    "sem_pzt_actuator": (_i, [_i, _i, _i64, _pi32, _i64, _pd, _pd, _pd, _i, _d, _pd]),
    "sem_pzt_sensor": (_i, [_i, _i, _i64, _pi32, _i64, _pd, _pd, _pd, _i, _pd]),
    "sem_pzt_shell_actuator": (_i, [_i, _i, _i64, _pi32, _i64, _pd, _d, _d, _pd, _pd, _d, _pd]),
    "sem_laminate_shell_add_pzt": (_i, [_pd, _d, _d, _d, _pd, _pd]),
    "sem_shell_pzt_sensor": (_i, [_p, _i64, _pi32, _d, _d, _d, _d, _d, _pd]),
    "sem_adjoint": (_i, [_p, _d, _i64, _i64, _i, _pd, _i64, _pd, _pd, _i64, _pd]),
}

_lib = None


class SemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sem error {code}: {msg}")
        self.code = code


def load():
    """Load the HIP library; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python ma-shm_amd/build.py` first")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)
