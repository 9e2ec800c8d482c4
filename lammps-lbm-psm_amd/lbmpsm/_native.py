"""ctypes binding of liblbmpsm.so (the C ABI declared in include/lbmpsm.h).

The library is the only compute path: if it is missing, importing this
module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbmpsm.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_STATE = -3
ERR_NOMEM = -4
ERR_NODEVICE = -5
ERR_LIMIT = -6

STORE_F64, STORE_F32, STORE_F16 = 0, 1, 2
STORAGE = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16}

BC_NONE, BC_SHEAR, BC_XFLOW, BC_CHANNEL, BC_CHANNEL_VELIN_PRESSOUT, BC_CLOSEDBOX = range(6)
# `fix lbm-psm-bc` keywords, src/fix_LBM_PSM_BC.cpp:24-39
BC_KEYWORDS = {
    "shear": BC_SHEAR,
    "xFlow": BC_XFLOW,
    "channel": BC_CHANNEL,
    "channel-velIn-pressOut": BC_CHANNEL_VELIN_PRESSOUT,
    "closedBox": BC_CLOSEDBOX,
}

STEP_STORE_MACROS = 1


class Config(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("periodic", C.c_int * 3),
        ("tau", C.c_double),
        ("F_lbm", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("dx", C.c_double),
        ("Uc", C.c_double),
        ("u_lb", C.c_double),
        ("tc", C.c_double),
        ("dt_d", C.c_double),
        ("force_factor", C.c_double),
        ("torque_factor", C.c_double),
        ("storage", C.c_int),
        ("device", C.c_int),
    ]


class Info(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("q", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("sx", C.c_int),
        ("sy", C.c_int),
        ("sz", C.c_int),
        ("ref_offset", C.c_int * 3),
        ("stored_nodes", C.c_int64),
        ("storage", C.c_int),
        ("device_bytes", C.c_int64),
    ]


class LbmPsmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lbmpsm error {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run make -C lammps-lbm-psm_amd"
        )
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    sigs = {
        "lbmpsm_config_from_fix": (
            C.c_int,
            [C.POINTER(Config), C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
             C.c_double, C.c_double, dp, dp, dp, ip],
        ),
        "lbmpsm_create": (C.c_int, [C.POINTER(Config), C.POINTER(P)]),
        "lbmpsm_destroy": (C.c_int, [P]),
        "lbmpsm_last_error": (C.c_char_p, [P]),
        "lbmpsm_get_info": (C.c_int, [P, C.POINTER(Info)]),
        "lbmpsm_init_equilibrium": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.c_double]),
        "lbmpsm_set_bc": (C.c_int, [P, C.c_int, C.c_double, C.c_double]),
        "lbmpsm_set_bc_ranks": (C.c_int, [P, C.c_int, C.c_double, C.c_double, ip]),
        "lbmpsm_set_particles": (C.c_int, [P, C.c_int, ip, dp, dp, dp, dp]),
        "lbmpsm_set_particles_device": (C.c_int, [P, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_double]),
        "lbmpsm_step": (C.c_int, [P, C.c_int]),
        "lbmpsm_step_noalloc": (C.c_int, [P, C.c_int, P]),
        "lbmpsm_get_hydro": (C.c_int, [P, C.c_int, dp, dp, dp]),
        "lbmpsm_get_populations": (C.c_int, [P, dp]),
        "lbmpsm_set_populations": (C.c_int, [P, dp]),
        "lbmpsm_get_populations_ref": (C.c_int, [P, dp]),
        "lbmpsm_set_populations_ref": (C.c_int, [P, dp]),
        "lbmpsm_get_macros": (C.c_int, [P, dp, dp]),
        "lbmpsm_get_slots": (C.c_int, [P, ip, dp, dp, dp]),
        "lbmpsm_mass": (C.c_int, [P, dp]),
        "lbmpsm_sync": (C.c_int, [P]),
        "lbmpsm_profile_enable": (C.c_int, [P, C.c_int]),
        "lbmpsm_profile_read": (C.c_int, [P, C.c_char_p, dp, C.POINTER(C.c_int64)]),
        "lbmpsm_profile_reset": (C.c_int, [P]),
        "lbmpsm_adjoint_forward": (C.c_int, [P, C.c_int, C.c_int, C.c_int, ip, dp, dp, dp, dp,
                                             dp, dp, dp]),
        "lbmpsm_adjoint_backward": (C.c_int, [P, C.c_int, dp, dp, dp, dp, dp, dp, dp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        msg = LIB.lbmpsm_last_error(ctx)
        raise LbmPsmError(rc, msg.decode() if msg else "")
