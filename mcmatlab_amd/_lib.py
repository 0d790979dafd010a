"""Loader for libmcm.so (the in-tree gfx950 build).  There is no fallback:
if the library is missing or no gfx950 device is visible, calls raise."""
from __future__ import annotations

import ctypes as C
import os

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmcm.so")

# every function include/mcm_abi.h declares, with its ctypes signature
SIGNATURES = {
    "mcm_last_error": (C.c_char_p, []),
    "mcm_version": (C.c_int, []),
    "mcm_device_count": (C.c_int, []),
    "mcm_free": (None, [C.c_void_p]),
    "mcm_mc_output_sizes": (C.c_int, [C.POINTER(abi.mcm_mc_input), C.POINTER(abi.mcm_mc_sizes)]),
    "mcm_mc_run": (C.c_int, [C.POINTER(abi.mcm_mc_input), C.POINTER(abi.mcm_mc_output)]),
    "mcm_mc_plan_create": (C.c_int, [C.POINTER(abi.mcm_mc_input), C.POINTER(C.c_void_p)]),
    "mcm_mc_plan_simulate": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]),
    "mcm_mc_plan_normalize": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(abi.mcm_mc_output)]),
    "mcm_mc_plan_tallies": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]),
    "mcm_mc_plan_counters": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "mcm_mc_plan_destroy": (C.c_int, [C.c_void_p]),
    "mcm_mc_plan_autotune": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.POINTER(C.c_double)]),
    "mcm_mc_plan_set_fluence_copies": (C.c_int, [C.c_void_p, C.c_int32]),
    "mcm_mc_batch_run": (C.c_int, [C.c_void_p, C.POINTER(abi.mcm_mc_item), C.c_int64, C.POINTER(abi.mcm_mc_output)]),
    "mcm_mc_multi_run": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64, C.POINTER(abi.mcm_mc_multi_item), C.c_int64,
                                   C.POINTER(abi.mcm_mc_output)]),
    "mcm_mc_grad_run": (C.c_int, [C.c_void_p, C.POINTER(abi.mcm_mc_grad_input), C.POINTER(abi.mcm_mc_grad_output)]),
    "mcm_heat_run": (C.c_int, [C.POINTER(abi.mcm_heat_input), C.POINTER(abi.mcm_heat_output)]),
    "mcm_heat_plan_create": (C.c_int, [C.POINTER(abi.mcm_heat_input), C.POINTER(C.c_void_p)]),
    "mcm_heat_plan_advance": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_int32]),
    "mcm_heat_plan_set_graph_steps": (C.c_int, [C.c_void_p, C.c_int64]),
    "mcm_heat_plan_set_explicit_zrun": (C.c_int, [C.c_void_p, C.c_int32]),
    "mcm_heat_plan_explicit_zrun": (C.c_int32, [C.c_void_p]),
    "mcm_heat_group_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64, C.POINTER(C.c_void_p)]),
    "mcm_heat_group_advance": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int64]),
    "mcm_heat_group_destroy": (C.c_int, [C.c_void_p]),
    "mcm_heat_plan_download": (C.c_int, [C.c_void_p, C.POINTER(abi.mcm_heat_output)]),
    "mcm_heat_plan_destroy": (C.c_int, [C.c_void_p]),
    "mcm_heat_adjoint_run": (C.c_int, [C.POINTER(abi.mcm_heat_input), C.POINTER(abi.mcm_heat_segment), C.c_int64,
                                       C.POINTER(C.c_uint8), C.c_int64, C.POINTER(abi.mcm_heat_adjoint_output)]),
    "mcm_downsample_params": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32]),
    "mcm_downsample_params_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
}

_lib = None


class MCMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmcm error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MCMError(abi.MCM_ERR_NO_DEVICE,
                           f"{LIB_PATH} is not built; run `python -m mcmatlab_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != abi.MCM_OK:
        raise MCMError(rc, lib().mcm_last_error().decode(errors="replace"))
    return rc
