"""ctypes binding of ``libpasif_hip.so`` (C ABI in ``include/pasif.h``).

There is no CPU fallback: if the library is missing or cannot be loaded the
import fails loudly with instructions to build it.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PASIF_LIB", os.path.join(HERE, "libpasif_hip.so"))

PASIF_OK = 0
PASIF_SOLVER_PERSISTENT = 0
PASIF_SOLVER_PLAIN = 1
PASIF_SOLVER_GRAPH = 2
ERROR_NAMES = {1: "argument", 2: "device", 3: "unsupported", 4: "memory", 5: "state"}


class PasifError(RuntimeError):
    """Failure reported by the HIP library (status code + message)."""

    def __init__(self, code: int, message: str):
        self.code = code
        super().__init__(f"[pasif {ERROR_NAMES.get(code, code)}] {message}")


_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_h = C.c_void_p

_SIGS = {
    "pasif_create": (C.c_int, [C.POINTER(C.c_void_p), _f64p, C.c_uint32, C.c_uint32, C.c_uint32,
                               C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pasif_destroy": (C.c_int, [_h]),
    "pasif_last_error": (C.c_char_p, [_h]),
    "pasif_set_fwd_K": (C.c_int, [_h, _u32p, _f64p, _u32p, _u32p, C.c_size_t]),
    "pasif_set_bwd_K": (C.c_int, [_h, _u32p, _f64p, _u32p, _u32p, C.c_size_t]),
    "pasif_set_fwd_gamma": (C.c_int, [_h, _u32p, _f64p, _u32p, C.c_size_t]),
    "pasif_set_fwd_lambda": (C.c_int, [_h, _u32p, _f64p, _u32p, C.c_size_t]),
    "pasif_set_fwd_psi": (C.c_int, [_h, _u32p, _f64p, _u32p, C.c_size_t]),
    "pasif_set_bwd_gamma": (C.c_int, [_h, _u32p, _f64p, _u32p, C.c_size_t]),
    "pasif_set_bwd_lambda": (C.c_int, [_h, _u32p, _f64p, _u32p, C.c_size_t]),
    "pasif_set_bwd_psi": (C.c_int, [_h, _u32p, _f64p, _u32p, C.c_size_t]),
    "pasif_set_fwd_force_pattern": (C.c_int, [_h, _f64p, C.c_size_t]),
    "pasif_set_fwd_initial_conditions": (C.c_int, [_h, _f64p, C.c_size_t]),
    "pasif_set_bwd_force_pattern": (C.c_int, [_h, _f64p, C.c_size_t]),
    "pasif_set_bwd_initial_conditions": (C.c_int, [_h, _f64p, C.c_size_t]),
    "pasif_allocate_system": (C.c_int, [_h]),
    "pasif_allocate_adjoint": (C.c_int, [_h]),
    "pasif_load_excitations": (C.c_int, [_h, _f64p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "pasif_set_interpolation_matrix": (C.c_int, [_h, _f64p, C.c_size_t, C.c_uint32]),
    "pasif_set_modulation_buffer": (C.c_int, [_h, _f64p, C.c_size_t]),
    "pasif_get_amplitudes": (C.c_int, [_h, _f64p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pasif_get_trajectory": (C.c_int, [_h, C.c_uint32, _f64p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pasif_get_gradient": (C.c_int, [_h, C.c_uint32, _f64p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pasif_set_systems_layout": (C.c_int, [_h, C.c_uint32, _u32p, _u32p, _i32p, _u32p]),
    "pasif_set_precision": (C.c_int, [_h, C.c_int]),
    "pasif_set_solver": (C.c_int, [_h, C.c_int]),
    "pasif_set_tuning": (C.c_int, [_h, C.c_uint32, C.c_uint32, C.c_uint32]),
    "pasif_set_amplitude_window": (C.c_int, [_h, C.c_uint32]),
    "pasif_set_graph_slicing": (C.c_int, [_h, C.c_uint32, C.c_uint32]),
    "pasif_get_max_amplitudes": (C.c_int, [_h, _f64p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pasif_last_timing": (C.c_int, [_h, _f64p, _u32p, _u32p, _u32p]),
    "pasif_set_kernel": (C.c_int, [_h, C.c_int]),
    "pasif_set_lane_group": (C.c_int, [_h, C.c_uint32]),
    "pasif_set_lane_bias": (C.c_int, [_h, C.c_int]),
    "pasif_lane_groups": (C.c_int, [_h, C.c_int, _u32p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pasif_kernel_info": (C.c_int, [_h, C.c_int, C.POINTER(C.c_int), _u32p, _u32p]),
    "pasif_lane_source": (C.c_int, [_h, C.c_int, C.c_char_p, C.c_size_t,
                                    C.POINTER(C.c_size_t), C.c_int]),
    "pasif_set_adjoint_mode": (C.c_int, [_h, C.c_int]),
    "pasif_run_footprint": (C.c_int, [_h, C.c_int, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "pasif_solve_noalloc": (C.c_int, [_h, C.c_int, C.c_void_p, C.c_void_p]),
    "pasif_version": (C.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


def exported_symbols():
    return list(_SIGS)


def load():
    """Load the library (once) and declare every C-ABI signature."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pasif_amd.build` "
                "(hipcc, gfx950).  There is no CPU fallback.")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(lib, handle, rc: int):
    if rc != PASIF_OK:
        msg = lib.pasif_last_error(handle)
        raise PasifError(rc, msg.decode() if msg else "")
