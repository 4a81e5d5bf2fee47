"""ctypes binding of the gfx950 HIP library (declared in include/aerokinetic.h).

The library is built in-tree (`aerokinetic_amd/lib/libaerokinetic_hip.so`).
There is no fallback: if it is missing or cannot be loaded, every solver entry
point raises `NativeLibraryError`.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libaerokinetic_hip.so"

STORE = {"fp64": 0, "fp32": 1, "fp16": 2}
FARFIELD = {"periodic": 0, "legacy": 1, "characteristic_transport_v1": 2, "exterior_regularized_fd_v1": 4}
HIST_FIXED = 9
HIST_EXT = 41  # exterior columns: wall mass accumulated, source mass, source mass accumulated
MAX_ELEM = 8


class NativeLibraryError(RuntimeError):
    pass


class HipError(RuntimeError):
    pass


_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

_SIGS = {
    "ak_lattice_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, _vp]),
    "ak_lattice_destroy": (ctypes.c_int, [_vp]),
    "ak_population_bytes": (ctypes.c_size_t, [_vp]),
    "ak_bind_populations": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "ak_set_solid": (ctypes.c_int, [_vp, _u8p]),
    "ak_set_walls": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _dp, _ip, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "ak_set_farfield": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "ak_init_seeded": (ctypes.c_int, [_vp, _dp]),
    "ak_set_sweep_tile": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ak_set_sweep_tile_form": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ak_next_sweep_odd": (ctypes.c_int, [_vp]),
    "ak_populations_written": (ctypes.c_int, [_vp]),
    "ak_set_sweep_waves": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ak_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int]),
    "ak_run_profiled": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, _dp]),
    "ak_set_graph_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ak_history_width": (ctypes.c_int, [_vp]),
    "ak_read_history": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "ak_get_populations": (ctypes.c_int, [_vp, _dp]),
    "ak_set_populations": (ctypes.c_int, [_vp, _dp]),
    "ak_get_macroscopic": (ctypes.c_int, [_vp, _dp, _dp]),
    "ak_diagnostics": (ctypes.c_int, [_vp, _dp]),
    "ak_diagnostics_noalloc": (ctypes.c_int, [_vp, _dp]),
    "ak_get_plane": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "ak_synchronize": (ctypes.c_int, [_vp]),
    "ak_last_error": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    "ak_version": (ctypes.c_char_p, []),
    "ak_device_copy": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "ak_state_bytes": (ctypes.c_size_t, [_vp]),
    "ak_residual": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]),
    "ak_exterior_state": (ctypes.c_int, [_vp, _dp]),
    "ak_save_state": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "ak_load_state": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "ak_refined_state_bytes": (ctypes.c_size_t, [_vp]),
    "ak_refined_save_state": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "ak_refined_load_state": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    "ak_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp]),
    "ak_batch_destroy": (ctypes.c_int, [_vp]),
    "ak_batch_add": (ctypes.c_int, [_vp, _vp]),
    "ak_batch_remove": (ctypes.c_int, [_vp, _vp]),
    "ak_batch_size": (ctypes.c_int, [_vp]),
    "ak_batch_set_graph_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ak_batch_set_sweep_grid": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ak_batch_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ak_batch_synchronize": (ctypes.c_int, [_vp]),
    "ak_refined_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _ip, _ip, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         _vp]),
    "ak_refined_destroy": (ctypes.c_int, [_vp]),
    "ak_refined_level": (_vp, [_vp, ctypes.c_int]),
    "ak_refined_set_walls": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _dp, _ip, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double]),
    "ak_refined_init_seeded": (ctypes.c_int, [_vp, _dp]),
    "ak_refined_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int]),
    "ak_refined_run_profiled": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, _dp]),
    "ak_refined_transfer_errors": (ctypes.c_int, [_vp]),
    "ak_refined_set_graph_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ak_refined_get_populations": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "ak_refined_budget": (ctypes.c_int, [_vp, _dp]),
}

_lib = None


def exported_symbols():
    return tuple(_SIGS)


def load():
    """Load the HIP library (no GPU call is made here)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NativeLibraryError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    try:
        lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
    except OSError as error:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {error}") from error
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error(handle=None):
    lib = load()
    buf = ctypes.create_string_buffer(1024)
    lib.ak_last_error(handle, buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, handle=None, what=""):
    if rc < 0:
        raise HipError(f"{what}: {last_error(handle)}")
    return rc


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(_u8p)
