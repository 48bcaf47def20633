"""Loads libminiiso.so (the HIP kernels behind the C ABI of include/miniiso.h) via ctypes.

There is no CPU fallback: if the library is missing or no GPU is present, calls fail loudly.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libminiiso.so")

MISO_OK = 0
ERRORS = {-1: "HIP error", -2: "allocation failed", -3: "invalid argument", -4: "call out of order",
          -5: "exceeds the admission budget"}



class MisoError(RuntimeError):
    pass


class Volume(ctypes.Structure):
    """struct miso_volume"""
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("nx", ctypes.c_int64),
        ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64),
        ("spacing", ctypes.c_float * 3),
        ("origin", ctypes.c_float * 3),
        ("on_device", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MisoError(f"{LIB_PATH} is not built; run `python -m miniisosurface_amd.build`")
    # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7. Loading torch first
    # makes libminiiso.so bind to that same runtime (matched by soname), so device pointers and
    # streams from torch tensors are valid in the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    PP = ctypes.POINTER(ctypes.c_void_p)
    U64P = ctypes.POINTER(ctypes.c_uint64)
    sig = {
        "miso_fe_create": [ctypes.POINTER(Volume), F, P, PP],
        "miso_fe_set_isoval": [P, F],
        "miso_fe_pass": [P, I],
        "miso_fe_run": [P],
        "miso_fe_counts": [P, U64P, U64P],
        "miso_fe_copy_out": [P, P, P, P],
        "miso_fe_device_output": [P, PP, PP, PP, ctypes.POINTER(I)],
        "miso_fe_set_index_bits": [P, I],
        "miso_fe_pass_times": [P, P],
        "miso_fe_set_timing": [P, ctypes.c_uint],
        "miso_fe_rows": [P, P, P, P, P, P, P, P, P],
        "miso_fe_cube_cases": [P, P, I],
        "miso_fe_slabs_create": [ctypes.POINTER(Volume), F, ctypes.c_int64, P, PP],
        "miso_fe_slabs_next": [P, ctypes.POINTER(I)],
        "miso_fe_slabs_restart": [P, F],
        "miso_fe_slabs_part": [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), U64P, U64P, U64P,
                               U64P],
        "miso_fe_slabs_copy_part": [P, P, P, P],
        "miso_fe_slabs_device_part": [P, PP, PP, PP],
        "miso_mc_create": [ctypes.POINTER(Volume), F, P, PP],
        "miso_mc_run": [P],
        "miso_mc_counts": [P, U64P, U64P],
        "miso_mc_copy_out": [P, P, P, P],
        "miso_mc_device_output": [P, PP, PP, PP, ctypes.POINTER(I)],
        "miso_mc_time_ms": [P, ctypes.POINTER(F)],
        "miso_mc_set_order": [P, I],
        "miso_batch_create": [ctypes.POINTER(Volume), ctypes.POINTER(F), ctypes.c_int64, P, PP],
        "miso_batch_run": [P],
        "miso_batch_totals": [P, U64P, U64P],
        "miso_batch_job_counts": [P, ctypes.c_int64, U64P, U64P],
        "miso_batch_job_output": [P, ctypes.c_int64, PP, PP, PP],
        "miso_batch_copy_out": [P, ctypes.c_int64, P, P, P],
        "miso_batch_set_tuning": [P, I, I],
        "miso_batch_get_tuning": [P, ctypes.POINTER(I), ctypes.POINTER(I)],
        "miso_batch_set_queue_order": [P, I],
        "miso_batch_get_queue_order": [P, ctypes.POINTER(I)],
        "miso_batch_set_launch": [P, I],
        "miso_batch_get_launch": [P, ctypes.POINTER(I)],
        "miso_batch_autotune": [P, I, ctypes.POINTER(F)],
        "miso_batch_time_ms": [P, ctypes.POINTER(F)],
        "miso_batch_create_empty": [P, PP],
        "miso_batch_set_jobs": [P, ctypes.POINTER(Volume), ctypes.POINTER(F), ctypes.c_int64],
        "miso_batch_count_async": [P],
        "miso_batch_query": [P],
        "miso_batch_count_wait": [P],
        "miso_batch_job_offsets": [P, ctypes.c_int64, U64P, U64P],
        "miso_batch_emit_async": [P, ctypes.c_int64, P, P, P, ctypes.c_uint64, ctypes.c_uint64],
        "miso_batch_wait": [P],
        "miso_fe_run_noalloc": [ctypes.POINTER(Volume), F, P, P, P, P, ctypes.c_uint64, ctypes.c_uint64, U64P,
                                U64P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    for name in ("miso_fe_destroy", "miso_fe_slabs_destroy", "miso_mc_destroy", "miso_batch_destroy"):
        getattr(L, name).argtypes = [P]
        getattr(L, name).restype = None
    L.miso_last_error.restype = ctypes.c_char_p
    L.miso_version.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != MISO_OK:
        msg = lib().miso_last_error().decode(errors="replace")
        raise MisoError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")
