"""Minimal HIP runtime access for the binding (device-to-device copies of
memory owned by the library)."""

from __future__ import annotations

import ctypes
import ctypes.util

_hip = None


def _runtime():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    rc = _runtime().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy device-to-device failed with code {rc}")
