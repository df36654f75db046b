"""Minimal HIP runtime access for the binding: device-to-device copies of
memory owned by the library, and non-blocking streams."""

from __future__ import annotations

import ctypes

_hip = None


def _runtime():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpyAsync.restype = ctypes.c_int
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_int, ctypes.c_void_p]
        _hip.hipStreamSynchronize.restype = ctypes.c_int
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipStreamCreateWithFlags.restype = ctypes.c_int
        _hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        _hip.hipStreamDestroy.restype = ctypes.c_int
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return _hip


def memcpy_d2d(dst: int, src: int, nbytes: int, stream: int = 0) -> None:
    """Copy on `stream` and wait for it (no implicit legacy-stream ordering,
    so a copy never waits on, or breaks, another thread's graph capture on a
    non-blocking stream)."""
    hip = _runtime()
    rc = hip.hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3,
                            ctypes.c_void_p(stream or None))
    if rc == 0:
        rc = hip.hipStreamSynchronize(ctypes.c_void_p(stream or None))
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync device-to-device failed with code {rc}")


def stream_create_nonblocking() -> int:
    """A new stream that does not synchronise with the legacy default stream
    (hipStreamNonBlocking); default priority."""
    h = ctypes.c_void_p()
    rc = _runtime().hipStreamCreateWithFlags(ctypes.byref(h), 1)
    if rc != 0:
        raise RuntimeError(f"hipStreamCreateWithFlags failed with code {rc}")
    return int(h.value)


def stream_destroy(handle: int) -> None:
    _runtime().hipStreamDestroy(ctypes.c_void_p(handle))
