"""ctypes binding to ``libmr_aha_c.so`` (the MI355X build of the C ABI).

Struct mirrors follow ``include/mr_aha/mr_aha.h`` field for field (the same
layouts as the reference's header, include/mr_aha/mr_aha.h:332-414, 814-952
there; python/mr_aha/_lib.py:213-448 mirrors them the same way). The library
is loaded at import and the import fails loudly when it is missing: there is
no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

__all__ = ["MRConfig", "MROpInputs", "MROpOptions", "MROpInfo", "MRError", "TrajType",
           "Triangulation", "OpMethod", "ToepDist", "lib", "check", "backend", "LIB_PATH",
           "DEFAULT_MASK_RADIUS", "OP_MAX_DEVICES", "OP_ALL_DEVICES", "OP_SHARE_PRIMARY",
           "OP_TOEP_SOURCES"]

DEFAULT_MASK_RADIUS = 2
OP_MAX_DEVICES = 16
OP_ALL_DEVICES = -1
OP_SHARE_PRIMARY = 1
OP_TOEP_SOURCES = {0: "supplied", 1: "broadcast", 2: "split", 3: "replicate", 4: "none"}


class _Enum:
    _FROM_STR: dict = {}
    _NAME = "value"

    @classmethod
    def parse(cls, value) -> int:
        if isinstance(value, int):
            return value
        try:
            return cls._FROM_STR[str(value).lower()]
        except KeyError:
            raise ValueError(f"{cls._NAME} must be one of {sorted(cls._FROM_STR)}; "
                             f"got {value!r}") from None


class TrajType(_Enum):
    TWO_D, THREE_D, STACK = 0, 1, 2
    _FROM_STR = {"2d": 0, "3d": 1, "stack": 2}
    _NAME = "traj_type"


class Triangulation(_Enum):
    LOWER, UPPER = 0, 1
    _FROM_STR = {"lower": 0, "upper": 1}
    _NAME = "triangulation"


class OpMethod(_Enum):
    TOEPLITZ, NUFFT = 0, 1
    _FROM_STR = {"toeplitz": 0, "nufft": 1}
    _TO_STR = {0: "toeplitz", 1: "nufft"}
    _NAME = "method"


class ToepDist(_Enum):
    AUTO, SPLIT, REPLICATE, BROADCAST = 0, 1, 2, 3
    _FROM_STR = {"auto": 0, "split": 1, "replicate": 2, "broadcast": 3, "off": 3}
    _NAME = "toep_dist"


class MRConfig(ctypes.Structure):
    _fields_ = [
        ("traj_type", ctypes.c_int),
        ("n_spatial", ctypes.c_int),
        ("n_dim", ctypes.c_int * 3),
        ("phi2_dim", ctypes.c_int * 3),
        ("L", ctypes.c_int),
        ("triangulation", ctypes.c_int),
        ("is_masked", ctypes.c_int),
        ("num_masked", ctypes.c_int64),
        ("n_coils", ctypes.c_int),
        ("eps", ctypes.c_double),
        ("gpu_device_id", ctypes.c_int),
        ("stream", ctypes.c_void_p),
        ("mem_fraction", ctypes.c_double),
        ("verbose", ctypes.c_int),
        ("mask_radius", ctypes.c_int),
    ]


class MROpInputs(ctypes.Structure):
    _fields_ = [
        ("d_SEs", ctypes.c_void_p),
        ("d_phi2", ctypes.c_void_p),
        ("d_traj", ctypes.c_void_p),
        ("Ntrajs", ctypes.c_int64),
        ("Nkx", ctypes.c_int64),
        ("d_mask", ctypes.c_void_p),
        ("d_phi2_toep", ctypes.c_void_p),
    ]


class MROpOptions(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("devices", ctypes.POINTER(ctypes.c_int)),
        ("n_devices", ctypes.c_int),
        ("toep_dist", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("nufft_upsampfac", ctypes.c_double),
        ("nufft_gpu_method", ctypes.c_int),
    ]


class MROpInfo(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("n_devices", ctypes.c_int),
        ("devices", ctypes.c_int * OP_MAX_DEVICES),
        ("coil_begin", ctypes.c_int * OP_MAX_DEVICES),
        ("coil_count", ctypes.c_int * OP_MAX_DEVICES),
        ("group", ctypes.c_int * OP_MAX_DEVICES),
        ("n_split_coils", ctypes.c_int),
        ("peer_access", ctypes.c_int),
        ("peer_bandwidth_gbs", ctypes.c_double),
        ("toep_source", ctypes.c_int),
        ("toep_devices", ctypes.c_int),
        ("toep_ms", ctypes.c_double),
        ("toep_component_ms", ctypes.c_double),
        ("toep_gather_gbs", ctypes.c_double),
        ("create_ms", ctypes.c_double),
    ]


class MRError(RuntimeError):
    """A non-zero status from the library; ``.status`` holds the code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def _find_library() -> Path:
    env = os.environ.get("MR_AHA_LIBRARY")
    if env:
        return Path(env)
    return Path(__file__).resolve().parent / "libmr_aha_c.so"


LIB_PATH = _find_library()
if not LIB_PATH.exists():
    raise OSError(f"libmr_aha_c.so not found at {LIB_PATH}; build it with "
                  f"`python -m mr_aha_amd.build` (hipcc, gfx950)")
lib = ctypes.CDLL(str(LIB_PATH))

_p = ctypes.c_void_p
_cfg = ctypes.POINTER(MRConfig)


def _decl(name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


_decl("mr_status_string", ctypes.c_char_p, [ctypes.c_int])
_decl("mr_last_error", ctypes.c_char_p, [])
_decl("mr_backend", ctypes.c_char_p, [])
_decl("mr_version", ctypes.c_int, [])
_decl("mr_config_defaults", ctypes.c_int, [_cfg, ctypes.c_int])
_decl("mr_config_validate", ctypes.c_int, [_cfg])
_decl("mr_num_tri", ctypes.c_int64, [ctypes.c_int])
for _n in ("mr_image_size", "mr_pad_size", "mr_phi2_toep_size", "mr_x_size", "mr_se_size"):
    _decl(_n, ctypes.c_int64, [_cfg])
_decl("mr_aha_plan_create", ctypes.c_int, [_cfg, ctypes.POINTER(ctypes.c_void_p)])
_decl("mr_aha_plan_execute", ctypes.c_int, [_p, _p, _p, _p, _p, _p])
_decl("mr_aha_plan_destroy", ctypes.c_int, [_p])
_decl("mr_aha_toeplitz", ctypes.c_int, [_cfg, _p, _p, _p, _p, _p])
_decl("mr_op_options_defaults", ctypes.c_int, [ctypes.POINTER(MROpOptions)])
_decl("mr_op_create", ctypes.c_int, [_cfg, ctypes.POINTER(MROpInputs),
                                     ctypes.POINTER(MROpOptions), ctypes.POINTER(ctypes.c_void_p)])
_decl("mr_op_execute", ctypes.c_int, [_p, _p, _p])
_decl("mr_op_info", ctypes.c_int, [_p, ctypes.POINTER(MROpInfo)])
_decl("mr_op_kernel", ctypes.c_int, [_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)])
_decl("mr_op_destroy", ctypes.c_int, [_p])
_decl("mr_op_abandon", ctypes.c_int, [_p])
_decl("mr_calc_toep", ctypes.c_int, [_cfg, _p, _p, ctypes.c_int64, ctypes.c_int64, _p, _p])
_decl("mr_calc_toep_from_psf", ctypes.c_int, [_cfg, _p, ctypes.c_int64, ctypes.c_int64, _p, _p])
_decl("mr_aha_plan_profile", ctypes.c_int, [_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_int), ctypes.c_int])
_decl("mr_op_profile", ctypes.c_int, [_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int), ctypes.c_int])
_decl("mr_aha_plan_xbd_group", ctypes.c_int, [_p])
_decl("mr_op_xbd_group", ctypes.c_int, [_p])
_decl("mr_aha_plan_xbd_form", ctypes.c_int, [_p])
_decl("mr_op_xbd_form", ctypes.c_int, [_p])
_decl("mr_aha_plan_autotune", ctypes.c_int, [_p, _p, _p, _p, _p, _p, ctypes.c_int])
_decl("mr_op_autotune", ctypes.c_int, [_p, _p, _p, ctypes.c_int])
_decl("mr_aha_plan_tiles", ctypes.c_int, [_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int])
_decl("mr_op_tiles", ctypes.c_int, [_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int])
_decl("mr_op_cg", ctypes.c_int, [_p, _p, _p, _p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(ctypes.c_double)])
_decl("mr_op_cg_noalloc", ctypes.c_int, [_p, _p, _p, _p, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)])
_decl("mr_stream_copy", ctypes.c_int, [_p, _p, ctypes.c_int64, _p])
CG_ZERO_INIT = 1
CG_GRAPH = 2
_decl("mr_op_set_kernel_format", ctypes.c_int, [_p, ctypes.c_int])
KERNEL_FP32, KERNEL_BF16, KERNEL_FP16 = 0, 1, 2
PROFILE_STAGES = ("F1", "F2", "F3", "BD", "I3", "I2", "I1", "other")
_decl("mr_calc_mask", ctypes.c_int, [_cfg, _p, ctypes.c_int64, ctypes.c_int64, _p,
                                     ctypes.POINTER(ctypes.c_int64)])


def backend() -> str:
    return lib.mr_backend().decode("ascii")


def check(status: int, what: str) -> None:
    """Raise MRError for a non-zero status (message: call, status text, detail)."""
    if status == 0:
        return
    summary = lib.mr_status_string(status).decode("utf-8", "replace")
    detail = lib.mr_last_error().decode("utf-8", "replace")
    raise MRError(status, f"{what}: {summary} ({status}): {detail}")
