"""Array-level Python interface (torch ROCm tensors in, torch tensors out).

Public names, argument meaning, shapes and error texts follow the
reference binding's Toeplitz path — ``make_config``, ``num_tri``,
``phi2_toep_shape``, ``calc_mask``, ``calc_toep``, ``aha_toeplitz``,
``create_op`` / ``AhAOp`` (reference: python/mr_aha/api.py:211-324,
380-701, 736-1255). Arrays are C-contiguous torch tensors on the current HIP
device, row-major with the column-major shapes of the C API reversed:

    x, AhAx      (L, [Nz,] Ny, Nx)         complex64
    SEs          (C, [Nz,] Ny, Nx)         complex64
    phi2         (L, L, [Nkz,] Ntrajs)     complex64, phi2[j, i, ...] = block (i, j)
    traj         (ndim, Nkx, Ntrajs)       float32, radians in [-pi, pi)
    phi2_toep    (T, [Mz,] My, Mx) or (T, num_masked)   complex64
    mask         (num_masked,)             int64, 0-based centered indices

Every call goes through ``libmr_aha_c.so``; nothing is computed in Python.

Layout of this module: ``_Arg`` checks one tensor argument and yields its
device pointer; ``_Geometry`` derives the grid sizes a configuration implies;
``_OpRequest`` gathers and checks everything ``mr_op_create`` takes, so that
``AhAOp.create_op`` only sequences it.
"""

from __future__ import annotations

import ctypes
import numbers
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from ._lib import (CG_GRAPH, CG_ZERO_INIT, DEFAULT_MASK_RADIUS, KERNEL_BF16, KERNEL_FP16,
                   KERNEL_FP32, OP_ALL_DEVICES, OP_SHARE_PRIMARY, OP_TOEP_SOURCES, MRConfig,
                   MRError, MROpInfo, MROpInputs, MROpOptions, OpMethod, ToepDist, TrajType,
                   Triangulation, check, lib)

__all__ = ["DEFAULT_MASK_RADIUS", "MRConfig", "MRError", "TrajType", "Triangulation",
           "make_config", "num_tri", "phi2_toep_shape", "calc_mask", "calc_toep",
           "calc_toep_from_psf", "aha_toeplitz", "AhAOp", "create_op"]


# ---------------------------------------------------------------------------
# Tensor arguments
# ---------------------------------------------------------------------------

class _Arg:
    """Checks of one array argument; the C ABI takes raw device pointers."""

    _DTYPE = {"complex64": torch.complex64, "float32": torch.float32, "int64": torch.int64}

    @classmethod
    def ptr(cls, t, name: str, dtype: str, optional: bool = False):
        """Device address of ``t`` (None for an omitted optional argument)."""
        if t is None and optional:
            return None
        if not isinstance(t, torch.Tensor):
            what = "None" if t is None else type(t).__name__
            raise TypeError(f"{name} must be a torch.Tensor on the GPU, got {what}")
        if t.dtype != cls._DTYPE[dtype]:
            raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
        if t.device.type != "cuda":
            raise TypeError(f"{name} must live on the GPU (device 'cuda'), got {t.device}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be C-contiguous; call .contiguous() first")
        return t.data_ptr()

    @staticmethod
    def numel(t, expected: int, name: str) -> None:
        """Element count check (the library indexes every buffer densely)."""
        got = None if t is None else int(t.numel())
        if got != int(expected):
            raise ValueError(f"{name} has {got} elements but {expected} were expected")

    @staticmethod
    def stream(stream):
        """Raw hipStream_t of None, an int handle or a torch.cuda.Stream."""
        if stream is None:
            return None
        if isinstance(stream, int):
            return stream or None
        handle = getattr(stream, "cuda_stream", None)
        if handle is None:
            raise TypeError("stream must be None, an int handle or a torch.cuda.Stream")
        return int(handle) or None


# ---------------------------------------------------------------------------
# Configuration
# ---------------------------------------------------------------------------

def _cfg_copy(cfg: MRConfig) -> MRConfig:
    out = MRConfig()
    ctypes.memmove(ctypes.byref(out), ctypes.byref(cfg), ctypes.sizeof(MRConfig))
    return out


def _check_cfg(cfg: MRConfig, *, mask_may_be_pending: bool = False) -> None:
    """mr_config_validate. With ``mask_may_be_pending`` a masked configuration
    whose location count is still 0 (its mask not built yet) is checked as
    unmasked, on a copy, so ``cfg`` keeps its flag."""
    if mask_may_be_pending and cfg.is_masked and cfg.num_masked == 0:
        cfg = _cfg_copy(cfg)
        cfg.is_masked = 0
    check(lib.mr_config_validate(ctypes.byref(cfg)), "mr_config_validate")


@dataclass(frozen=True)
class _Geometry:
    """Image and upsampled-grid sizes, column-major (x, y[, z])."""

    n: tuple
    M: tuple

    @classmethod
    def from_args(cls, tt: int, n_spatial: int, n_dim, phi2_dim) -> "_Geometry":
        n = tuple(int(v) for v in n_dim)
        if len(n) != n_spatial:
            raise ValueError(f"n_dim must have {n_spatial} entries for this trajectory type, "
                             f"got {len(n)}")
        if phi2_dim is not None:
            M = tuple(int(v) for v in phi2_dim)
            if len(M) != n_spatial:
                raise ValueError(f"phi2_dim must have {n_spatial} entries, got {len(M)}")
        else:
            # 2x oversampling, except the Cartesian kz axis of a stack
            M = tuple(2 * v if not (tt == TrajType.STACK and a == 2) else v
                      for a, v in enumerate(n))
        return cls(n, M)

    @classmethod
    def of(cls, cfg: MRConfig) -> "_Geometry":
        k = cfg.n_spatial
        return cls(tuple(cfg.n_dim[:k]), tuple(cfg.phi2_dim[:k]))

    def store(self, cfg: MRConfig) -> None:
        k = len(self.n)
        cfg.n_dim[:k] = list(self.n)
        cfg.phi2_dim[:k] = list(self.M)

    def image_shape(self) -> tuple:
        return tuple(reversed(self.n))

    def grid_shape(self) -> tuple:
        return tuple(reversed(self.M))


def make_config(traj_type, n_dim: Sequence[int], L: int, *, phi2_dim=None,
                triangulation="lower", is_masked: bool = False, num_masked: int = 0,
                n_coils: int = 1, eps: float = 1e-6, gpu_device_id: int = -1, stream=None,
                mem_fraction: float = 0.5, verbose: bool = False,
                mask_radius: int = DEFAULT_MASK_RADIUS) -> MRConfig:
    """Validated configuration (reference: python/mr_aha/api.py:211-324).

    ``n_dim`` and ``phi2_dim`` are column-major ``(Nx, Ny[, Nz])``; the
    default grid is ``2 * n_dim`` except ``Mz = Nz`` for a stack. A masked
    configuration may be built before its mask exists (num_masked 0): the
    count is adopted from the mask by ``calc_toep`` / ``create_op``.
    """
    tt = TrajType.parse(traj_type)
    cfg = MRConfig()
    check(lib.mr_config_defaults(ctypes.byref(cfg), tt), "mr_config_defaults")
    _Geometry.from_args(tt, cfg.n_spatial, n_dim, phi2_dim).store(cfg)
    fields = {
        "L": int(L),
        "triangulation": Triangulation.parse(triangulation),
        "is_masked": int(bool(is_masked)),
        "num_masked": int(num_masked),
        "n_coils": int(n_coils),
        "eps": float(eps),
        "gpu_device_id": int(gpu_device_id),
        "mem_fraction": float(mem_fraction),
        "verbose": int(bool(verbose)),
        "mask_radius": int(mask_radius),
        "stream": _Arg.stream(stream),
    }
    for key, value in fields.items():
        setattr(cfg, key, value)
    _check_cfg(cfg, mask_may_be_pending=True)
    return cfg


def num_tri(L: int) -> int:
    return int(lib.mr_num_tri(int(L)))


def phi2_toep_shape(cfg: MRConfig) -> tuple:
    """Row-major kernel shape: (T, [Mz,] My, Mx), or (T, num_masked)."""
    T = num_tri(cfg.L)
    return (T, int(cfg.num_masked)) if cfg.is_masked else (T,) + _Geometry.of(cfg).grid_shape()


def _size(fn, cfg: MRConfig) -> int:
    return int(fn(ctypes.byref(cfg)))


def _readouts(cfg: MRConfig, traj) -> tuple:
    """(Ntrajs, Nkx) of a trajectory shaped (ndim, Nkx, Ntrajs)."""
    ndim = 3 if cfg.traj_type == TrajType.THREE_D else 2
    shape = None if traj is None else tuple(traj.shape)
    if shape is None or len(shape) != 3 or shape[0] != ndim:
        raise ValueError(f"traj must be shaped (ndim, Nkx, Ntrajs) with ndim = {ndim}, "
                         f"got {shape}")
    return shape[2], shape[1]


def _adopt_mask(cfg: MRConfig, mask) -> None:
    """A masked configuration needs its mask; a count still 0 is taken from it."""
    if not cfg.is_masked:
        return
    if mask is None:
        raise ValueError("cfg.is_masked is set, so a mask is required")
    count = int(mask.numel())
    if cfg.num_masked == 0:
        cfg.num_masked = count
        _check_cfg(cfg)
    elif count != int(cfg.num_masked):
        raise ValueError(f"mask has {count} elements but cfg.num_masked is {cfg.num_masked}")


def _gram_inputs(cfg: MRConfig, phi2, traj, mask) -> tuple:
    """Check the Gram blocks against the trajectory; returns (Ntrajs, Nkx)."""
    ntr, nkx = _readouts(cfg, traj)
    planes = cfg.phi2_dim[2] if cfg.traj_type == TrajType.STACK else 1
    _Arg.numel(phi2, cfg.L * cfg.L * planes * ntr, "phi2")
    _adopt_mask(cfg, mask)
    return ntr, nkx


# ---------------------------------------------------------------------------
# Stateless primitives
# ---------------------------------------------------------------------------

def calc_mask(cfg: MRConfig, traj, radius: Optional[int] = None):
    """Neighbourhood mask of the trajectory (reference: api.py:461-505)."""
    c = _cfg_copy(cfg)
    if radius is not None:
        c.mask_radius = int(radius)
    ntr, nkx = _readouts(c, traj)
    out = torch.empty(_size(lib.mr_pad_size, c), dtype=torch.int64, device=traj.device)
    count = ctypes.c_int64(0)
    check(lib.mr_calc_mask(ctypes.byref(c), _Arg.ptr(traj, "traj", "float32"), ntr, nkx,
                           _Arg.ptr(out, "mask", "int64"), ctypes.byref(count)), "mr_calc_mask")
    return out[: count.value].clone()


def calc_toep(cfg: MRConfig, phi2, traj, mask=None, out=None):
    """Toeplitz kernel ``phi2_toep`` (reference: api.py:556-639)."""
    ntr, nkx = _gram_inputs(cfg, phi2, traj, mask)
    if out is None:
        out = torch.empty(phi2_toep_shape(cfg), dtype=torch.complex64, device=phi2.device)
    else:
        _Arg.numel(out, _size(lib.mr_phi2_toep_size, cfg), "out")
    check(lib.mr_calc_toep(ctypes.byref(cfg), _Arg.ptr(phi2, "phi2", "complex64"),
                           _Arg.ptr(traj, "traj", "float32"), ctypes.c_int64(ntr),
                           ctypes.c_int64(nkx), _Arg.ptr(mask, "mask", "int64", optional=True),
                           _Arg.ptr(out, "out", "complex64")), "mr_calc_toep")
    return out


def calc_toep_from_psf(cfg: MRConfig, psf, t0: int, phi2_toep, mask=None):
    """Finish components [t0, t0 + n_comp) of ``phi2_toep`` from centered
    point-spread functions ``psf`` (n_comp, [Mz,] My, Mx)."""
    per_comp = 1
    for v in _Geometry.of(cfg).M:
        per_comp *= int(v)
    check(lib.mr_calc_toep_from_psf(ctypes.byref(cfg), _Arg.ptr(psf, "psf", "complex64"),
                                    ctypes.c_int64(int(t0)),
                                    ctypes.c_int64(int(psf.numel()) // per_comp),
                                    _Arg.ptr(mask, "mask", "int64", optional=True),
                                    _Arg.ptr(phi2_toep, "phi2_toep", "complex64")),
          "mr_calc_toep_from_psf")
    return phi2_toep


def aha_toeplitz(cfg: MRConfig, x, SEs, phi2_toep, mask=None, out=None):
    """One-shot Toeplitz A^H A (reference: api.py:642-701)."""
    for t, fn, name in ((x, lib.mr_x_size, "x"), (SEs, lib.mr_se_size, "SEs"),
                        (phi2_toep, lib.mr_phi2_toep_size, "phi2_toep")):
        _Arg.numel(t, _size(fn, cfg), name)
    if cfg.is_masked and mask is None:
        raise ValueError("cfg.is_masked is set, so a mask is required")
    out = _output_like(x, out)
    check(lib.mr_aha_toeplitz(ctypes.byref(cfg), _Arg.ptr(x, "x", "complex64"),
                              _Arg.ptr(SEs, "SEs", "complex64"),
                              _Arg.ptr(phi2_toep, "phi2_toep", "complex64"),
                              _Arg.ptr(mask, "mask", "int64", optional=True),
                              _Arg.ptr(out, "out", "complex64")), "mr_aha_toeplitz")
    return out


def _output_like(x, out):
    if out is None:
        return torch.empty(x.shape, dtype=torch.complex64, device=x.device)
    _Arg.numel(out, x.numel(), "out")
    return out


# ---------------------------------------------------------------------------
# Operator
# ---------------------------------------------------------------------------

_NUFFT_OPTS = ("upsampfac", "gpu_method")


def _device_ids(devices):
    """(ctypes int array or None, count) for the ``devices`` argument:
    None (the current device), "all", one id or a sequence of ids."""
    if isinstance(devices, str):
        if devices.lower() == "all":
            return None, OP_ALL_DEVICES
        raise ValueError(f"devices must be None, 'all', an int or a list; got {devices!r}")
    if devices is None:
        ids = []
    elif isinstance(devices, numbers.Integral):
        ids = [int(devices)]
    else:
        ids = [int(d) for d in devices]
    return ((ctypes.c_int * len(ids))(*ids), len(ids)) if ids else (None, 0)


@dataclass
class _OpRequest:
    """Everything mr_op_create receives, checked before the call."""

    cfg: MRConfig
    method: str
    SEs: object
    phi2: object
    traj: object
    mask: object
    kernel: object

    @property
    def nufft(self) -> bool:
        return self.method == "nufft"

    def check_method(self, nufft_opts: dict) -> None:
        if self.nufft and self.cfg.is_masked:
            raise ValueError("method='nufft' has no kernel to mask; build the configuration "
                             "with is_masked=False")
        if nufft_opts and not self.nufft:
            raise ValueError("nufft_opts is only used with method='nufft'")
        unknown = sorted(set(nufft_opts) - set(_NUFFT_OPTS))
        if unknown:
            raise ValueError(f"unknown nufft_opts {unknown}; accepted: {list(_NUFFT_OPTS)}")
        if self.kernel is not None and self.nufft:
            raise ValueError("_phi2_toep is only used with method='toeplitz'")

    def readouts(self) -> tuple:
        """(Ntrajs, Nkx); (0, 0) when a supplied kernel replaces phi2 and traj."""
        kernel_only = (self.kernel is not None and self.phi2 is None and self.traj is None
                       and not self.nufft)
        if kernel_only:
            _adopt_mask(self.cfg, self.mask)
            return 0, 0
        if self.phi2 is None or self.traj is None:
            raise ValueError("phi2 and traj are required (the operator is built from the "
                             "Gram blocks)")
        return _gram_inputs(self.cfg, self.phi2, self.traj, self.mask)

    def check_sizes(self) -> None:
        _Arg.numel(self.SEs, _size(lib.mr_se_size, self.cfg), "SEs")
        if self.kernel is not None:
            _Arg.numel(self.kernel, _size(lib.mr_phi2_toep_size, self.cfg), "_phi2_toep")

    def options(self, devices, toep_dist, share_primary: bool, nufft_opts: dict):
        opt = MROpOptions()
        check(lib.mr_op_options_defaults(ctypes.byref(opt)), "mr_op_options_defaults")
        ids, count = _device_ids(devices)
        opt.method = OpMethod.parse(self.method)
        opt.devices = ctypes.cast(ids, ctypes.POINTER(ctypes.c_int)) if ids else None
        opt.n_devices = count
        opt.toep_dist = ToepDist.parse(toep_dist)
        opt.flags = OP_SHARE_PRIMARY if share_primary else 0
        opt.nufft_upsampfac = float(nufft_opts.get("upsampfac", 0.0))
        opt.nufft_gpu_method = int(nufft_opts.get("gpu_method", 0))
        return opt, ids

    def inputs(self, ntr: int, nkx: int) -> MROpInputs:
        have_kernel = self.kernel is not None
        return MROpInputs(
            _Arg.ptr(self.SEs, "SEs", "complex64"),
            _Arg.ptr(self.phi2, "phi2", "complex64", optional=have_kernel),
            _Arg.ptr(self.traj, "traj", "float32", optional=have_kernel),
            int(ntr), int(nkx),
            _Arg.ptr(self.mask, "mask", "int64", optional=True),
            _Arg.ptr(self.kernel, "_phi2_toep", "complex64", optional=True))


class AhAOp:
    """The A^H A operator (reference: python/mr_aha/api.py:736-1223).

    Create with :func:`create_op`; apply with ``op(x)`` or ``op.apply(x, out)``.
    """

    def __init__(self, *args, **kwargs):
        raise TypeError("AhAOp is created with mr_aha_amd.create_op(cfg, SEs, phi2, traj, ...) "
                        "(or AhAOp.create_op), not by calling the class")

    @classmethod
    def create_op(cls, cfg: MRConfig, SEs, phi2, traj, mask=None, *, method: str = "toeplitz",
                  devices=None, toep_dist: str = "auto", share_primary: bool = True,
                  nufft_opts: Optional[dict] = None, _phi2_toep=None) -> "AhAOp":
        req = _OpRequest(cfg, OpMethod._TO_STR[OpMethod.parse(method)], SEs, phi2, traj, mask,
                         _phi2_toep)
        nufft_opts = dict(nufft_opts or {})
        req.check_method(nufft_opts)
        req.check_sizes()
        ntr, nkx = req.readouts()
        opt, ids = req.options(devices, toep_dist, share_primary, nufft_opts)
        inp = req.inputs(ntr, nkx)
        handle = ctypes.c_void_p()
        check(lib.mr_op_create(ctypes.byref(cfg), ctypes.byref(inp), ctypes.byref(opt),
                               ctypes.byref(handle)), "mr_op_create")
        del ids  # the library copied the device list
        op = object.__new__(cls)
        op._handle = handle
        op._cfg = _cfg_copy(cfg)
        op._method = req.method
        op._device = SEs.device
        # arrays the operator aliases (MR_OP_SHARE_PRIMARY) stay alive with it
        op._keep = (SEs, mask, _phi2_toep) if share_primary else ()
        op._kfmt = "fp32"
        return op

    @property
    def config(self) -> MRConfig:
        return self._cfg

    @property
    def method(self) -> str:
        return self._method

    def _open(self):
        if not self._handle:
            raise RuntimeError("this AhAOp has already been closed")
        return self._handle

    def apply(self, x, out=None):
        h = self._open()
        _Arg.numel(x, _size(lib.mr_x_size, self._cfg), "x")
        out = _output_like(x, out)
        check(lib.mr_op_execute(h, _Arg.ptr(x, "x", "complex64"),
                                _Arg.ptr(out, "out", "complex64")), "mr_op_execute")
        return out

    __call__ = apply

    @property
    def info(self) -> dict:
        inf = MROpInfo()
        check(lib.mr_op_info(self._open(), ctypes.byref(inf)), "mr_op_info")
        n = inf.n_devices
        per_device = {k: list(getattr(inf, k)[:n])
                      for k in ("devices", "coil_begin", "coil_count", "group")}
        return dict(method=OpMethod._TO_STR.get(int(inf.method), "unknown"), **per_device,
                    n_split_coils=int(inf.n_split_coils), peer_access=bool(inf.peer_access),
                    peer_bandwidth_gbs=float(inf.peer_bandwidth_gbs),
                    toep_source=OP_TOEP_SOURCES.get(int(inf.toep_source), "unknown"),
                    toep_devices=int(inf.toep_devices), toep_ms=float(inf.toep_ms),
                    toep_component_ms=float(inf.toep_component_ms),
                    toep_gather_gbs=float(inf.toep_gather_gbs), create_ms=float(inf.create_ms))

    def kernel(self, k: int = 0):
        """Copy of the kernel the operator holds, shaped ``phi2_toep_shape``.

        The reference returns a view of operator memory; torch cannot wrap a
        foreign device pointer, so a device-to-device copy is returned.
        """
        ptr = ctypes.c_void_p()
        check(lib.mr_op_kernel(self._open(), int(k), ctypes.byref(ptr)), "mr_op_kernel")
        out = torch.empty(phi2_toep_shape(self._cfg), dtype=torch.complex64, device=self._device)
        _memcpy_d2d(out.data_ptr(), ptr.value, out.numel() * 8)
        return out

    @property
    def xbd_group(self) -> int:
        """Coils per workgroup of the fused x pass (0: the separate passes run).
        No counterpart in the reference."""
        return int(lib.mr_op_xbd_group(self._open()))

    @property
    def xbd_form(self) -> int:
        """Form of the fused x pass: 0 none, 1 one workgroup per row and coil
        group, 2 persistent, 3 block step on the matrix cores. No counterpart
        in the reference."""
        return int(lib.mr_op_xbd_form(self._open()))

    def set_kernel_format(self, fmt: str) -> None:
        """Kernel storage read by the applies: "fp32", "bf16" or "fp16"
        (library entry mr_op_set_kernel_format; 16-bit entries with one FP32
        scale per 4x4 block, for the shapes with a compiled fused x pass).
        No counterpart in the reference."""
        codes = {"fp32": KERNEL_FP32, "bf16": KERNEL_BF16, "fp16": KERNEL_FP16}
        if fmt not in codes:
            raise ValueError(f"kernel format {fmt!r}: expected 'fp32', 'bf16' or 'fp16'")
        check(lib.mr_op_set_kernel_format(self._open(), codes[fmt]), "mr_op_set_kernel_format")
        self._kfmt = fmt

    @property
    def kernel_format(self) -> str:
        return self._kfmt

    def autotune(self, x=None, reps: int = 2) -> dict:
        """Time the tile widths of the apply's FFT passes and the fused x
        pass's two-coil form on this operator and keep the fastest (library
        entry mr_op_autotune; results do not depend on the choice). ``x`` is
        the image the trials apply to (default: a seeded complex Gaussian of
        the operator's shape). Returns the tiling now in use (`tiles`). No
        counterpart in the reference."""
        h = self._open()
        nx = _size(lib.mr_x_size, self._cfg)
        if x is None:
            g = torch.Generator(device="cpu").manual_seed(1234)
            x = torch.complex(torch.randn(nx, generator=g), torch.randn(nx, generator=g))
            x = x.to(self._device)
        _Arg.numel(x, nx, "x")
        out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
        torch.cuda.synchronize(self._device)   # x written on torch's stream is complete
        check(lib.mr_op_autotune(h, _Arg.ptr(x, "x", "complex64"),
                                 _Arg.ptr(out, "out", "complex64"), int(reps)), "mr_op_autotune")
        return self.tiles

    @property
    def tiles(self) -> dict:
        """Tiling in use (library entry mr_op_tiles): pencils per workgroup
        requested for the outer-axis, y, inverse y and inverse outer-axis FFT
        passes (0: the built-in rule), whether the fused x pass runs
        two-coil workgroups and whether its block step runs on the matrix
        cores."""
        t = (ctypes.c_int * 6)()
        check(lib.mr_op_tiles(self._open(), t, 6), "mr_op_tiles")
        return dict(fo=t[0], fy=t[1], iy=t[2], io=t[3], wide=bool(t[4]), mfma=bool(t[5]))

    def cg(self, b, x=None, *, lam: float = 0.0, iters: int = 40, refresh: int = 10,
           op_refresh: Optional["AhAOp"] = None, graph: bool = False, noalloc: bool = False):
        """Conjugate gradients on (A^H A + lam I) x = b (library entry mr_op_cg;
        ``noalloc``: mr_op_cg_noalloc, the same solve on the library's static
        work array, without graphs).

        FP32 vectors, FP64 inner products and step lengths; every ``refresh``
        iterations the residual is recomputed as b - (A^H A + lam I) x with
        ``op_refresh`` (default: this operator). ``x`` is the initial guess
        (zeros when None) and is overwritten. ``graph`` replays HIP graphs of
        the iterations (needs a configuration built with a non-default stream).
        Returns ``(x, hist)`` with hist[k] = ||r_k|| / ||b||, k = 0..iters.
        No counterpart in the reference, whose callers run CG in Python.
        """
        h = self._open()
        n = _size(lib.mr_x_size, self._cfg)
        _Arg.numel(b, n, "b")
        flags = CG_GRAPH if graph else 0
        if x is None:
            x = torch.empty(b.shape, dtype=torch.complex64, device=b.device)
            flags |= CG_ZERO_INIT
        else:
            _Arg.numel(x, n, "x")
        ref = op_refresh._open() if op_refresh is not None else None
        hist = (ctypes.c_double * (int(iters) + 1))()
        fn, name = (lib.mr_op_cg_noalloc, "mr_op_cg_noalloc") if noalloc else (lib.mr_op_cg,
                                                                                "mr_op_cg")
        check(fn(h, ref, _Arg.ptr(b, "b", "complex64"), _Arg.ptr(x, "x", "complex64"),
                 float(lam), int(iters), int(refresh), flags, hist), name)
        return x, list(hist)

    def close(self) -> None:
        if getattr(self, "_handle", None):
            lib.mr_op_destroy(self._handle)
            self._handle = ctypes.c_void_p()
        self._keep = ()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass


def _memcpy_d2d(dst: int, src: int, nbytes: int) -> None:
    from ._hiprt import memcpy_d2d
    memcpy_d2d(dst, src, nbytes, int(torch.cuda.current_stream().cuda_stream))


def create_op(cfg: MRConfig, SEs, phi2, traj, mask=None, **kwargs) -> AhAOp:
    """Create the A^H A operator; alias of :meth:`AhAOp.create_op`."""
    return AhAOp.create_op(cfg, SEs, phi2, traj, mask, **kwargs)
