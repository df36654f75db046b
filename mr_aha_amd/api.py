"""Array-level Python interface (torch ROCm tensors in, torch tensors out).

Mirrors the reference binding's public functions for the Toeplitz path —
``make_config``, ``num_tri``, ``phi2_toep_shape``, ``calc_mask``,
``calc_toep``, ``aha_toeplitz``, ``create_op`` / ``AhAOp`` (reference:
python/mr_aha/api.py:211-324, 380-701, 736-1255) — with the same argument
meaning, shapes and errors. Arrays are C-contiguous torch tensors on the
current HIP device, row-major with the column-major shapes of the C API
reversed:

    x, AhAx      (L, [Nz,] Ny, Nx)         complex64
    SEs          (C, [Nz,] Ny, Nx)         complex64
    phi2         (L, L, [Nkz,] Ntrajs)     complex64, phi2[j, i, ...] = block (i, j)
    traj         (ndim, Nkx, Ntrajs)       float32, radians in [-pi, pi)
    phi2_toep    (T, [Mz,] My, Mx) or (T, num_masked)   complex64
    mask         (num_masked,)             int64, 0-based centered indices

Every call goes through ``libmr_aha_c.so``; nothing is computed in Python.
"""

from __future__ import annotations

import ctypes
import numbers
from typing import Optional, Sequence

import torch

from ._lib import (CG_GRAPH, CG_ZERO_INIT, DEFAULT_MASK_RADIUS, KERNEL_BF16, KERNEL_FP16,
                   KERNEL_FP32, OP_ALL_DEVICES, OP_SHARE_PRIMARY, OP_TOEP_SOURCES, MRConfig, MRError, MROpInfo, MROpInputs, MROpOptions, OpMethod, ToepDist,
                   TrajType, Triangulation, check, lib)

__all__ = ["DEFAULT_MASK_RADIUS", "MRConfig", "MRError", "TrajType", "Triangulation",
           "make_config", "num_tri", "phi2_toep_shape", "calc_mask", "calc_toep",
           "calc_toep_from_psf", "aha_toeplitz", "AhAOp", "create_op"]

_DTYPES = {"complex64": torch.complex64, "float32": torch.float32, "int64": torch.int64}


def _ptr(t, name: str, dtype: str, allow_none: bool = False):
    """Device pointer of a contiguous HIP tensor of the required dtype."""
    if t is None:
        if allow_none:
            return None
        raise TypeError(f"{name} must not be None")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on the GPU, got {type(t).__name__}")
    if t.dtype != _DTYPES[dtype]:
        raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if t.device.type != "cuda":
        raise TypeError(f"{name} must live on the GPU (device 'cuda'), got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous; call .contiguous() first")
    return t.data_ptr()


def _check_size(t, expected: int, name: str) -> None:
    if t is None or int(t.numel()) != int(expected):
        got = None if t is None else int(t.numel())
        raise ValueError(f"{name} has {got} elements but {expected} were expected")


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream or None
    if hasattr(stream, "cuda_stream"):
        return int(stream.cuda_stream) or None
    raise TypeError("stream must be None, an int handle or a torch.cuda.Stream")


def make_config(traj_type, n_dim: Sequence[int], L: int, *, phi2_dim=None,
                triangulation="lower", is_masked: bool = False, num_masked: int = 0,
                n_coils: int = 1, eps: float = 1e-6, gpu_device_id: int = -1, stream=None,
                mem_fraction: float = 0.5, verbose: bool = False,
                mask_radius: int = DEFAULT_MASK_RADIUS) -> MRConfig:
    """Validated configuration (reference: python/mr_aha/api.py:211-324).

    ``n_dim`` and ``phi2_dim`` are column-major ``(Nx, Ny[, Nz])``; the
    default grid is ``2 * n_dim`` except ``Mz = Nz`` for a stack. A masked
    configuration may be built before its mask exists (num_masked 0).
    """
    tt = TrajType.parse(traj_type)
    cfg = MRConfig()
    check(lib.mr_config_defaults(ctypes.byref(cfg), tt), "mr_config_defaults")
    nsp = cfg.n_spatial
    n_dim = [int(v) for v in n_dim]
    if len(n_dim) != nsp:
        raise ValueError(f"n_dim must have {nsp} entries for this trajectory type, got {len(n_dim)}")
    if phi2_dim is None:
        phi2_dim = [2 * v for v in n_dim]
        if tt == TrajType.STACK:
            phi2_dim[2] = n_dim[2]
    else:
        phi2_dim = [int(v) for v in phi2_dim]
    if len(phi2_dim) != nsp:
        raise ValueError(f"phi2_dim must have {nsp} entries, got {len(phi2_dim)}")
    for i in range(nsp):
        cfg.n_dim[i] = n_dim[i]
        cfg.phi2_dim[i] = phi2_dim[i]
    cfg.L = int(L)
    cfg.triangulation = Triangulation.parse(triangulation)
    cfg.is_masked = 1 if is_masked else 0
    cfg.num_masked = int(num_masked)
    cfg.n_coils = int(n_coils)
    cfg.eps = float(eps)
    cfg.gpu_device_id = int(gpu_device_id)
    cfg.mem_fraction = float(mem_fraction)
    cfg.verbose = 1 if verbose else 0
    cfg.mask_radius = int(mask_radius)
    cfg.stream = _stream_ptr(stream)
    probe = MRConfig.from_buffer_copy(cfg)
    if probe.is_masked and probe.num_masked == 0:
        probe.is_masked = 0
    check(lib.mr_config_validate(ctypes.byref(probe)), "mr_config_validate")
    return cfg


def num_tri(L: int) -> int:
    return int(lib.mr_num_tri(int(L)))


def phi2_toep_shape(cfg: MRConfig) -> tuple:
    """Row-major kernel shape: (T, [Mz,] My, Mx), or (T, num_masked)."""
    T = num_tri(cfg.L)
    if cfg.is_masked:
        return (T, int(cfg.num_masked))
    return (T,) + tuple(int(cfg.phi2_dim[i]) for i in reversed(range(cfg.n_spatial)))


def _img_shape(cfg: MRConfig) -> tuple:
    return tuple(int(cfg.n_dim[i]) for i in reversed(range(cfg.n_spatial)))


def _check_traj(cfg: MRConfig, traj):
    ndim = 3 if cfg.traj_type == TrajType.THREE_D else 2
    if traj is None or traj.dim() != 3 or traj.shape[0] != ndim:
        raise ValueError(f"traj must be shaped (ndim, Nkx, Ntrajs) with ndim = {ndim}, "
                         f"got {None if traj is None else tuple(traj.shape)}")
    return int(traj.shape[2]), int(traj.shape[1])


def _check_gram(cfg: MRConfig, phi2, traj, mask):
    Ntrajs, Nkx = _check_traj(cfg, traj)
    nkz = cfg.phi2_dim[2] if cfg.traj_type == TrajType.STACK else 1
    _check_size(phi2, cfg.L * cfg.L * nkz * Ntrajs, "phi2")
    if cfg.is_masked:
        if mask is None:
            raise ValueError("cfg.is_masked is set, so a mask is required")
        if cfg.num_masked == 0:
            cfg.num_masked = int(mask.numel())
            check(lib.mr_config_validate(ctypes.byref(cfg)), "mr_config_validate")
        elif int(mask.numel()) != int(cfg.num_masked):
            raise ValueError(f"mask has {mask.numel()} elements but cfg.num_masked is "
                             f"{cfg.num_masked}")
    return Ntrajs, Nkx


def calc_mask(cfg: MRConfig, traj, radius: Optional[int] = None):
    """Neighbourhood mask of the trajectory (reference: api.py:461-505)."""
    c = MRConfig.from_buffer_copy(cfg)
    if radius is not None:
        c.mask_radius = int(radius)
    Ntrajs, Nkx = _check_traj(c, traj)
    n = ctypes.c_int64(0)
    pad = int(lib.mr_pad_size(ctypes.byref(c)))
    buf = torch.empty(pad, dtype=torch.int64, device=traj.device)
    check(lib.mr_calc_mask(ctypes.byref(c), _ptr(traj, "traj", "float32"), Ntrajs, Nkx,
                           _ptr(buf, "mask", "int64"), ctypes.byref(n)), "mr_calc_mask")
    return buf[: n.value].clone()


def calc_toep(cfg: MRConfig, phi2, traj, mask=None, out=None):
    """Toeplitz kernel ``phi2_toep`` (reference: api.py:556-639)."""
    Ntrajs, Nkx = _check_gram(cfg, phi2, traj, mask)
    shape = phi2_toep_shape(cfg)
    expected = int(lib.mr_phi2_toep_size(ctypes.byref(cfg)))
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=phi2.device)
    else:
        _check_size(out, expected, "out")
    check(lib.mr_calc_toep(ctypes.byref(cfg), _ptr(phi2, "phi2", "complex64"),
                           _ptr(traj, "traj", "float32"), ctypes.c_int64(Ntrajs),
                           ctypes.c_int64(Nkx), _ptr(mask, "mask", "int64", allow_none=True),
                           _ptr(out, "out", "complex64")), "mr_calc_toep")
    return out


def calc_toep_from_psf(cfg: MRConfig, psf, t0: int, phi2_toep, mask=None):
    """Finish components [t0, t0 + n_comp) of ``phi2_toep`` from centered
    point-spread functions ``psf`` (n_comp, [Mz,] My, Mx)."""
    grid = 1
    for i in range(cfg.n_spatial):
        grid *= int(cfg.phi2_dim[i])
    n_comp = int(psf.numel()) // grid
    check(lib.mr_calc_toep_from_psf(ctypes.byref(cfg), _ptr(psf, "psf", "complex64"),
                                    ctypes.c_int64(int(t0)), ctypes.c_int64(n_comp),
                                    _ptr(mask, "mask", "int64", allow_none=True),
                                    _ptr(phi2_toep, "phi2_toep", "complex64")),
          "mr_calc_toep_from_psf")
    return phi2_toep


def aha_toeplitz(cfg: MRConfig, x, SEs, phi2_toep, mask=None, out=None):
    """One-shot Toeplitz A^H A (reference: api.py:642-701)."""
    _check_size(x, int(lib.mr_x_size(ctypes.byref(cfg))), "x")
    _check_size(SEs, int(lib.mr_se_size(ctypes.byref(cfg))), "SEs")
    _check_size(phi2_toep, int(lib.mr_phi2_toep_size(ctypes.byref(cfg))), "phi2_toep")
    if cfg.is_masked and mask is None:
        raise ValueError("cfg.is_masked is set, so a mask is required")
    if out is None:
        out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
    else:
        _check_size(out, x.numel(), "out")
    check(lib.mr_aha_toeplitz(ctypes.byref(cfg), _ptr(x, "x", "complex64"),
                              _ptr(SEs, "SEs", "complex64"),
                              _ptr(phi2_toep, "phi2_toep", "complex64"),
                              _ptr(mask, "mask", "int64", allow_none=True),
                              _ptr(out, "out", "complex64")), "mr_aha_toeplitz")
    return out


def _devices_arg(devices):
    if devices is None:
        return None, 0
    if isinstance(devices, str):
        if devices.lower() != "all":
            raise ValueError(f"devices must be None, 'all', an int or a list; got {devices!r}")
        return None, OP_ALL_DEVICES
    if isinstance(devices, numbers.Integral):
        devices = [devices]
    devices = [int(d) for d in devices]
    if not devices:
        return None, 0
    return (ctypes.c_int * len(devices))(*devices), len(devices)


_NUFFT_OPTS = ("upsampfac", "gpu_method")


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
        self = cls.__new__(cls)
        self._handle = ctypes.c_void_p()
        self._keep = ()
        self._method = OpMethod._TO_STR[OpMethod.parse(method)]
        is_nufft = self._method == "nufft"
        if is_nufft and cfg.is_masked:
            raise ValueError("method='nufft' has no kernel to mask; build the configuration "
                             "with is_masked=False")
        nufft_opts = dict(nufft_opts or {})
        if not is_nufft and nufft_opts:
            raise ValueError("nufft_opts is only used with method='nufft'")
        unknown = set(nufft_opts) - set(_NUFFT_OPTS)
        if unknown:
            raise ValueError(f"unknown nufft_opts {sorted(unknown)}; accepted: {list(_NUFFT_OPTS)}")
        _check_size(SEs, int(lib.mr_se_size(ctypes.byref(cfg))), "SEs")
        if _phi2_toep is not None and phi2 is None and traj is None and not is_nufft:
            Ntrajs, Nkx = 0, 0
            if cfg.is_masked:
                if mask is None:
                    raise ValueError("cfg.is_masked is set, so a mask is required")
                if cfg.num_masked == 0:
                    cfg.num_masked = int(mask.numel())
        else:
            if phi2 is None or traj is None:
                raise ValueError("phi2 and traj are required (the operator is built from the "
                                 "Gram blocks)")
            Ntrajs, Nkx = _check_gram(cfg, phi2, traj, mask)
        if _phi2_toep is not None:
            if is_nufft:
                raise ValueError("_phi2_toep is only used with method='toeplitz'")
            _check_size(_phi2_toep, int(lib.mr_phi2_toep_size(ctypes.byref(cfg))), "_phi2_toep")
        devs, ndev = _devices_arg(devices)
        self._cfg = MRConfig.from_buffer_copy(cfg)
        opt = MROpOptions()
        check(lib.mr_op_options_defaults(ctypes.byref(opt)), "mr_op_options_defaults")
        opt.method = OpMethod.parse(method)
        opt.devices = ctypes.cast(devs, ctypes.POINTER(ctypes.c_int)) if devs else None
        opt.n_devices = ndev
        opt.toep_dist = ToepDist.parse(toep_dist)
        opt.flags = OP_SHARE_PRIMARY if share_primary else 0
        opt.nufft_upsampfac = float(nufft_opts.get("upsampfac", 0.0))
        opt.nufft_gpu_method = int(nufft_opts.get("gpu_method", 0))
        inp = MROpInputs(
            _ptr(SEs, "SEs", "complex64"),
            _ptr(phi2, "phi2", "complex64", allow_none=_phi2_toep is not None),
            _ptr(traj, "traj", "float32", allow_none=_phi2_toep is not None),
            int(Ntrajs), int(Nkx),
            _ptr(mask, "mask", "int64", allow_none=True),
            _ptr(_phi2_toep, "_phi2_toep", "complex64", allow_none=True),
        )
        if share_primary:
            self._keep = (SEs, mask, _phi2_toep)
        self._device = SEs.device
        check(lib.mr_op_create(ctypes.byref(cfg), ctypes.byref(inp), ctypes.byref(opt),
                               ctypes.byref(self._handle)), "mr_op_create")
        return self

    @property
    def config(self) -> MRConfig:
        return self._cfg

    @property
    def method(self) -> str:
        return self._method

    def _check_open(self):
        if not self._handle:
            raise RuntimeError("this AhAOp has already been closed")

    def apply(self, x, out=None):
        self._check_open()
        _check_size(x, int(lib.mr_x_size(ctypes.byref(self._cfg))), "x")
        if out is not None:
            _check_size(out, x.numel(), "out")
        else:
            out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
        check(lib.mr_op_execute(self._handle, _ptr(x, "x", "complex64"),
                                _ptr(out, "out", "complex64")), "mr_op_execute")
        return out

    __call__ = apply

    @property
    def info(self) -> dict:
        self._check_open()
        inf = MROpInfo()
        check(lib.mr_op_info(self._handle, ctypes.byref(inf)), "mr_op_info")
        n = inf.n_devices
        return dict(method=OpMethod._TO_STR.get(int(inf.method), "unknown"),
                    devices=list(inf.devices[:n]), coil_begin=list(inf.coil_begin[:n]),
                    coil_count=list(inf.coil_count[:n]), group=list(inf.group[:n]),
                    n_split_coils=int(inf.n_split_coils), peer_access=bool(inf.peer_access),
                    peer_bandwidth_gbs=float(inf.peer_bandwidth_gbs),
                    toep_source=OP_TOEP_SOURCES.get(int(inf.toep_source), "unknown"),
                    toep_devices=int(inf.toep_devices), toep_ms=float(inf.toep_ms),
                    toep_component_ms=float(inf.toep_component_ms),
                    toep_gather_gbs=float(inf.toep_gather_gbs), create_ms=float(inf.create_ms))

    def kernel(self, k: int = 0):
        """Read-only copy of the kernel the operator holds (as a tensor).

        The reference returns a view of operator memory; torch cannot wrap a
        foreign device pointer without a copy, so a device-to-device copy is
        returned instead (same values, shape of phi2_toep_shape).
        """
        self._check_open()
        ptr = ctypes.c_void_p()
        check(lib.mr_op_kernel(self._handle, int(k), ctypes.byref(ptr)), "mr_op_kernel")
        shape = phi2_toep_shape(self._cfg)
        out = torch.empty(shape, dtype=torch.complex64, device=self._device)
        nbytes = out.numel() * 8
        _memcpy_d2d(out.data_ptr(), ptr.value, nbytes)
        return out

    def set_kernel_format(self, fmt: str) -> None:
        """Kernel storage read by the applies: "fp32", "bf16" or "fp16"
        (library entry mr_op_set_kernel_format; 16-bit entries with one FP32
        scale per 4x4 block, for the shapes with a compiled fused x pass).
        No counterpart in the reference."""
        self._check_open()
        code = {"fp32": KERNEL_FP32, "bf16": KERNEL_BF16, "fp16": KERNEL_FP16}.get(fmt)
        if code is None:
            raise ValueError(f"kernel format {fmt!r}: expected 'fp32', 'bf16' or 'fp16'")
        check(lib.mr_op_set_kernel_format(self._handle, code), "mr_op_set_kernel_format")
        self._kfmt = fmt

    @property
    def kernel_format(self) -> str:
        return getattr(self, "_kfmt", "fp32")

    def autotune(self, x=None, reps: int = 2) -> dict:
        """Time the tile widths of the apply's FFT passes and the fused x
        pass's two-coil form on this operator and keep the fastest (library
        entry mr_op_autotune; results do not depend on the choice). ``x`` is
        the image the trials apply to (default: a seeded complex Gaussian of
        the operator's shape). Returns the tiling now in use (`tiles`). No
        counterpart in the reference."""
        self._check_open()
        nx = int(lib.mr_x_size(ctypes.byref(self._cfg)))
        if x is None:
            g = torch.Generator(device="cpu").manual_seed(1234)
            x = torch.complex(torch.randn(nx, generator=g), torch.randn(nx, generator=g))
            x = x.to(self._device)
        _check_size(x, nx, "x")
        out = torch.empty(x.shape, dtype=torch.complex64, device=x.device)
        torch.cuda.synchronize(self._device)   # x written on torch's stream is complete
        check(lib.mr_op_autotune(self._handle, _ptr(x, "x", "complex64"),
                                 _ptr(out, "out", "complex64"), int(reps)), "mr_op_autotune")
        return self.tiles

    @property
    def tiles(self) -> dict:
        """Tiling in use (library entry mr_op_tiles): pencils per workgroup
        requested for the outer-axis, y, inverse y and inverse outer-axis FFT
        passes (0: the built-in rule) and whether the fused x pass runs
        two-coil workgroups."""
        self._check_open()
        t = (ctypes.c_int * 5)()
        check(lib.mr_op_tiles(self._handle, t, 5), "mr_op_tiles")
        return dict(fo=t[0], fy=t[1], iy=t[2], io=t[3], wide=bool(t[4]))

    def cg(self, b, x=None, *, lam: float = 0.0, iters: int = 40, refresh: int = 10,
           op_refresh: Optional["AhAOp"] = None, graph: bool = False):
        """Conjugate gradients on (A^H A + lam I) x = b (library entry mr_op_cg).

        FP32 vectors, FP64 inner products and step lengths; every ``refresh``
        iterations the residual is recomputed as b - (A^H A + lam I) x with
        ``op_refresh`` (default: this operator). ``x`` is the initial guess
        (zeros when None) and is overwritten. ``graph`` replays HIP graphs of
        the iterations (needs a configuration built with a non-default stream).
        Returns ``(x, hist)`` with hist[k] = ||r_k|| / ||b||, k = 0..iters.
        No counterpart in the reference, whose callers run CG in Python.
        """
        self._check_open()
        n = int(lib.mr_x_size(ctypes.byref(self._cfg)))
        _check_size(b, n, "b")
        flags = 0
        if x is None:
            x = torch.empty(b.shape, dtype=torch.complex64, device=b.device)
            flags |= CG_ZERO_INIT
        else:
            _check_size(x, n, "x")
        if graph:
            flags |= CG_GRAPH
        hist = (ctypes.c_double * (int(iters) + 1))()
        ref = op_refresh._handle if op_refresh is not None else None
        if op_refresh is not None:
            op_refresh._check_open()
        check(lib.mr_op_cg(self._handle, ref, _ptr(b, "b", "complex64"), _ptr(x, "x", "complex64"),
                           float(lam), int(iters), int(refresh), flags, hist), "mr_op_cg")
        return x, list(hist)

    def close(self) -> None:
        if self._handle:
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
