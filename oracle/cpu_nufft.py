"""CPU NUFFT normal operator: the reference's CPU path (method="nufft").

Test infrastructure: see ``oracle/__init__.py``. Timed as the benchmark's
CPU baseline and pinned against the direct NUDFT (``nudft.aha_nudft``).

The same A^H A as src/cuda/mr_nufft_op.cu:613-700 and
python/mr_aha/nufft.py:473-552, on the host: per coil, type-2 NUFFT (sign -1)
of S_c x, the per-readout Gram multiply k'_l = sum_j k_j B(j, l), type-1
NUFFT (sign +1), conj(S_c) and the coil sum; for a stack, the z modulation
zmod(z) = exp(2 pi i floor(Nz/2) z / Nz) and an unnormalised z-DFT pair
around per-plane 2D NUFFTs (mr_nufft_op.cu:436-449). The NUFFTs are the
FINUFFT method (Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 41(5),
2019): exponential-of-semicircle kernel of width w = ceil(log10(1/eps)) + 1,
beta = 2.30 w, upsampling 2; the kernel's Fourier transform by Gauss-Legendre
quadrature for the deconvolution. Spreading and interpolation are C++ with
OpenMP (oracle/cpu_nufft.cpp); the FFTs are pocketfft (scipy.fft) over the
same threads. complex64 at eps 1e-6 is the timed baseline; complex128 at a
small eps is cross-checked against the exact NUDFT.
"""

from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

import numpy as np
import scipy.fft as sfft

from .conventions import Geometry, gram_blocks

_HERE = Path(__file__).resolve().parent
_LIB = None


def _cpu_has(flags) -> bool:
    try:
        text = Path("/proc/cpuinfo").read_text()
    except OSError:
        return False
    line = next((ln for ln in text.splitlines() if ln.startswith("flags")), "")
    have = set(line.split(":", 1)[-1].split())
    return all(f in have for f in flags)


def lib():
    """The compiled spreader (oracle/_build), built by oracle/build_c.py."""
    global _LIB
    if _LIB is None:
        v3 = _HERE / "_build" / "libmr_cpu_nufft_v3.so"
        base = _HERE / "_build" / "libmr_cpu_nufft.so"
        if not base.exists():
            from . import build_c
            build_c.build()
        path = v3 if v3.exists() and _cpu_has(("avx2", "fma", "bmi2")) else base
        L = ctypes.CDLL(str(path))
        p = ctypes.c_void_p
        i64p = ctypes.POINTER(ctypes.c_int64)
        for name in ("mr_cpu_interp_f", "mr_cpu_spread_f", "mr_cpu_interp_d", "mr_cpu_spread_d"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, i64p, ctypes.c_int64, p, ctypes.c_int64, p, p, p,
                           ctypes.c_int, ctypes.c_double, p]
        L.mr_cpu_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def es_params(eps: float):
    w = int(math.ceil(math.log10(1.0 / eps))) + 1
    w = max(2, min(16, w))
    return w, 2.30 * w


def es_fourier(w: int, beta: float, nf: int, N: int) -> np.ndarray:
    """phi_hat(2 pi k / nf) for the centered modes k = -floor(N/2) .. ceil(N/2)-1,
    phi_hat(xi) = int_{-w/2}^{w/2} phi(z) cos(xi z) dz (Gauss-Legendre)."""
    q = 4 * w + 8
    t, wq = np.polynomial.legendre.leggauss(q)
    z = (t + 1.0) * (w / 4.0)                      # nodes on [0, w/2]
    wz = wq * (w / 4.0)
    phi = np.exp(beta * (np.sqrt(np.maximum(0.0, 1.0 - (2.0 * z / w) ** 2)) - 1.0))
    k = np.arange(N) - N // 2
    xi = 2.0 * np.pi * k / nf
    return 2.0 * (np.cos(np.outer(xi, z)) * (phi * wz)[None, :]).sum(axis=1)


class _Plan2:
    """2D NUFFT pair on one trajectory: type 2 (sign -1) and type 1 (sign +1)
    for a batch of B images (Ny, Nx), samples as (M, B)."""

    def __init__(self, N, traj_xy, eps, dtype, threads):
        self.Ny, self.Nx = N
        self.nfy, self.nfx = 2 * self.Ny, 2 * self.Nx
        self.w, self.beta = es_params(eps)
        self.dtype = np.dtype(dtype)
        self.double = self.dtype == np.complex128
        self.threads = threads
        self.px = np.ascontiguousarray(traj_xy[0].ravel(), dtype=np.float32)
        self.py = np.ascontiguousarray(traj_xy[1].ravel(), dtype=np.float32)
        self.M = self.px.size
        fy = es_fourier(self.w, self.beta, self.nfy, self.Ny)
        fx = es_fourier(self.w, self.beta, self.nfx, self.Nx)
        self.deconv = (1.0 / np.outer(fy, fx)).astype(np.float64 if self.double else np.float32)
        self.iy = np.mod(np.arange(self.Ny) - self.Ny // 2, self.nfy)
        self.ix = np.mod(np.arange(self.Nx) - self.Nx // 2, self.nfx)
        self.nf = (ctypes.c_int64 * 3)(self.nfx, self.nfy, 1)
        L = lib()
        self._interp = L.mr_cpu_interp_d if self.double else L.mr_cpu_interp_f
        self._spread = L.mr_cpu_spread_d if self.double else L.mr_cpu_spread_f

    def type2(self, im):
        """im (Ny, Nx, B) -> samples (M, B)."""
        B = im.shape[2]
        g = np.zeros((self.nfy, self.nfx, B), dtype=self.dtype)
        g[np.ix_(self.iy, self.ix)] = im * self.deconv[:, :, None]
        g = sfft.fft2(g, axes=(0, 1), workers=self.threads, overwrite_x=True)
        out = np.empty((self.M, B), dtype=self.dtype)
        rc = self._interp(2, self.nf, B, g.ctypes.data, self.M, self.px.ctypes.data,
                          self.py.ctypes.data, None, self.w, self.beta, out.ctypes.data)
        assert rc == 0
        return out

    def type1(self, k):
        """samples (M, B) -> image (Ny, Nx, B)."""
        B = k.shape[1]
        k = np.ascontiguousarray(k, dtype=self.dtype)
        g = np.zeros((self.nfy, self.nfx, B), dtype=self.dtype)
        rc = self._spread(2, self.nf, B, g.ctypes.data, self.M, self.px.ctypes.data,
                          self.py.ctypes.data, None, self.w, self.beta, k.ctypes.data)
        assert rc == 0
        g = sfft.ifft2(g, axes=(0, 1), workers=self.threads, norm="forward", overwrite_x=True)
        return g[np.ix_(self.iy, self.ix)] * self.deconv[:, :, None]


class _Plan3:
    """3D NUFFT pair for a batch of B volumes (Nz, Ny, Nx)."""

    def __init__(self, N, traj, eps, dtype, threads):
        self.N = tuple(N)
        self.nf = tuple(2 * n for n in self.N)
        self.w, self.beta = es_params(eps)
        self.dtype = np.dtype(dtype)
        self.double = self.dtype == np.complex128
        self.threads = threads
        self.p = [np.ascontiguousarray(traj[a].ravel(), dtype=np.float32) for a in range(3)]
        self.M = self.p[0].size
        f = [es_fourier(self.w, self.beta, self.nf[i], self.N[i]) for i in range(3)]
        d = 1.0 / (f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :])
        self.deconv = d.astype(np.float64 if self.double else np.float32)
        self.idx = [np.mod(np.arange(self.N[i]) - self.N[i] // 2, self.nf[i]) for i in range(3)]
        self.nfc = (ctypes.c_int64 * 3)(self.nf[2], self.nf[1], self.nf[0])
        L = lib()
        self._interp = L.mr_cpu_interp_d if self.double else L.mr_cpu_interp_f
        self._spread = L.mr_cpu_spread_d if self.double else L.mr_cpu_spread_f

    def type2(self, im):
        B = im.shape[3]
        g = np.zeros(self.nf + (B,), dtype=self.dtype)
        g[np.ix_(*self.idx)] = im * self.deconv[..., None]
        g = sfft.fftn(g, axes=(0, 1, 2), workers=self.threads, overwrite_x=True)
        out = np.empty((self.M, B), dtype=self.dtype)
        rc = self._interp(3, self.nfc, B, g.ctypes.data, self.M, self.p[0].ctypes.data,
                          self.p[1].ctypes.data, self.p[2].ctypes.data, self.w, self.beta,
                          out.ctypes.data)
        assert rc == 0
        return out

    def type1(self, k):
        B = k.shape[1]
        k = np.ascontiguousarray(k, dtype=self.dtype)
        g = np.zeros(self.nf + (B,), dtype=self.dtype)
        rc = self._spread(3, self.nfc, B, g.ctypes.data, self.M, self.p[0].ctypes.data,
                          self.p[1].ctypes.data, self.p[2].ctypes.data, self.w, self.beta,
                          k.ctypes.data)
        assert rc == 0
        g = sfft.ifftn(g, axes=(0, 1, 2), workers=self.threads, norm="forward", overwrite_x=True)
        return g[np.ix_(*self.idx)] * self.deconv[..., None]


def _gram(k, B, Nkx, Ntr):
    """k (M, L) with sample m = r Ntr + s -> k'(m) = k(m) B[s] (B: (Ntr, L, L))."""
    L = k.shape[1]
    kk = k.reshape(Nkx, Ntr, L).transpose(1, 0, 2)          # (Ntr, Nkx, L)
    return np.matmul(kk, B).transpose(1, 0, 2).reshape(Nkx * Ntr, L)


def aha_nufft(g: Geometry, x, SEs, phi2, traj, eps: float = 1e-6, dtype=np.complex64,
              threads: int | None = None, coils=None, planes=None):
    """A^H A x through NUFFTs (method="nufft"), restricted to ``coils``
    (default all). ``x`` (L, *img), ``SEs`` (C, *img), ``phi2`` and ``traj``
    in the binding layout. For a stack, ``planes`` restricts the kz planes
    whose 2D NUFFTs run (a timing sample; the result is then partial)."""
    threads = threads or os.cpu_count() or 1
    L = g.L
    img = g.img_shape()
    dt = np.dtype(dtype)
    x = np.asarray(x).astype(dt, copy=False).reshape((L,) + img)
    S = np.asarray(SEs).astype(dt, copy=False).reshape((-1,) + img)
    Bm = gram_blocks(phi2).astype(dt)                    # (Ntr, [Nkz,] L, L)
    ndim, Nkx, Ntr = traj.shape
    coils = range(S.shape[0]) if coils is None else coils
    out = np.zeros((L,) + img, dtype=dt)
    if g.traj_type == "3D":
        plan = _Plan3(img, [traj[a] for a in range(3)], eps, dt, threads)
        for c in coils:
            im = np.moveaxis(x * S[c][None], 0, -1)       # (Nz, Ny, Nx, L)
            k = _gram(plan.type2(np.ascontiguousarray(im)), Bm, Nkx, Ntr)
            out += np.moveaxis(plan.type1(k), -1, 0) * np.conj(S[c])[None]
        return out
    plan = _Plan2(img[-2:], [traj[0], traj[1]], eps, dt, threads)
    if g.traj_type == "2D":
        for c in coils:
            im = np.moveaxis(x * S[c][None], 0, -1)       # (Ny, Nx, L)
            k = _gram(plan.type2(np.ascontiguousarray(im)), Bm, Nkx, Ntr)
            out += np.moveaxis(plan.type1(k), -1, 0) * np.conj(S[c])[None]
        return out
    nz = img[0]
    zmod = np.exp(2j * np.pi * (nz // 2) * np.arange(nz) / nz).astype(dt)[:, None, None]
    planes = range(nz) if planes is None else planes
    for c in coils:
        se = S[c] * zmod
        im = sfft.fft(x * se[None], axis=1, workers=threads)          # (L, Nz(kz), Ny, Nx)
        res = np.zeros_like(im)
        for z in planes:
            p = np.ascontiguousarray(np.moveaxis(im[:, z], 0, -1))       # (Ny, Nx, L)
            k = _gram(plan.type2(p), Bm[:, z], Nkx, Ntr)
            res[:, z] = np.moveaxis(plan.type1(k), -1, 0)
        res = sfft.ifft(res, axis=1, norm="forward", workers=threads)
        out += res * np.conj(se)[None]
    return out
