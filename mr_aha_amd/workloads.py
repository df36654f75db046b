"""Synthetic benchmark workloads.

Shapes of the benchmark configurations: golden-angle radial spokes (2D and
stack-of-stars, the same spokes on every kz plane) or Fibonacci-sphere
kooshball spokes (3D); a rank x time basis Phi from the QR of a seeded
complex Gaussian matrix; every readout (and kz plane) acquired at one frame
with the rank-1 Gram block Phi(:,t) Phi(:,t)^H; smooth Gaussian-bump coil
maps normalised to unit root-sum-of-squares; complex Gaussian
coefficients x. All arrays use the binding layout of ``mr_aha_amd.api``.

The random draws (the basis, then x) come from numpy's default_rng(seed) in
a fixed order, so a workload equals the committed golden fixtures built
from the same recipe and seed; the trajectories and coil maps are closed
forms evaluated in float64 on the device. The tests check both against the
fixtures' generator.
"""

from __future__ import annotations

import math

import numpy as np
import torch

GOLDEN = (math.sqrt(5.0) - 1.0) / 2.0


def _readout(nkx, device):
    return math.pi * (2.0 * torch.arange(nkx, device=device, dtype=torch.float64) / nkx - 1.0)


def radial_traj(n_spokes, nkx, device):
    r = _readout(nkx, device)[:, None]
    th = math.pi * GOLDEN * torch.arange(n_spokes, device=device, dtype=torch.float64)[None, :]
    return torch.stack([r * torch.cos(th), r * torch.sin(th)]).to(torch.float32).contiguous()


def kooshball_traj(n_spokes, nkx, device):
    r = _readout(nkx, device)[:, None]
    s = torch.arange(n_spokes, device=device, dtype=torch.float64)[None, :]
    zdir = 1.0 - (2.0 * s + 1.0) / (2.0 * n_spokes)
    rho = torch.sqrt(torch.clamp(1.0 - zdir ** 2, min=0.0))
    az = math.pi * (3.0 - math.sqrt(5.0)) * s
    return torch.stack([r * rho * torch.cos(az), r * rho * torch.sin(az), r * zdir]) \
        .to(torch.float32).contiguous()


def basis_and_x(n_frames, L, seed, x_shape=None):
    """Phi (L, n_frames): Q^T of the QR of a complex Gaussian (n_frames, L)
    matrix (orthonormal rows), then x of ``x_shape`` (complex64), both from
    numpy's default_rng(seed) in that order."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n_frames, L)) + 1j * rng.standard_normal((n_frames, L))
    Q, _ = np.linalg.qr(G)
    Phi = torch.from_numpy(Q.T.copy())
    x = None
    if x_shape is not None:
        x = (rng.standard_normal(x_shape) + 1j * rng.standard_normal(x_shape)).astype(np.complex64)
        x = torch.from_numpy(x)
    return Phi, x


def rank1_phi2(Phi, frames):
    """(L, L, *frames.shape) with phi2[j, i, ...] = Phi(i,t) conj(Phi(j,t))."""
    p = Phi[:, frames]
    B = p[:, None] * torch.conj(p[None, :])
    return B.transpose(0, 1).to(torch.complex64).contiguous()


def coil_maps(img_shape, n_coils, three_d, device):
    axes = [torch.arange(s, device=device, dtype=torch.float64) - s // 2 for s in img_shape]
    grids = torch.meshgrid(*axes, indexing="ij")
    if len(img_shape) == 2:
        y, x = grids
        z = torch.zeros_like(x)
        nz = 1
        ny, nx = img_shape
    else:
        z, y, x = grids
        nz, ny, nx = img_shape
    maps = []
    for c in range(n_coils):
        if three_d:
            zz = 1.0 - 2.0 * (c + 0.5) / n_coils
            rr = math.sqrt(max(0.0, 1.0 - zz * zz))
            ang = math.pi * (3.0 - math.sqrt(5.0)) * c
            cx, cy, cz = 0.45 * nx * rr * math.cos(ang), 0.45 * ny * rr * math.sin(ang), 0.45 * nz * zz
        else:
            ang = 2.0 * math.pi * c / n_coils
            cx, cy, cz = 0.45 * nx * math.cos(ang), 0.45 * ny * math.sin(ang), 0.0
        sx, sy, sz = 0.45 * nx, 0.45 * ny, 0.6 * max(nz, 1)
        mag = torch.exp(-0.5 * (((x - cx) / sx) ** 2 + ((y - cy) / sy) ** 2 + ((z - cz) / sz) ** 2))
        ph = 0.6 * math.pi * (math.cos(0.7 * c) * x / nx + math.sin(1.3 * c) * y / ny) + 0.4 * c
        maps.append(torch.polar(mag, ph))
    S = torch.stack(maps)
    S = S / torch.sqrt(torch.sum(S.abs() ** 2, dim=0, keepdim=True))
    return S.to(torch.complex64).contiguous()


def complex_gaussian(shape, seed, device):
    g = torch.Generator(device=device).manual_seed(seed)
    re = torch.randn(shape, generator=g, device=device, dtype=torch.float32)
    im = torch.randn(shape, generator=g, device=device, dtype=torch.float32)
    return torch.complex(re, im).contiguous()


def stack_of_stars(n=320, nz=56, L=16, C=12, n_spokes=400, nkx=640, n_frames=200,
                   seed=20261203, device="cuda", with_x=True):
    """320x320x56 image, in-plane grid 640x640, L=16, C=12, 400 spokes x 640
    samples on every kz plane, frame (spoke + 400 kz) mod 200 by default."""
    traj = radial_traj(n_spokes, nkx, device)
    Phi, x = basis_and_x(n_frames, L, seed, (L, nz, n, n) if with_x else None)
    Phi = Phi.to(device)
    frames = (torch.arange(n_spokes, device=device)[None, :] +
              n_spokes * torch.arange(nz, device=device)[:, None]) % n_frames
    phi2 = rank1_phi2(Phi, frames)
    S = coil_maps((nz, n, n), C, False, device)
    return dict(traj_type="stack", n_dim=(n, n, nz), L=L, n_coils=C, traj=traj, phi2=phi2,
                SEs=S, x=x.to(device) if with_x else None)


def radial_2d(n=128, L=4, C=4, n_spokes=64, nkx=256, n_frames=64, seed=20261202, device="cuda",
              with_x=True):
    """128x128 image, L=4, C=4, 64 golden-angle spokes x 256 samples by default."""
    traj = radial_traj(n_spokes, nkx, device)
    Phi, x = basis_and_x(n_frames, L, seed, (L, n, n) if with_x else None)
    phi2 = rank1_phi2(Phi.to(device), torch.arange(n_spokes, device=device) % n_frames)
    S = coil_maps((n, n), C, False, device)
    return dict(traj_type="2D", n_dim=(n, n), L=L, n_coils=C, traj=traj, phi2=phi2, SEs=S,
                x=x.to(device) if with_x else None)


def kooshball(n=160, L=8, C=8, n_spokes=20000, nkx=320, n_frames=100, seed=20261204,
              device="cuda", with_x=True):
    """160^3 image, grid 320^3, L=8, C=8, 20000 Fibonacci-sphere spokes x 320."""
    traj = kooshball_traj(n_spokes, nkx, device)
    Phi, x = basis_and_x(n_frames, L, seed, (L, n, n, n) if with_x else None)
    phi2 = rank1_phi2(Phi.to(device), torch.arange(n_spokes, device=device) % n_frames)
    S = coil_maps((n, n, n), C, True, device)
    return dict(traj_type="3D", n_dim=(n, n, n), L=L, n_coils=C, traj=traj, phi2=phi2, SEs=S,
                x=x.to(device) if with_x else None)
