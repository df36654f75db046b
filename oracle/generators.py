"""Seeded synthetic inputs (oracle side).

Test infrastructure: see ``oracle/__init__.py``.

Restates the reference's generators so that the same seeds give the same
inputs: golden-angle radial / golden-means kooshball trajectories and PSD
Gram blocks of tests/make_reference.py:160-236, the uniform-random problems
of tests/check_op.py:82-172, plus the benchmark workloads (radial and
stack-of-stars with an orthonormal temporal basis, Gaussian-bump coil maps).
All outputs use the binding layout of ``oracle.conventions``.
"""

from __future__ import annotations

import numpy as np

from .conventions import Geometry, make_geometry

GOLDEN = (np.sqrt(5.0) - 1.0) / 2.0


def radial_readout(n_kx: int) -> np.ndarray:
    """pi * (2 s / n_kx - 1), s = 0..n_kx-1 (tests/make_reference.py:160-170)."""
    return np.pi * (2.0 * np.arange(n_kx) / n_kx - 1.0)


def make_traj(traj_type: str, n_traj: int, n_kx: int) -> np.ndarray:
    """Golden-angle radial (2D/stack) or golden-means kooshball (3D) spokes.

    tests/make_reference.py:172-201. Returns float32 (ndim, Nkx, Ntrajs).
    """
    r = radial_readout(n_kx)[:, None]
    t = np.arange(n_traj)[None, :]
    if traj_type in ("2D", "stack"):
        th = np.pi * GOLDEN * t
        traj = np.stack([r * np.cos(th), r * np.sin(th)])
    else:
        phi1, phi2 = 0.4656, 0.6823
        z = np.mod(t * phi1, 1.0)
        az = 2.0 * np.pi * np.mod(t * phi2, 1.0)
        s = np.sqrt(1.0 - z ** 2)
        traj = np.stack([r * s * np.cos(az), r * s * np.sin(az), r * z])
    return np.ascontiguousarray(traj, dtype=np.float32)


def make_phi2(rng, L: int, n_traj: int, n_kz: int | None) -> np.ndarray:
    """Random PSD Gram blocks A^T conj(A) / 2 of a two-frame basis.

    tests/make_reference.py:204-231. Returns complex64 (L, L, [Nkz,] Ntrajs)
    with phi2[j, i, ...] = B(i, j).
    """
    lead = (n_traj,) if n_kz is None else (n_traj, n_kz)
    A = rng.standard_normal(lead + (2, L)) + 1j * rng.standard_normal(lead + (2, L))
    B = np.einsum("...ki,...kj->...ij", A, A.conj()) / 2.0
    phi2 = B.transpose(tuple(range(B.ndim))[::-1])
    return np.ascontiguousarray(phi2, dtype=np.complex64)


def crandn(rng, shape) -> np.ndarray:
    """Standard complex Gaussian complex64 (tests/make_reference.py:234-236)."""
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


#: tests/make_reference.py:130-146 — (name, traj_type, n_dim, L, n_coils,
#: n_traj, n_kx, seed, phi2_dim)
REFERENCE_CASES = [
    ("2d_50x40_L3_c3", "2D", (50, 40), 3, 3, 61, 100, 1, None),
    ("2d_33x27_L1_c2_odd", "2D", (33, 27), 1, 2, 45, 66, 2, None),
    ("2d_50x40_L4_c2_M101x81", "2D", (50, 40), 4, 2, 61, 100, 3, (101, 81)),
    ("3d_24x20x16_L4_c2", "3D", (24, 20, 16), 4, 2, 400, 48, 4, None),
    ("3d_25x21x15_L3_c2_odd", "3D", (25, 21, 15), 3, 2, 350, 50, 5, None),
    ("stack_40x36x10_L3_c3", "stack", (40, 36, 10), 3, 3, 40, 80, 6, None),
    ("stack_31x36x9_L4_c2_odd", "stack", (31, 36, 9), 4, 2, 35, 62, 7, None),
]


def reference_case(name: str):
    """Inputs of one tests/make_reference.py case, in its RNG draw order.

    tests/make_reference.py:383-396: traj (deterministic), then phi2, x, SEs
    from ``default_rng(seed)``. Returns (geometry, traj, phi2, x, SEs).
    """
    for case in REFERENCE_CASES:
        if case[0] == name:
            _, tt, n, L, nc, ntr, nkx, seed, pdim = case
            break
    else:
        raise KeyError(name)
    rng = np.random.default_rng(seed)
    g = make_geometry(tt, n, L, nc, pdim)
    traj = make_traj(tt, ntr, nkx)
    phi2 = make_phi2(rng, L, ntr, n[2] if tt == "stack" else None)
    x = crandn(rng, (L,) + g.img_shape())
    SEs = crandn(rng, (nc,) + g.img_shape())
    return g, traj, phi2, x, SEs


#: tests/check_op.py:82-89 — name -> (traj_type, n_dim, L, n_coils, Ntrajs, Nkx)
CHECK_OP_PROBLEMS = {
    "2d": ("2D", (40, 36), 5, 3, 60, 48),
    "2d_odd_L33": ("2D", (33, 27), 33, 5, 40, 40),
    "stack": ("stack", (32, 30, 8), 4, 3, 40, 40),
    "stack_odd": ("stack", (31, 36, 9), 3, 2, 30, 40),
    "3d": ("3D", (20, 18, 16), 3, 4, 300, 24),
    "3d_odd": ("3D", (21, 19, 15), 4, 3, 300, 24),
}


def check_op_problem(name: str, rng):
    """Uniform-random problem of tests/check_op.py:125-172.

    Trajectory uniform in [-pi, pi), Gram blocks P P^H / 3 of a random L x 3
    basis per readout (and kz), random x and maps. Returns
    (geometry, traj, phi2, x, SEs).
    """
    tt, n, L, C, Ntr, Nkx = CHECK_OP_PROBLEMS[name]
    nkz = n[2] if tt == "stack" else 1
    ndim = 3 if tt == "3D" else 2
    traj = rng.uniform(-np.pi, np.pi, (ndim, Nkx, Ntr)).astype(np.float32)
    phi2 = np.empty((L, L, nkz, Ntr), np.complex64)
    for t in range(Ntr):
        for z in range(nkz):
            P = rng.standard_normal((L, 3)) + 1j * rng.standard_normal((L, 3))
            phi2[:, :, z, t] = (P @ P.conj().T).T / 3
    if tt != "stack":
        phi2 = phi2[:, :, 0, :]
    g = make_geometry(tt, n, L, C)
    img = g.img_shape()
    x = rng.standard_normal((L,) + img) + 1j * rng.standard_normal((L,) + img)
    S = rng.standard_normal((C,) + img) + 1j * rng.standard_normal((C,) + img)
    return (g, traj, np.ascontiguousarray(phi2), x.astype(np.complex64),
            S.astype(np.complex64))


# --------------------------------------------------------------------------
# Benchmark workloads
# --------------------------------------------------------------------------

def orthonormal_basis(rng, n_frames: int, L: int) -> np.ndarray:
    """Rank x time basis Phi (L, n_frames): Q^T of the QR of a complex Gaussian.

    Rows of Phi are orthonormal in time (Phi Phi^H = I_L over all frames).
    """
    G = rng.standard_normal((n_frames, L)) + 1j * rng.standard_normal((n_frames, L))
    Q, _ = np.linalg.qr(G)
    return Q.T.copy()  # (L, n_frames)


def rank1_phi2(Phi: np.ndarray, frames: np.ndarray) -> np.ndarray:
    """Gram blocks Phi(:,t) Phi(:,t)^H of the frame of every readout.

    ``frames`` has shape (Ntrajs,) or (Nkz, Ntrajs). Returns the binding
    layout (L, L, [Nkz,] Ntrajs) with phi2[j, i, ...] = Phi(i,t) conj(Phi(j,t)).
    """
    p = Phi[:, frames]  # (L, ...) with ... = frames.shape
    B = p[:, None] * np.conj(p[None, :])  # B[i, j, ...] = Phi_i conj(Phi_j)
    return np.ascontiguousarray(np.swapaxes(B, 0, 1), dtype=np.complex64)


def gaussian_coil_maps(g: Geometry, n_coils: int, rng=None) -> np.ndarray:
    """Smooth complex coil maps: Gaussian bumps centred on a circle (sphere
    in 3D) around the field of view, each with a gentle linear phase,
    normalised to unit root-sum-of-squares. Returns complex64 (C, *img)."""
    img = g.img_shape()
    axes = [np.arange(s) - s // 2 for s in img]
    grids = np.meshgrid(*axes, indexing="ij")  # row-major axes ([z,] y, x)
    if g.n_spatial == 2:
        y, x = grids
        z = np.zeros_like(x)
    else:
        z, y, x = grids
    nx, ny, nz = g.n
    maps = []
    for c in range(n_coils):
        if g.traj_type == "3D":
            # Fibonacci points on the unit sphere.
            zz = 1.0 - 2.0 * (c + 0.5) / n_coils
            rr = np.sqrt(max(0.0, 1.0 - zz * zz))
            ang = np.pi * (3.0 - np.sqrt(5.0)) * c
            cx, cy, cz = 0.45 * nx * rr * np.cos(ang), 0.45 * ny * rr * np.sin(ang), 0.45 * nz * zz
        else:
            ang = 2.0 * np.pi * c / n_coils
            cx, cy, cz = 0.45 * nx * np.cos(ang), 0.45 * ny * np.sin(ang), 0.0
        sx, sy, sz = 0.45 * nx, 0.45 * ny, 0.6 * max(nz, 1)
        mag = np.exp(-0.5 * (((x - cx) / sx) ** 2 + ((y - cy) / sy) ** 2 + ((z - cz) / sz) ** 2))
        ph = 0.6 * np.pi * (np.cos(0.7 * c) * x / nx + np.sin(1.3 * c) * y / ny) + 0.4 * c
        maps.append(mag * np.exp(1j * ph))
    S = np.stack(maps)
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=0, keepdims=True))
    return S.astype(np.complex64)


def radial_workload(n=128, L=4, C=4, n_spokes=64, n_kx=256, n_frames=64, seed=20261202):
    """Plain 2D radial A^H A: golden-angle spokes, frame (spoke mod n_frames),
    rank-1 Gram blocks of an orthonormal basis, Gaussian maps, Gaussian x.

    Default sizes: 128x128 image, L=4, C=4, 64 spokes x 256 samples, 64 frames.
    Returns (geometry, traj, phi2, x, SEs, Phi).
    """
    rng = np.random.default_rng(seed)
    g = make_geometry("2D", (n, n), L, C)
    traj = make_traj("2D", n_spokes, n_kx)
    Phi = orthonormal_basis(rng, n_frames, L)
    frames = np.arange(n_spokes) % n_frames
    phi2 = rank1_phi2(Phi, frames)
    SEs = gaussian_coil_maps(g, C)
    x = crandn(rng, (L,) + g.img_shape())
    return g, traj, phi2, x, SEs, Phi


def stack_workload(n=320, nz=56, L=16, C=12, n_spokes=400, n_kx=640, n_frames=200,
                   seed=20261203, with_x=True):
    """Stack-of-stars A^H A: the same golden-angle spokes on every kz plane,
    frame (spoke + n_spokes * kz) mod n_frames, rank-1 Gram blocks.

    Default sizes: 320x320x56, in-plane grid 640x640, L=16, C=12,
    400 spokes x 640 samples, 200 frames. Returns
    (geometry, traj, phi2, x, SEs, Phi); x is None when ``with_x`` is False.
    """
    rng = np.random.default_rng(seed)
    g = make_geometry("stack", (n, n, nz), L, C)
    traj = make_traj("stack", n_spokes, n_kx)
    Phi = orthonormal_basis(rng, n_frames, L)
    frames = (np.arange(n_spokes)[None, :] + n_spokes * np.arange(nz)[:, None]) % n_frames
    phi2 = rank1_phi2(Phi, frames)  # (L, L, Nkz, Ntrajs)
    SEs = gaussian_coil_maps(g, C)
    x = crandn(rng, (L,) + g.img_shape()) if with_x else None
    return g, traj, phi2, x, SEs, Phi


def fibonacci_kooshball(n_spokes: int, n_kx: int) -> np.ndarray:
    """3D radial spokes along Fibonacci-sphere directions, float32 (3, Nkx, Ntrajs)."""
    r = radial_readout(n_kx)[:, None]
    s = np.arange(n_spokes)[None, :]
    zdir = 1.0 - (2.0 * s + 1.0) / (2.0 * n_spokes)  # upper half-sphere, full lines
    rho = np.sqrt(np.maximum(0.0, 1.0 - zdir ** 2))
    az = np.pi * (3.0 - np.sqrt(5.0)) * s
    traj = np.stack([r * rho * np.cos(az), r * rho * np.sin(az), r * zdir])
    return np.ascontiguousarray(traj, dtype=np.float32)


def kooshball_workload(n=160, L=8, C=8, n_spokes=20000, n_kx=320, n_frames=100,
                       seed=20261204, with_x=True):
    """3D radial A^H A on Fibonacci-sphere spokes (default 160^3, L=8, C=8)."""
    rng = np.random.default_rng(seed)
    g = make_geometry("3D", (n, n, n), L, C)
    traj = fibonacci_kooshball(n_spokes, n_kx)
    Phi = orthonormal_basis(rng, n_frames, L)
    phi2 = rank1_phi2(Phi, np.arange(n_spokes) % n_frames)
    SEs = gaussian_coil_maps(g, C)
    x = crandn(rng, (L,) + g.img_shape()) if with_x else None
    return g, traj, phi2, x, SEs, Phi
