"""Direct (exact) NUDFT normal operator and a dense brute-force forward model.

Test infrastructure: see ``oracle/__init__.py``.
"""

from __future__ import annotations

import numpy as np

from .conventions import Geometry, gram_blocks


def _exp_tables(g: Geometry, traj: np.ndarray):
    """e^{-i r_a x_{m,a}} per spatial NUFFT axis, row-major axis order.

    r_a = i - floor(N_a/2) are the centered image coordinates (the CMCL mode
    order, modeord = 0, of python/mr_aha/nufft.py:374-377).
    """
    nd = traj.shape[0]
    tabs = []
    for a in reversed(range(nd)):
        r = np.arange(g.n[a]) - g.n[a] // 2
        xm = traj[a].astype(np.float64).ravel()
        tabs.append(np.exp(-1j * np.outer(xm, r)))  # (Mnu, N_a)
    return tabs


def _type2(tabs, im):
    """k[m] = sum_r im[r] e^{-i r.x_m} for a batch im (B, [Nz,] Ny, Nx)."""
    if len(tabs) == 2:
        Ey, Ex = tabs
        t = np.einsum("byx,mx->bym", im, Ex)
        return np.einsum("bym,my->bm", t, Ey)
    Ez, Ey, Ex = tabs
    t = np.einsum("bzyx,mx->bzym", im, Ex)
    t = np.einsum("bzym,my->bzm", t, Ey)
    return np.einsum("bzm,mz->bm", t, Ez)


def _type1(tabs, k):
    """im[r] = sum_m k[m] e^{+i r.x_m} for a batch k (B, Mnu)."""
    if len(tabs) == 2:
        Ey, Ex = tabs
        t = np.einsum("bm,my->bmy", k, np.conj(Ey))
        return np.einsum("bmy,mx->byx", t, np.conj(Ex))
    Ez, Ey, Ex = tabs
    t = np.einsum("bm,mz->bmz", k, np.conj(Ez))
    t = np.einsum("bmz,my->bmzy", t, np.conj(Ey))
    return np.einsum("bmzy,mx->bzyx", t, np.conj(Ex))


def aha_nudft(g: Geometry, x: np.ndarray, SEs: np.ndarray, phi2: np.ndarray,
              traj: np.ndarray) -> np.ndarray:
    """Exact A^H A x = sum_c S_c^H F_+ B F_- S_c x (python/mr_aha/nufft.py:473-552).

    F_- is the type-2 NUDFT (sign -1), B multiplies the rank row vector of
    every sample by its readout's Gram block from the right
    (k'_l = sum_j k_j B(j, l), python/mr_aha/nufft.py:439-471), F_+ is the
    type-1 NUDFT (sign +1). For a stack the image is modulated by
    zmod(z) = exp(2 pi i floor(Nz/2) z / Nz) and transformed with the
    unnormalised z-DFT pair, and kz index K uses the Gram blocks of plane K
    (src/cuda/mr_nufft_op.cu:436-449, python/mr_aha/nufft.py:392-397, 531-549).

    Pinned by the dense brute-force forward model (``dense_aha``) and the
    full-Cartesian closed form (tests/test_oracle.py).
    """
    L = g.L
    img = g.img_shape()
    x = np.asarray(x, dtype=np.complex128).reshape((L,) + img)
    S = np.asarray(SEs, dtype=np.complex128).reshape((-1,) + img)
    B = gram_blocks(phi2).astype(np.complex128)  # (Ntr, [Nkz,] L, L)
    ndim, Nkx, Ntr = traj.shape
    tabs = _exp_tables(g, traj)
    out = np.zeros((L,) + img, dtype=np.complex128)
    stack = g.traj_type == "stack"
    if stack:
        nz = g.n[2]
        zmod = np.exp(2j * np.pi * (nz // 2) * np.arange(nz) / nz)[:, None, None]
    for c in range(S.shape[0]):
        se = S[c] * zmod if stack else S[c]
        im = x * se[None]
        if stack:
            im = np.fft.fft(im, axis=1)                   # (L, Nz(kz), Ny, Nx)
            k = _type2(tabs, im.reshape((L * nz,) + img[1:]))
            k = k.reshape(L, nz, Nkx, Ntr)
            # k'_l[z, r, s] = sum_j k_j[z, r, s] B[s, z, j, l]
            k = np.einsum("jzrs,szjl->lzrs", k, B)
            im = _type1(tabs, k.reshape(L * nz, Nkx * Ntr)).reshape((L,) + img)
            im = np.fft.ifft(im, axis=1, norm="forward")
        else:
            k = _type2(tabs, im).reshape(L, Nkx, Ntr)
            k = np.einsum("jrs,sjl->lrs", k, B)
            im = _type1(tabs, k.reshape(L, Nkx * Ntr))
        out += im * np.conj(se)[None]
    return out


def dense_aha(g: Geometry, x: np.ndarray, SEs: np.ndarray, Phi: np.ndarray,
              frames: np.ndarray, traj: np.ndarray) -> np.ndarray:
    """Brute force from the data model of include/mr_aha/mr_aha.h:48-60.

    The forward model of coil c is the explicit matrix
    A_c[m, (l, r)] = Phi(l, t_m) e^{-i r.x_m} S_c(r): every sample m of
    readout s is acquired at time frame t_m = frames[s] with rank x time
    basis Phi (L, T). Returns sum_c A_c^H A_c x, formed with dense matrices.
    2D and 3D only (tiny sizes).
    """
    assert g.traj_type in ("2D", "3D")
    L = g.L
    img = g.img_shape()
    ndim, Nkx, Ntr = traj.shape
    # E[m, r] = e^{-i r.x_m} over the flattened row-major image.
    grids = np.meshgrid(*[np.arange(s) - s // 2 for s in img], indexing="ij")
    coords = np.stack([gr.ravel() for gr in grids])  # (nsp, Npix), row-major axes
    pts = np.stack([traj[a].astype(np.float64).ravel() for a in reversed(range(ndim))])
    E = np.exp(-1j * (pts.T @ coords))  # (Mnu, Npix)
    t_m = np.tile(np.asarray(frames), Nkx)  # sample m = r*Ntr + s
    PhiM = Phi[:, t_m].T  # (Mnu, L)
    xv = np.asarray(x, dtype=np.complex128).reshape(L, -1)
    S = np.asarray(SEs, dtype=np.complex128).reshape(SEs.shape[0], -1)
    out = np.zeros_like(xv)
    for c in range(S.shape[0]):
        Ac = np.concatenate([PhiM[:, l:l + 1] * E * S[c][None, :] for l in range(L)], axis=1)
        y = Ac @ xv.ravel()
        out += (Ac.conj().T @ y).reshape(L, -1)
    return out.reshape((L,) + img)
