"""Exact Toeplitz kernel and five-stage Toeplitz apply (oracle, FP64).

Test infrastructure: see ``oracle/__init__.py``.
"""

from __future__ import annotations

import numpy as np

from .conventions import Geometry, gram_blocks, hermitian_block, pad_index, tri_pairs


def dirichlet_matrix(coords: np.ndarray, M: int) -> np.ndarray:
    """D[m, k] = sum_{d=-floor(M/2)}^{ceil(M/2)-1} exp(i d (2 pi kappa/M - x_m)).

    ``kappa = k - floor(M/2)`` is the signed coordinate of centered index k.
    Evaluated as the direct sum over d (no closed form, so no singular points).
    This is the centered DFT of one type-1 NUFFT term followed by the
    centered inverse DFT of src/cuda/mr_calc_toep.cu:10-24, 569-618, with the
    mode range of include/mr_aha/mr_aha.h:1169-1178.
    """
    d = np.arange(M) - M // 2                      # signed mode coordinates
    kappa = np.arange(M) - M // 2
    E = np.exp(-1j * np.outer(coords.astype(np.float64), d))   # e^{-i d x_m}
    F = np.exp(2j * np.pi * np.outer(d, kappa) / M)            # e^{+2 pi i kappa d / M}
    return E @ F


def strengths(phi2: np.ndarray, g: Geometry, upper: bool) -> np.ndarray:
    """Packed Gram entries per readout: (T, Ntrajs) or (T, Nkz, Ntrajs).

    Component t is stored block entry tri_pairs[t] of every readout's block
    (src/cuda/mr_kernels.cu:307-333 packing, :335-362 broadcast; the
    broadcast along the readout is implicit in ``calc_kernel``).
    """
    B = gram_blocks(phi2)  # (Ntrajs, [Nkz,] L, L) with B[..., i, j] = entry (i, j)
    pairs = tri_pairs(g.L, upper)
    c = np.stack([B[..., i, j] for (i, j) in pairs])  # (T, Ntrajs, [Nkz])
    if g.traj_type == "stack":
        c = np.swapaxes(c, 1, 2)  # (T, Nkz, Ntrajs)
    return c.astype(np.complex128)


def calc_kernel(g: Geometry, phi2: np.ndarray, traj: np.ndarray, upper: bool = False,
                mask: np.ndarray | None = None) -> np.ndarray:
    """Exact ``phi2_toep`` (T, [Mz,] My, Mx), centered order, complex128.

    phi2_toep[kappa, t] = (1/prod M_tr) sum_m c_{t,m} prod_a D_{M_a}(2 pi kappa_a/M_a - x_{m,a})
    over the transformed axes: x, y for 2D and stack (each kz plane with its
    own Gram blocks, no transform or shift along z, scale 1/(Mx My)); x, y, z
    for 3D (scale 1/(Mx My Mz)) — src/cuda/mr_calc_toep.cu:171-200.
    With ``mask`` (0-based centered linear indices), returns (T, num_masked)
    gathered from the full kernel (src/cuda/mr_kernels.cu:427-462).

    Pinned by: the closed form of full Cartesian sampling, agreement with the
    direct NUDFT normal operator (``oracle.nudft``) through ``apply``, and the
    dense brute-force forward model (tests/test_oracle.py).
    """
    Mx, My, Mz = g.M
    T = g.T
    c = strengths(phi2, g, upper)
    ndim, Nkx, Ntr = traj.shape
    if g.traj_type in ("2D", "stack"):
        Dx = dirichlet_matrix(traj[0].ravel(), Mx).reshape(Nkx, Ntr, Mx)
        Dy = dirichlet_matrix(traj[1].ravel(), My).reshape(Nkx, Ntr, My)
        # Per-readout contribution G_s[ky, kx] = sum_r Dy[r,s,ky] Dx[r,s,kx].
        G = np.matmul(np.transpose(Dy, (1, 2, 0)), np.transpose(Dx, (1, 0, 2)))  # (Ntr, My, Mx)
        Gf = G.reshape(Ntr, My * Mx)
        if g.traj_type == "2D":
            K = (c @ Gf).reshape(T, My, Mx)
        else:
            K = (c.reshape(T * Mz, Ntr) @ Gf).reshape(T, Mz, My, Mx)
        K /= Mx * My
    else:
        Dx = dirichlet_matrix(traj[0].ravel(), Mx)   # (M_nu, Mx)
        Dy = dirichlet_matrix(traj[1].ravel(), My)
        Dz = dirichlet_matrix(traj[2].ravel(), Mz)
        Mnu = Dx.shape[0]
        cm = np.repeat(c[:, None, :], Nkx, axis=1).reshape(T, Mnu)  # sample m = r*Ntr + s
        DyDx = (Dy[:, :, None] * Dx[:, None, :]).reshape(Mnu, My * Mx)
        A = (cm[:, :, None] * Dz[None, :, :])                        # (T, Mnu, Mz)
        A = np.transpose(A, (0, 2, 1)).reshape(T * Mz, Mnu)
        K = (A @ DyDx).reshape(T, Mz, My, Mx) / (Mx * My * Mz)
    if mask is not None:
        K = K.reshape(T, -1)[:, np.asarray(mask)]
    return K


def _dirichlet_cols(x: np.ndarray, kappas: np.ndarray, M: int) -> np.ndarray:
    """D_M(2 pi kappa/M - x_m) for samples x (S,) and signed coordinates
    kappas (K,), shape (S, K): the closed form
    e^{-i th (M - 2 ceil(M/2) + 1)/2} sin(M th/2) / sin(th/2), and the direct
    sum over the M modes where |sin(th/2)| < 1e-6 (the closed form's
    removable singularity). Same function as ``dirichlet_matrix``."""
    th = 2.0 * np.pi * kappas[None, :].astype(np.float64) / M - x[:, None].astype(np.float64)
    s = np.sin(th / 2)
    shift = (M - 2 * ((M + 1) // 2) + 1) / 2.0
    near = np.abs(s) < 1e-6
    out = np.exp(-1j * th * shift) * np.sin(M * th / 2) / np.where(near, 1.0, s)
    if near.any():
        d = np.arange(M) - M // 2
        out[near] = np.exp(1j * np.outer(th[near], d)).sum(axis=1)
    return out


def kernel_entries(g: Geometry, phi2: np.ndarray, traj: np.ndarray, sel: np.ndarray,
                   upper: bool = False, chunk: int = 1 << 18) -> np.ndarray:
    """Exact entries of ``phi2_toep`` at a few locations, without the full grid.

    ``sel`` is an int array (E, 1 + n_spatial) of (t, [kz,] ky, kx) centered
    indices, the row-major index order of ``calc_kernel``'s result. Returns
    complex128 (E,). Same sum as ``calc_kernel`` (per transformed axis a
    Dirichlet factor of the signed coordinate kappa_a = k_a - floor(M_a/2),
    strengths constant along a readout, scale 1/prod M_tr), evaluated per
    entry over all samples in chunks, so the full-size workloads can be
    checked at sampled entries. Cheap when the entries share few distinct
    coordinates per axis (each distinct value costs one pass over the samples).

    Pinned against ``calc_kernel`` at every entry of small 2D, stack and 3D
    grids with odd and even M (tests/test_oracle.py).
    """
    sel = np.asarray(sel, dtype=np.int64)
    Mx, My, Mz = g.M
    c = strengths(phi2, g, upper)            # (T, Ntr) or (T, Nkz, Ntr)
    ndim, Nkx, Ntr = traj.shape
    three = g.traj_type == "3D"
    if three:
        t, kz, ky, kx = sel.T
        axes = [(kx, Mx, traj[0]), (ky, My, traj[1]), (kz, Mz, traj[2])]
        scale = 1.0 / (Mx * My * Mz)
    else:
        if g.traj_type == "stack":
            t, kz, ky, kx = sel.T
        else:
            t, ky, kx = sel.T
            kz = np.zeros_like(t)
        axes = [(kx, Mx, traj[0]), (ky, My, traj[1])]
        scale = 1.0 / (Mx * My)
    # distinct coordinate tuples: their per-readout Dirichlet products G[s]
    coords = np.stack([a[0] - a[1] // 2 for a in axes], axis=1)     # signed kappa per entry
    uniq, inv = np.unique(coords, axis=0, return_inverse=True)
    inv = inv.ravel()
    G = np.zeros((len(uniq), Ntr), dtype=np.complex128)
    xs = [np.asarray(a[2], dtype=np.float64).ravel() for a in axes]  # sample m = r*Ntr + s
    nsamp = Nkx * Ntr
    for m0 in range(0, nsamp, chunk):
        m1 = min(nsamp, m0 + chunk)
        prod = np.ones((m1 - m0, len(uniq)), dtype=np.complex128)
        for ai, (_, M, _) in enumerate(axes):
            vals, vinv = np.unique(uniq[:, ai], return_inverse=True)
            D = _dirichlet_cols(xs[ai][m0:m1], vals, M)
            prod *= D[:, vinv.ravel()]
        s_idx = np.arange(m0, m1) % Ntr
        for u in range(len(uniq)):
            G[u] += np.bincount(s_idx, weights=prod[:, u].real, minlength=Ntr) + \
                1j * np.bincount(s_idx, weights=prod[:, u].imag, minlength=Ntr)
    if g.traj_type == "stack":
        cs = c[t, kz]                        # (E, Ntr)
    else:
        cs = c[t]
    return scale * np.einsum("es,es->e", cs, G[inv])


def _img_to_grid_index(g: Geometry):
    """Natural padded indices of every image index, per row-major axis."""
    return [pad_index(np.arange(g.n[a]), g.n[a], g.M[a]) for a in reversed(range(g.n_spatial))]


def apply(g: Geometry, x: np.ndarray, SEs: np.ndarray, phi2_toep: np.ndarray,
          upper: bool = False, mask: np.ndarray | None = None) -> np.ndarray:
    """Five-stage Toeplitz A^H A in FP64 (src/cuda/mr_aha.cu:694-790).

    Per coil: pad(S_c x) onto the natural-order upsampled grid
    (mr_kernels.cu:98-142), unnormalised forward DFT over the spatial axes,
    per-point v <- v A with A rebuilt Hermitian from the packed triangle at
    the centered index of the natural frequency (mr_kernels.cu:73-92,
    178-221), unnormalised inverse DFT, crop and conj(S_c) accumulate
    (mr_kernels.cu:144-176). With ``mask``, the kernel is [T, num_masked]
    and every location outside the mask multiplies by zero
    (src/cuda/mr_aha.cu:673-690).

    Pinned by: the direct NUDFT (``oracle.nudft.aha_nudft``), the dense
    forward model, Hermitian/PSD invariants and upper == lower.
    """
    L = g.L
    nsp = g.n_spatial
    img = g.img_shape()
    grid = g.grid_shape()
    x = np.asarray(x, dtype=np.complex128).reshape((L,) + img)
    S = np.asarray(SEs, dtype=np.complex128).reshape((-1,) + img)
    # Full Hermitian blocks on the centered grid, then permuted to natural order.
    if mask is not None:
        full = np.zeros((g.T, g.P), dtype=np.complex128)
        full[:, np.asarray(mask)] = phi2_toep
        Kc = full.reshape((g.T,) + grid)
    else:
        Kc = np.asarray(phi2_toep, dtype=np.complex128).reshape((g.T,) + grid)
    A = hermitian_block(Kc, L, upper)  # (L, L, *grid) centered, A[i, j] = entry (i, j)
    for ax in range(nsp):
        M = grid[ax]
        # natural index j holds centered index (j + M//2) mod M
        A = np.take(A, np.mod(np.arange(M) + M // 2, M), axis=2 + ax)
    idx = _img_to_grid_index(g)
    axes = tuple(range(1, 1 + nsp))
    out = np.zeros((L,) + img, dtype=np.complex128)
    for c in range(S.shape[0]):
        pad = np.zeros((L,) + grid, dtype=np.complex128)
        pad[(slice(None),) + np.ix_(*idx)] = S[c][None] * x
        V = np.fft.fftn(pad, axes=axes)
        # y_i = sum_j v_j A[j, i]
        Y = np.einsum("j...,ji...->i...", V, A)
        u = np.fft.ifftn(Y, axes=axes, norm="forward")
        out += np.conj(S[c])[None] * u[(slice(None),) + np.ix_(*idx)]
    return out
