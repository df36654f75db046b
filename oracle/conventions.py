"""Index and storage conventions of the Toeplitz A^H A path (oracle, FP64).

Test infrastructure: see ``oracle/__init__.py``.

Arrays follow the reference's Python binding: row-major shapes that are the
reversed column-major shapes of the C API (include/mr_aha/mr_aha.h:19-31):

    x, AhAx      (L, [Nz,] Ny, Nx)           == C [Nx, Ny, (Nz,) L]
    SEs          (C, [Nz,] Ny, Nx)           == C [Nx, Ny, (Nz,) C]
    phi2         (L, L, [Nkz,] Ntrajs)       == C [Ntrajs, (Nkz,) L, L]
    traj         (ndim, Nkx, Ntrajs)         == C [Ntrajs, Nkx, ndim]
    phi2_toep    (T, [Mz,] My, Mx)           == C [Mx, My, (Mz,) T]

``phi2[j, i, ...]`` holds entry (i, j) of the Hermitian Gram block
(include/mr_aha/mr_aha.h:33-69).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TRAJ_TYPES = ("2D", "3D", "stack")


def num_tri(L: int) -> int:
    """(L^2+L)/2, or 0 for L <= 0 (src/common/mr_config.cpp:179-183).

    Pinned by the known answer num_tri(16) == 136 in tests/test_oracle.py.
    """
    return L * (L + 1) // 2 if L > 0 else 0


def tri_pairs(L: int, upper: bool) -> list[tuple[int, int]]:
    """Block entries (i, j) of the stored triangle in packed order.

    Column-major packing of MATLAB ``A(tril(true(L)))`` / ``A(triu(...))``
    (include/mr_aha/mr_aha.h:299-317). Pinned by the literal L=3 fixture of
    python/mr_aha/_psf.py:104-105 and by the closed-form offsets of
    ``tri_offset`` (test_oracle.py).
    """
    if upper:
        return [(i, j) for j in range(L) for i in range(j + 1)]
    return [(i, j) for j in range(L) for i in range(j, L)]


def tri_offset(i: int, j: int, L: int, upper: bool) -> tuple[int, bool]:
    """Packed offset of block entry (i, j) and whether to conjugate.

    Closed forms of include/mr_aha/mr_aha.h:307-309 (lower
    ``t = j*L - j*(j-1)/2 + (i-j)``, upper ``t = j*(j+1)/2 + i``); an entry of
    the unstored triangle is answered by its mirror, conjugated
    (src/cuda/mr_internal.cuh:155-181).
    """
    r, c, cj = i, j, False
    if upper:
        if r > c:
            r, c, cj = c, r, True
        return c * (c + 1) // 2 + r, cj
    if r < c:
        r, c, cj = c, r, True
    return c * L - c * (c - 1) // 2 + (r - c), cj


def cent2nat(k, M):
    """Centered index -> natural (FFT-order) index (mr_internal.cuh:119-127)."""
    return np.mod(np.asarray(k) - M // 2, M)


def nat2cent(j, M):
    """Natural index -> centered index (mr_internal.cuh:134-139)."""
    return np.mod(np.asarray(j) + M // 2, M)


def pad_index(i, N, M):
    """Image index -> natural padded index j = (i - floor(N/2)) mod M.

    src/cuda/mr_kernels.cu:111-128 (pad) and :158-165 (crop, its adjoint).
    """
    return np.mod(np.asarray(i) - N // 2, M)


@dataclass(frozen=True)
class Geometry:
    """Problem geometry in column-major axis order (x, y, z).

    ``n`` and ``M`` always have three entries; unused trailing axes are 1, as
    in mr_config (include/mr_aha/mr_aha.h:332-414).
    """

    traj_type: str
    n: tuple[int, int, int]
    M: tuple[int, int, int]
    L: int
    n_coils: int = 1

    @property
    def n_spatial(self) -> int:
        return 2 if self.traj_type == "2D" else 3

    @property
    def ndim_traj(self) -> int:
        return 3 if self.traj_type == "3D" else 2

    @property
    def T(self) -> int:
        return num_tri(self.L)

    @property
    def N(self) -> int:
        return int(np.prod(self.n))

    @property
    def P(self) -> int:
        return int(np.prod(self.M))

    @property
    def nkz(self) -> int:
        return self.n[2] if self.traj_type == "stack" else 1

    def img_shape(self) -> tuple:
        """Row-major image shape without the rank/coil axis."""
        return tuple(reversed(self.n[: self.n_spatial]))

    def grid_shape(self) -> tuple:
        """Row-major padded-grid shape."""
        return tuple(reversed(self.M[: self.n_spatial]))


def make_geometry(traj_type: str, n_dim, L: int, n_coils: int = 1,
                  phi2_dim=None) -> Geometry:
    """Geometry with the binding's default upsampled grid.

    Default ``phi2_dim = 2 * n_dim`` except ``Mz = Nz`` for a stack
    (python/mr_aha/api.py:286-290).
    """
    if traj_type not in TRAJ_TYPES:
        raise ValueError(traj_type)
    nsp = 2 if traj_type == "2D" else 3
    n_dim = [int(v) for v in n_dim]
    assert len(n_dim) == nsp
    if phi2_dim is None:
        M = [2 * v for v in n_dim]
        if traj_type == "stack":
            M[2] = n_dim[2]
    else:
        M = [int(v) for v in phi2_dim]
    n3 = tuple(n_dim + [1] * (3 - nsp))
    M3 = tuple(M + [1] * (3 - nsp))
    return Geometry(traj_type, n3, M3, int(L), int(n_coils))


def hermitian_block(packed: np.ndarray, L: int, upper: bool) -> np.ndarray:
    """Full L x L blocks from packed storage ``(T, ...)`` -> ``(L, L, ...)``.

    Entry [i, j] of the result is block entry (i, j) (NOT the row-major
    transpose used by phi2); the unstored half is the conjugate mirror
    (src/cuda/mr_internal.cuh:155-181).
    """
    out = np.empty((L, L) + packed.shape[1:], dtype=packed.dtype)
    for i in range(L):
        for j in range(L):
            t, cj = tri_offset(i, j, L, upper)
            out[i, j] = np.conj(packed[t]) if cj else packed[t]
    return out


def gram_blocks(phi2: np.ndarray) -> np.ndarray:
    """Gram blocks B[..., i, j] = block entry (i, j) from a binding-layout phi2.

    ``phi2`` is (L, L, [Nkz,] Ntrajs) with phi2[j, i, ...] = B(i, j); the
    result is (Ntrajs, [Nkz,] L, L).
    """
    return np.ascontiguousarray(np.transpose(phi2, tuple(range(phi2.ndim))[::-1]))
