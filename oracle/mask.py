"""Neighbourhood mask restated in numpy (oracle, integer work: bit-exact).

Test infrastructure: see ``oracle/__init__.py``.
"""

from __future__ import annotations

import numpy as np

from .conventions import Geometry

TAU = 6.283185307179586


def calc_mask(g: Geometry, traj: np.ndarray, radius: int) -> np.ndarray:
    """0-based centered linear indices kept by the neighbourhood mask.

    src/cuda/mr_calc_mask.cu:63-181, 329-341: every sample maps to grid index
    rintf(x * (float)(M / 2pi)) + M/2 (mod M) per transformed axis, computed
    in float32 with round-half-even; marks are dilated by a (2r+1)^d box with
    periodic wrap (Chebyshev radius r); indices ascend; a stack replicates the
    in-plane mask over its Mz planes (plane-major).

    Pinned by the closed form of a single on-grid sample (a (2r+1)^2 box) and
    by brute force on tiny grids (tests/test_oracle.py).
    """
    three = g.traj_type == "3D"
    m = [g.M[0], g.M[1], g.M[2] if three else 1]
    nd = 3 if three else 2
    shape_rm = tuple(reversed(m))  # row-major (z, y, x)
    marks = np.zeros(shape_rm, dtype=bool)
    idx = []
    for a in range(nd):
        sc = np.float32(m[a] / TAU)
        prod = (traj[a].astype(np.float32).ravel() * sc).astype(np.float32)
        i = np.rint(prod).astype(np.int64) + m[a] // 2
        idx.append(np.mod(i, m[a]))
    if nd == 2:
        marks[0, idx[1], idx[0]] = True
    else:
        marks[idx[2], idx[1], idx[0]] = True
    r = g_radius = int(radius)
    if g_radius > 0:
        for ax_cm in range(3):
            if m[ax_cm] <= 1:
                continue
            ax_rm = 2 - ax_cm
            acc = marks.copy()
            for d in range(-r, r + 1):
                if d:
                    acc |= np.roll(marks, d, axis=ax_rm)
            marks = acc
    flat = np.flatnonzero(marks.ravel())  # row-major (z,y,x) flat == column-major linear
    if g.traj_type == "stack":
        plane = m[0] * m[1]
        flat = np.concatenate([flat + plane * kz for kz in range(g.M[2])])
    return flat.astype(np.int64)
