"""CPU baseline: the oracle's Toeplitz A^H A restated for speed (oracle side).

Test infrastructure: see ``oracle/__init__.py``. Used only by bench.py's
CPU-baseline leg, timed on the host cores.

Same five stages as ``oracle.toeplitz.apply`` (src/cuda/mr_aha.cu:694-790),
in complex64 with multi-threaded pocketfft (scipy.fft, ``workers``), the
block multiply walking the packed triangle once per z-slab so that no full
L x L x grid block array is formed.
"""

from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from .conventions import pad_index, tri_pairs


def toeplitz_apply_coils(n, M, L, x, S, K, upper, coils, workers, out=None):
    """A^H A restricted to ``coils`` (an iterable of coil indices).

    n, M: column-major extents (3 entries). x (L, *img), S (C, *img),
    K (T, *grid) centered, all row-major numpy complex64. Returns the partial
    sum over ``coils``.
    """
    nsp = len(x.shape) - 1
    img = x.shape[1:]
    grid = K.shape[1:]
    idx = [pad_index(np.arange(n[a]), n[a], M[a]) for a in reversed(range(nsp))]
    # natural-order kernel view: natural j holds centered (j + M//2) mod M
    roll = [int(M[a] // 2) for a in reversed(range(nsp))]
    pairs = tri_pairs(L, upper)
    axes = tuple(range(1, 1 + nsp))
    if out is None:
        out = np.zeros_like(x)
    for c in coils:
        pad = np.zeros((L,) + grid, dtype=np.complex64)
        pad[(slice(None),) + np.ix_(*idx)] = S[c][None] * x
        V = sfft.fftn(pad, axes=axes, workers=workers, overwrite_x=True)
        del pad
        Y = np.zeros_like(V)
        # slabs along the slowest grid axis
        lead = grid[0]
        step = max(1, lead // 8)
        for s0 in range(0, lead, step):
            s1 = min(lead, s0 + step)
            nat = np.arange(s0, s1)
            cen0 = np.mod(nat + roll[0], grid[0])
            Ks = K[:, cen0]
            for a in range(1, nsp):
                Ks = np.roll(Ks, -roll[a], axis=1 + a)
            Vs = V[:, s0:s1]
            Ys = Y[:, s0:s1]
            for t, (i, j) in enumerate(pairs):
                # stored entry (i, j): y_j += v_i A_ij ; mirror y_i += v_j conj(A_ij)
                Ys[j] += Vs[i] * Ks[t]
                if i != j:
                    Ys[i] += Vs[j] * np.conj(Ks[t])
        del V
        u = sfft.ifftn(Y, axes=axes, workers=workers, norm="forward", overwrite_x=True)
        out += np.conj(S[c])[None] * u[(slice(None),) + np.ix_(*idx)]
    return out
