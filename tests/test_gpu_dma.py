"""The double-buffered axis pass (global-to-LDS row loads, first stage out of
place, later stages in place on [element][16 pencils]) against the one-tile
16-pencil pass it replaces where the tile rows are 128-byte lines: the same
butterflies and twiddle products (relative 2-norm 1e-6: the two kernels are
compiled separately, and the compiler may contract different products into
fused multiply-adds, so results agree to rounding, not bit for bit); and
against the FP64 five-stage apply (relative 2-norm 1e-5).

Also the double-buffered outer (z) passes of a stack with 56 partitions
(tiles of 16 xy x 8 ranks: S x formed in the first stage; the inverse sums
conj(S) times the coils in registers), one coil group and several (the later
groups add to the sum), against the one-tile outer passes.

Shapes: a stack whose y pass runs at length 512 (x-last route: y pass with
zero padding, inverse y pass with the crop) and a 2D rank-17 case on the
separate-pass route (y passes at length 640, x rows 640 apart; compared with
the one-tile pass only, its rank kernel's parity is tested at small grids).
"""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def mr():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import mr_aha_amd
    return mr_aha_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


CASES = [("stack", (256, 256, 4), 4, 2), ("2D", (320, 320), 17, 2)]


@pytest.mark.parametrize("traj,n,L,C", CASES)
def test_dma_pass_matches_one_tile_pass(mr, monkeypatch, traj, n, L, C):
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[dma] {traj} n={n} L={L} C={C}", flush=True)
    rng = np.random.default_rng(7 + L)
    g = cv.make_geometry(traj, n, L, C)
    kshape = (g.T,) + g.grid_shape()
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape())
         + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((C,) + g.img_shape())
         + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    cfg = mr.make_config(traj, n, L, n_coils=C)
    dx, dS, dK = dev(x), dev(S), dev(K)
    outs = {}
    for dma in ("1", "0"):
        monkeypatch.setenv("MR_AHA_HIP_FFT_DMA", dma)
        outs[dma] = mr.aha_toeplitz(cfg, dx, dS, dK).cpu().numpy()
    d = rel(outs["1"], outs["0"])
    print(f"[dma]   vs one-tile pass {d:.2e}", flush=True)
    assert d < 1e-6
    if L > 16:
        return   # the rank kernels' parity is tested at small grids
    ref = toeplitz.apply(g, x, S, K.astype(np.complex128))
    e = rel(outs["1"], ref)
    print(f"[dma]   vs FP64 {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("group", ["0", "2"])
def test_zdma_outer_passes(mr, monkeypatch, group):
    from oracle import conventions as cv
    from oracle import toeplitz
    traj, n, L, C = "stack", (48, 40, 56), 8, 5
    print(f"[zdma] {traj} n={n} L={L} C={C} coil group {group}", flush=True)
    if group != "0":
        monkeypatch.setenv("MR_AHA_HIP_COIL_GROUP", group)
    rng = np.random.default_rng(11)
    g = cv.make_geometry(traj, n, L, C)
    kshape = (g.T,) + g.grid_shape()
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape())
         + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((C,) + g.img_shape())
         + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    cfg = mr.make_config(traj, n, L, n_coils=C)
    dx, dS, dK = dev(x), dev(S), dev(K)
    outs = {}
    for zdma in ("1", "0"):
        monkeypatch.setenv("MR_AHA_HIP_ZDMA", zdma)
        outs[zdma] = mr.aha_toeplitz(cfg, dx, dS, dK).cpu().numpy()
    d = rel(outs["1"], outs["0"])
    print(f"[zdma]   vs one-tile passes {d:.2e}", flush=True)
    assert d < 1e-6
    ref = toeplitz.apply(g, x, S, K.astype(np.complex128))
    e = rel(outs["1"], ref)
    print(f"[zdma]   vs FP64 {e:.2e}", flush=True)
    assert e < 1e-5

