"""Ranks above the compiled block kernels (L = 17 .. 512): the tiled block
multiply (outputs x coils blocked in registers, kernel entries streamed,
coil batches of up to 12).

Ranks of the reference's own rank checks (tests/check_bd_kernels.py:52: 17,
20, 30-33, 47, 72, 130) plus 256 and 512, both triangles, masked and
unmasked, random packed kernels with complex diagonals, against the FP64
five-stage apply (relative 2-norm 1e-5); every coil-batch size and output
blocking against the others (1e-6: the same sums in the same order, so in
practice bit-equal).
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


def host(t):
    return t.detach().cpu().numpy()


# (L, image n, coils): small grids for the large ranks (the FP64 oracle forms
# the full L x L block of every location)
RANKS = [(17, (12, 10), 3), (20, (12, 10), 5), (30, (10, 8), 3), (31, (10, 8), 2),
         (32, (10, 8), 4), (33, (10, 8), 3), (47, (8, 8), 5), (72, (8, 6), 3),
         (130, (6, 5), 2), (256, (4, 4), 2), (512, (4, 3), 1),
         # one batch of 12 coils, and 20 coils in batches
         (40, (8, 6), 12), (20, (10, 8), 20)]


@pytest.mark.parametrize("L,n,C", RANKS)
@pytest.mark.parametrize("masked", [False, True])
def test_large_rank_block_multiply(mr, monkeypatch, L, n, C, masked):
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[rank] L={L} n={n} C={C} masked={masked}", flush=True)
    rng = np.random.default_rng(1000 + L + C)
    g = cv.make_geometry("2D", n, L, C)
    mask = None
    if masked:
        keep = rng.random(g.P) < 0.6
        mask = np.flatnonzero(keep).astype(np.int64)
        kshape = (g.T, mask.size)
    else:
        kshape = (g.T,) + g.grid_shape()
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    ish = (L,) + g.img_shape()
    x = (rng.standard_normal(ish) + 1j * rng.standard_normal(ish)).astype(np.complex64)
    csh = (C,) + g.img_shape()
    S = (rng.standard_normal(csh) + 1j * rng.standard_normal(csh)).astype(np.complex64)
    dx, dS, dK = dev(x), dev(S), dev(K)
    dmask = dev(mask) if masked else None
    for tri in ("lower", "upper"):
        cfg = mr.make_config("2D", n, L, n_coils=C, triangulation=tri, is_masked=masked,
                             num_masked=int(mask.size) if masked else 0)
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128), upper=(tri == "upper"), mask=mask)
        outs = []
        for ncb, opt in (("12", "0"), ("4", "0"), ("2", "0"), ("1", "0"), ("12", "4")):
            monkeypatch.setenv("MR_AHA_HIP_BDT_NCB", ncb)
            monkeypatch.setenv("MR_AHA_HIP_BDT_OPT", opt)
            y = host(mr.aha_toeplitz(cfg, dx, dS, dK, dmask))
            outs.append(y)
        e = rel(outs[0], ref)
        print(f"[rank]   {tri}: vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5, (tri, e)
        for y in outs[1:]:
            assert rel(y, outs[0]) < 1e-6
