"""Block step of the fused x pass on the matrix cores (v_mfma_f32_4x4x1_16b_f32).

Every compiled x length (256, 320, 384, 512, 640) with ranks 4, 8 and 16, an
even and an odd coil count (a last coil without a partner runs with the
unused rows of its blocks zeroed), the two-coil 1024-thread workgroups and
the 512-thread ones, both triangles. Random packed kernels (complex
diagonal included) and data. Against:
  * the FP64 five-stage apply on the same kernel (relative 2-norm 1e-5);
  * the FMA block walks of the same kernel (MR_AHA_HIP_MFMA=0), 1e-6: both
    FP32, different summation order;
  * the plan's own report that the matrix-core step is in use.
A kernel pointer that is not 16-byte aligned falls back to the FMA walks and
must still match.
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


# (image n, grid M, L, C, wide): x grid lengths 256..640
CASES = [((128, 10), (256, 20), 16, 3, True), ((160, 8), (320, 16), 16, 2, True),
         ((192, 6), (384, 12), 16, 3, True), ((256, 6), (512, 12), 16, 3, True),
         ((320, 6), (640, 12), 16, 2, True), ((320, 6), (640, 12), 16, 3, False),
         ((256, 6), (512, 12), 16, 2, False),
         ((160, 8), (320, 16), 8, 5, True), ((320, 6), (640, 12), 8, 3, True),
         ((128, 8), (256, 16), 4, 5, True), ((192, 6), (384, 12), 4, 2, True),
         ((320, 6), (640, 12), 4, 3, True)]


def _inputs(g, seed):
    rng = np.random.default_rng(seed)
    kshape = (g.T,) + g.grid_shape()
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    ish = (g.L,) + g.img_shape()
    x = (rng.standard_normal(ish) + 1j * rng.standard_normal(ish)).astype(np.complex64)
    csh = (g.n_coils,) + g.img_shape()
    S = (rng.standard_normal(csh) + 1j * rng.standard_normal(csh)).astype(np.complex64)
    return K, x, S


@pytest.mark.parametrize("n,M,L,C,wide", CASES)
def test_mfma_block_step(mr, monkeypatch, n, M, L, C, wide):
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[mfma] n={n} M={M} L={L} C={C} wide={wide}", flush=True)
    if not wide:
        monkeypatch.setenv("MR_AHA_HIP_XBD_WIDE", "0")
    g = cv.make_geometry("2D", n, L, C, phi2_dim=M)
    K, x, S = _inputs(g, 31 * L + C + M[0])
    dK, dx, dS = dev(K), dev(x), dev(S)
    assert dK.data_ptr() % 16 == 0
    for tri in ("lower", "upper"):
        cfg = mr.make_config("2D", n, L, phi2_dim=M, n_coils=C, triangulation=tri)
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128), upper=(tri == "upper"))
        monkeypatch.setenv("MR_AHA_HIP_MFMA", "1")
        with mr.create_op(cfg, dS, None, None, _phi2_toep=dK) as op:
            t = op.tiles
            assert t["mfma"], t
            assert t["wide"] == (wide and L == 16 and M[0] in (512, 640)), t
            y = host(op(dx))
        monkeypatch.setenv("MR_AHA_HIP_MFMA", "0")
        with mr.create_op(cfg, dS, None, None, _phi2_toep=dK) as op:
            assert not op.tiles["mfma"]
            yf = host(op(dx))
        e, ef = rel(y, ref), rel(y, yf)
        print(f"[mfma]   {tri}: vs FP64 {e:.2e}, vs FMA walks {ef:.2e}", flush=True)
        assert e < 1e-5, (tri, e)
        assert ef < 1e-6, (tri, ef)


def test_mfma_unaligned_kernel_falls_back(mr, monkeypatch):
    """Kernel pointer 8 bytes past a 16-byte boundary: the 16-byte entry loads
    cannot serve it, the launch takes the FMA walks and the result is the
    same."""
    from oracle import conventions as cv
    from oracle import toeplitz
    n, M, L, C = (320, 6), (640, 12), 16, 2
    g = cv.make_geometry("2D", n, L, C, phi2_dim=M)
    K, x, S = _inputs(g, 5)
    buf = torch.empty(K.size + 1, dtype=torch.complex64, device="cuda")
    dK = buf[1:].view(K.shape)
    dK.copy_(dev(K))
    assert dK.data_ptr() % 16 == 8
    cfg = mr.make_config("2D", n, L, phi2_dim=M, n_coils=C)
    ref = toeplitz.apply(g, x, S, K.astype(np.complex128))
    monkeypatch.setenv("MR_AHA_HIP_MFMA", "1")
    y_off = host(mr.aha_toeplitz(cfg, dev(x), dev(S), dK))
    y_al = host(mr.aha_toeplitz(cfg, dev(x), dev(S), dev(K)))
    assert rel(y_off, ref) < 1e-5
    assert rel(y_al, ref) < 1e-5
    assert rel(y_off, y_al) < 1e-6


def test_mfma_stack_and_3d(mr, monkeypatch):
    """Stack-of-stars and 3D grids (x length 320 and 256) through the
    matrix-core step, against the FP64 apply and the FMA walks."""
    from oracle import conventions as cv
    from oracle import toeplitz
    for tt, n, L, C in (("stack", (160, 6, 4), 16, 3), ("3D", (128, 6, 4), 8, 2)):
        g = cv.make_geometry(tt, n, L, C)
        K, x, S = _inputs(g, 3 * L + C)
        cfg = mr.make_config(tt, n, L, n_coils=C)
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128))
        outs = []
        for mf in ("1", "0"):
            monkeypatch.setenv("MR_AHA_HIP_MFMA", mf)
            outs.append(host(mr.aha_toeplitz(cfg, dev(x), dev(S), dev(K))))
        assert rel(outs[0], ref) < 1e-5, (tt, rel(outs[0], ref))
        assert rel(outs[0], outs[1]) < 1e-6, (tt, rel(outs[0], outs[1]))
