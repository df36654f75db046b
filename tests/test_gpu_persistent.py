"""Persistent form of the fused x pass (one 1024-thread workgroup per CU
walking a range of (row, coil-pair) items, the next item's image rows
streamed into the LDS slots the current result leaves free).

It runs where the image fills the centred half of the x grid (x grid 512 or
640, L = 16, FP32 kernel, two or more coils). The shapes below give every
workgroup several items, so the prefetch of the next item is exercised, and
three coils make the items alternate between a coil pair and a lone coil
(the two store counts the prefetch wait depends on). Its result must equal
the one-item form within 1e-6 (the same operations in the same order; the
compiler may contract a product into an FMA in one kernel and not the other,
so the last bit can differ) and the FP64 five-stage apply within 1e-5;
masked and unmasked, both triangles. A row computed from stale LDS slots
would miss both bounds by orders of magnitude.
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


def xbd_form(cfg):
    """Fused-pass form a plan for cfg launches (mr_aha_plan_xbd_form)."""
    import ctypes
    from mr_aha_amd import _lib
    plan = ctypes.c_void_p()
    _lib.check(_lib.lib.mr_aha_plan_create(ctypes.byref(cfg), ctypes.byref(plan)), "plan_create")
    try:
        return int(_lib.lib.mr_aha_plan_xbd_form(plan))
    finally:
        _lib.lib.mr_aha_plan_destroy(plan)


# (image n, coils): x grid 640 and 512 (grid 2n x 2n x nz); rows x coil
# groups = 1024 and 512 items over the workgroups, one per CU
SHAPES = [((320, 32, 8), 3), ((256, 32, 8), 2)]


@pytest.mark.parametrize("n,C", SHAPES)
@pytest.mark.parametrize("masked", [False, True])
def test_persistent_fused_pass(mr, monkeypatch, n, C, masked):
    from oracle import conventions as cv
    from oracle import toeplitz
    L = 16
    print(f"[persistent] stack {n} L={L} C={C} masked={masked}", flush=True)
    rng = np.random.default_rng(n[0] + C + int(masked))
    g = cv.make_geometry("stack", n, L, C)
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
    nn = g.n[: g.n_spatial]
    M = g.M[: g.n_spatial]
    for tri in ("lower", "upper"):
        cfg = mr.make_config("stack", nn, L, phi2_dim=M, n_coils=C, triangulation=tri,
                             is_masked=masked, num_masked=int(mask.size) if masked else 0)
        outs = {}
        for pd in ("1", "0"):
            monkeypatch.setenv("MR_AHA_HIP_XBD_PD", pd)
            assert xbd_form(cfg) == (2 if pd == "1" else 1)
            outs[pd] = host(mr.aha_toeplitz(cfg, dx, dS, dK, dmask))
        d = rel(outs["1"], outs["0"])
        print(f"[persistent]   {tri}: persistent vs one-item form {d:.2e}", flush=True)
        assert d < 1e-6, (tri, d)
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128), upper=(tri == "upper"), mask=mask)
        e = rel(outs["1"], ref)
        print(f"[persistent]   {tri}: vs FP64 {e:.2e}", flush=True)
        assert e < 1e-5, (tri, e)
