"""GPU parity of the HIP path against the FP64 oracle and the golden fixtures.

Tolerances (relative 2-norm), as in the reference's tests
(tests/check_op.py:92-94, tests/check_reference.py:42-52):
  1e-5  HIP Toeplitz apply vs the FP64 five-stage apply on the same kernel,
        HIP kernel (ES gridding, eps 1e-6) vs the exact Dirichlet-sum kernel,
        HIP end-to-end vs the exact Toeplitz result
  1e-4  HIP Toeplitz vs the exact direct-NUDFT normal operator
  1e-6  operator object vs the stateless primitives (same method)
  exact integer equality for masks, bit equality for an injected kernel.
"""

import ctypes
import os

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


def cfg_for(mr, g, **kw):
    n = g.n[: g.n_spatial]
    M = g.M[: g.n_spatial]
    return mr.make_config(g.traj_type, n, g.L, phi2_dim=M, n_coils=g.n_coils, **kw)


GOLDEN_CASES = ["2d_50x40_L3_c3", "2d_33x27_L1_c2_odd", "2d_50x40_L4_c2_M101x81",
                "3d_24x20x16_L4_c2", "3d_25x21x15_L3_c2_odd", "stack_40x36x10_L3_c3",
                "stack_31x36x9_L4_c2_odd", "checkop_2d", "checkop_stack", "radial_128_L4_C4"]


def case(name):
    from tests.golden.make_golden import case_inputs
    return case_inputs(name)


@pytest.mark.parametrize("name", GOLDEN_CASES)
@pytest.mark.parametrize("tri", ["lower", "upper"])
def test_toeplitz_pipeline_vs_oracle(mr, golden, name, tri):
    from oracle import toeplitz
    print(f"[parity] {name} {tri}", flush=True)
    gd = golden(name)
    g, traj, phi2, x, S = case(name)
    cfg = cfg_for(mr, g, triangulation=tri)
    K = mr.calc_toep(cfg, dev(phi2), dev(traj))
    torch.cuda.synchronize()
    Kh = host(K)
    assert Kh.shape == (g.T,) + g.grid_shape()
    # kernel build vs exact Dirichlet-sum kernel
    e_k = rel(Kh.reshape(g.T, -1)[:, gd["kidx"]], gd[f"K_{tri}"])
    assert e_k < 1e-5, e_k
    y = host(mr.aha_toeplitz(cfg, dev(x), dev(S), K))
    # apply alone vs the FP64 five-stage apply on the same kernel
    e_a = rel(y, toeplitz.apply(g, x, S, Kh.astype(np.complex128), upper=(tri == "upper")))
    assert e_a < 1e-5, e_a
    # end to end vs the exact Toeplitz and the exact NUDFT normal operator
    assert rel(y, gd[f"AhAx_{tri}"]) < 1e-5
    assert rel(y, gd["AhAx_nudft"]) < 1e-4


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_calc_mask_bit_exact(mr, name):
    from oracle import mask as omask
    g, traj, phi2, x, S = case(name)
    cfg = cfg_for(mr, g)
    m = host(mr.calc_mask(cfg, dev(traj)))
    assert np.array_equal(m, omask.calc_mask(g, traj, cfg.mask_radius))


@pytest.mark.parametrize("name", ["2d_50x40_L3_c3", "3d_25x21x15_L3_c2_odd", "stack_31x36x9_L4_c2_odd"])
def test_masked_pipeline(mr, name):
    from oracle import toeplitz
    g, traj, phi2, x, S = case(name)
    cfg = cfg_for(mr, g)
    K = mr.calc_toep(cfg, dev(phi2), dev(traj))
    cfgm = cfg_for(mr, g, is_masked=True)
    mask = mr.calc_mask(cfgm, dev(traj))
    Km = mr.calc_toep(cfgm, dev(phi2), dev(traj), mask)
    assert cfgm.num_masked == mask.numel()
    # masked kernel == unmasked gathered at the mask (same arithmetic)
    assert np.array_equal(host(Km), host(K).reshape(g.T, -1)[:, host(mask)])
    y = host(mr.aha_toeplitz(cfgm, dev(x), dev(S), Km, mask))
    ref = toeplitz.apply(g, x, S, host(Km).astype(np.complex128), mask=host(mask))
    assert rel(y, ref) < 1e-5


@pytest.mark.parametrize("name", ["2d_50x40_L3_c3", "3d_24x20x16_L4_c2", "stack_40x36x10_L3_c3"])
def test_operator_object(mr, name):
    print(f"[op] {name}", flush=True)
    g, traj, phi2, x, S = case(name)
    dx, dS, dphi2, dtraj = dev(x), dev(S), dev(phi2), dev(traj)
    for tri in ("lower", "upper"):
        cfg = cfg_for(mr, g, triangulation=tri)
        K = mr.calc_toep(cfg, dphi2, dtraj)
        ref = mr.aha_toeplitz(cfg, dx, dS, K)
        for share in (True, False):
            with mr.create_op(cfg, dS, dphi2, dtraj, _phi2_toep=K, share_primary=share) as op:
                y = op(dx)
                assert torch.equal(y, ref), "injected kernel must be bit-identical"
                assert op.info["toep_source"] == "supplied"
        with mr.create_op(cfg, dS, dphi2, dtraj) as op:
            assert rel(host(op.kernel(0)), host(K)) < 1e-6
            y = op(dx)
            assert rel(host(y), host(ref)) < 1e-6
            inf = op.info
            assert inf["devices"] == [torch.cuda.current_device()]
            assert inf["toep_source"] == "broadcast" and inf["method"] == "toeplitz"
            assert inf["coil_count"] == [g.n_coils] and inf["group"][0] >= 1
            out = torch.empty_like(dx)
            assert op.apply(dx, out=out) is out and torch.equal(out, y)


def test_coil_group_invariance(mr, monkeypatch):
    """Coil groups of 1, 2 and all coils give the same result (the reference's
    coil-group invariance, tests/check_bd_kernels.py:13-17)."""
    g, traj, phi2, x, S = case("stack_40x36x10_L3_c3")
    cfg = cfg_for(mr, g)
    K = mr.calc_toep(cfg, dev(phi2), dev(traj))
    outs = []
    for grp in ("1", "2", "3"):
        monkeypatch.setenv("MR_AHA_HIP_COIL_GROUP", grp)
        outs.append(host(mr.aha_toeplitz(cfg, dev(x), dev(S), K)))
    assert rel(outs[0], outs[2]) < 1e-6 and rel(outs[1], outs[2]) < 1e-6


def _smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def _xbd_group(mr, cfg):
    """Coils per workgroup of the fused x pass a plan for cfg would use (0: separate passes)."""
    from mr_aha_amd import _lib
    plan = ctypes.c_void_p()
    _lib.check(_lib.lib.mr_aha_plan_create(ctypes.byref(cfg), ctypes.byref(plan)), "plan_create")
    try:
        g = _lib.lib.mr_aha_plan_xbd_group(plan)
        assert g >= 0
        return g
    finally:
        _lib.lib.mr_aha_plan_destroy(plan)


@pytest.mark.parametrize("name", ["2d_50x40_L3_c3", "3d_24x20x16_L4_c2", "3d_25x21x15_L3_c2_odd",
                                  "stack_40x36x10_L3_c3", "stack_31x36x9_L4_c2_odd",
                                  "radial_128_L4_C4", "2d_33x27_L1_c2_odd"])
@pytest.mark.parametrize("masked", [False, True])
def test_fused_and_separate_routes(mr, monkeypatch, name, masked):
    """The fused x pass (x-FFT, block multiply, inverse x-FFT per k-space row)
    and the separate FFT/block passes agree with the FP64 apply and with each
    other, for 1, 2 and all coils per fused workgroup, both triangles. An x
    grid length with a prime factor above 7 (66 = 2*3*11, 62 = 2*31) runs the
    separate passes whatever is asked."""
    from oracle import toeplitz
    print(f"[routes] {name} masked={masked}", flush=True)
    g, traj, phi2, x, S = case(name)
    dx, dS = dev(x), dev(S)
    for tri in ("lower", "upper"):
        cfg = cfg_for(mr, g, triangulation=tri, is_masked=masked)
        mask = mr.calc_mask(cfg, dev(traj)) if masked else None
        K = mr.calc_toep(cfg, dev(phi2), dev(traj), mask)
        ref = toeplitz.apply(g, x, S, host(K).astype(np.complex128), upper=(tri == "upper"),
                             mask=host(mask) if masked else None)
        outs = {}
        for route, G in (("separate", None), ("fused", "1"), ("fused", "2"), ("fused", "64")):
            monkeypatch.setenv("MR_AHA_HIP_XBD", "0" if route == "separate" else "1")
            if G is None:
                monkeypatch.delenv("MR_AHA_HIP_XBD_G", raising=False)
            else:
                monkeypatch.setenv("MR_AHA_HIP_XBD_G", G)
            got = _xbd_group(mr, cfg)
            if route == "separate":
                assert got == 0
            else:
                assert got == (min(int(G), g.n_coils) if _smooth(g.M[0]) else 0), (route, G, got)
            y = host(mr.aha_toeplitz(cfg, dx, dS, K, mask))
            assert rel(y, ref) < 1e-5, (route, G, tri)
            outs[(route, G)] = y
        base = outs[("separate", None)]
        for k, y in outs.items():
            assert rel(y, base) < 1e-6, (k, tri)


SPECIALISED = [("2D", (320, 12), 16, 2), ("2D", (160, 10), 8, 3), ("2D", (128, 8), 4, 5),
               ("2D", (192, 6), 16, 2), ("2D", (256, 6), 8, 3), ("stack", (320, 8, 3), 16, 2),
               ("3D", (160, 6, 4), 4, 3), ("2D", (200, 6), 8, 2),
               ("2D", (320, 10), 8, 2), ("2D", (192, 8), 8, 3),
               # coil pairs sharing kernel loads, with an odd last coil: x length
               # 320 (two coils in 80 KB) and 512 (1024-thread workgroups, 160 KB)
               ("2D", (160, 8), 16, 3), ("2D", (256, 8), 16, 3)]


@pytest.mark.parametrize("tt,n,L,C", SPECIALISED)
@pytest.mark.parametrize("masked", [False, True])
def test_fused_specialised_lengths(mr, monkeypatch, tt, n, L, C, masked):
    """x grid lengths 640, 320, 256, 384 and 512 with ranks 16, 8 and 4 run
    the fused kernels compiled for their length; 400 (the last case) has no
    such kernel and runs the run-time-length one. Random packed kernel and
    data, both triangles, against the FP64 apply, the run-time-length fused
    kernel and the separate passes."""
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[specialised] {tt} {n} L={L} C={C} masked={masked}", flush=True)
    rng = np.random.default_rng(7 * L + C + len(n))
    g = cv.make_geometry(tt, n, L, C)
    T = g.T
    grid = g.grid_shape()
    mask = None
    if masked:
        keep = rng.random(g.P) < 0.6
        mask = np.flatnonzero(keep).astype(np.int64)
        kshape = (T, mask.size)
    else:
        kshape = (T,) + grid
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape()) + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((C,) + g.img_shape()) + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    dx, dS, dK = dev(x), dev(S), dev(K)
    dmask = dev(mask) if masked else None
    for tri in ("lower", "upper"):
        cfg = cfg_for(mr, g, triangulation=tri, is_masked=masked,
                      num_masked=int(mask.size) if masked else 0)
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128), upper=(tri == "upper"), mask=mask)
        outs = []
        for xbd, ct in (("1", "1"), ("1", "0"), ("0", "1")):
            monkeypatch.setenv("MR_AHA_HIP_XBD", xbd)
            monkeypatch.setenv("MR_AHA_HIP_XBD_CT", ct)
            assert (_xbd_group(mr, cfg) > 0) == (xbd == "1")
            y = host(mr.aha_toeplitz(cfg, dx, dS, dK, dmask))
            assert rel(y, ref) < 1e-5, (tri, xbd, ct, rel(y, ref))
            outs.append(y)
        assert rel(outs[0], outs[2]) < 1e-6 and rel(outs[1], outs[2]) < 1e-6


CT_AXIS = [("stack", (20, 128, 56), 4, 2), ("3D", (20, 16, 160), 4, 2), ("2D", (24, 160), 8, 3),
           ("2D", (40, 192), 4, 2), ("stack", (18, 256, 8), 8, 2), ("stack", (21, 128, 5), 3, 1),
           ("stack", (21, 320, 4), 3, 2)]


@pytest.mark.parametrize("tt,n,L,C", CT_AXIS)
def test_fft_compile_time_lengths(mr, monkeypatch, tt, n, L, C):
    """y and z grid lengths 256, 56, 320, 384 and 512 run the axis passes
    compiled for their length (outer-axis pass with the sensitivity product,
    y pass with padding, inverse y pass with crop, inverse outer pass with the
    coil sum); against the FP64 apply and the run-time-length passes. The y
    passes run both in 16-pencil tiles (swizzled LDS order, twiddles from the
    global table; 1024 threads at y grid 512 and 640; (21, 128, 5) and
    (21, 320, 4) leave a partial last tile) and in the padded tiles of at most
    8 pencils."""
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[ct-axis] {tt} {n} L={L} C={C}", flush=True)
    rng = np.random.default_rng(11 * L + C)
    g = cv.make_geometry(tt, n, L, C)
    kshape = (g.T,) + g.grid_shape()
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape()) + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((C,) + g.img_shape()) + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    cfg = cfg_for(mr, g)
    ref = toeplitz.apply(g, x, S, K.astype(np.complex128))
    outs = []
    for ct, w16 in (("1", "1"), ("1", "0"), ("0", "1")):
        monkeypatch.setenv("MR_AHA_HIP_FFT_CT", ct)
        monkeypatch.setenv("MR_AHA_HIP_FFT_W16", w16)
        y = host(mr.aha_toeplitz(cfg, dev(x), dev(S), dev(K)))
        assert rel(y, ref) < 1e-5, (ct, w16, rel(y, ref))
        outs.append(y)
    assert rel(outs[0], outs[2]) < 1e-6 and rel(outs[1], outs[2]) < 1e-6


@pytest.mark.parametrize("L", [1, 2, 5, 7, 12, 13, 16, 17, 20])
def test_block_ranks(mr, L):
    """Every rank (specialised kernels up to 16, the any-rank kernel above)
    against the FP64 apply, both triangles, random packed kernel."""
    from oracle import conventions as cv
    from oracle import toeplitz
    rng = np.random.default_rng(100 + L)
    g = cv.make_geometry("2D", (12, 10), L, 3)
    T = g.T
    K = (rng.standard_normal((T,) + g.grid_shape()) + 1j * rng.standard_normal((T,) + g.grid_shape())).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape()) + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((3,) + g.img_shape()) + 1j * rng.standard_normal((3,) + g.img_shape())).astype(np.complex64)
    for tri in ("lower", "upper"):
        cfg = cfg_for(mr, g, triangulation=tri)
        y = host(mr.aha_toeplitz(cfg, dev(x), dev(S), dev(K)))
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128), upper=(tri == "upper"))
        assert rel(y, ref) < 1e-5, (L, tri)


def test_calc_toep_from_psf(mr):
    """Two-stage build: the exact sign -1 point-spread functions finished on the
    GPU equal the one-shot kernel."""
    from oracle import conventions as cv
    from oracle import toeplitz
    g, traj, phi2, x, S = case("stack_31x36x9_L4_c2_odd")
    cfg = cfg_for(mr, g)
    K = toeplitz.calc_kernel(g, phi2, traj)
    # psf_t[delta] = prod M_tr * (centered forward DFT of phi2_toep) restated:
    Mx, My = g.M[0], g.M[1]
    Kc = np.fft.ifftshift(K, axes=(-2, -1))
    psf = np.fft.fftshift(np.fft.fft2(Kc), axes=(-2, -1))
    out = torch.zeros((g.T,) + g.grid_shape(), dtype=torch.complex64, device="cuda")
    for t0 in range(0, g.T, 3):
        n = min(3, g.T - t0)
        mr.calc_toep_from_psf(cfg, dev(psf[t0:t0 + n].astype(np.complex64)), t0, out)
    assert rel(host(out), K) < 1e-5


def test_argument_errors(mr, monkeypatch):
    g, traj, phi2, x, S = case("checkop_2d")
    dx, dS, dphi2, dtraj = dev(x), dev(S), dev(phi2), dev(traj)
    cfg = cfg_for(mr, g)
    with pytest.raises(ValueError, match="method"):
        mr.create_op(cfg, dS, dphi2, dtraj, method="dft")
    cfgm = cfg_for(mr, g, is_masked=True)
    mask = mr.calc_mask(cfgm, dtraj)
    with pytest.raises(ValueError, match="mask"):
        mr.create_op(cfgm, dS, dphi2, dtraj, mask, method="nufft")
    with pytest.raises(ValueError, match="mask"):
        mr.create_op(cfgm, dS, dphi2, dtraj)
    with pytest.raises(ValueError, match="nufft"):
        mr.create_op(cfg, dS, dphi2, dtraj, nufft_opts={"upsampfac": 2})
    with pytest.raises(ValueError, match="nufft_opts"):
        mr.create_op(cfg, dS, dphi2, dtraj, method="nufft", nufft_opts={"spread": 1})
    with pytest.raises(ValueError, match="devices"):
        mr.create_op(cfg, dS, dphi2, dtraj, devices="some")
    with pytest.raises(mr.MRError, match="not visible"):
        mr.create_op(cfg, dS, dphi2, dtraj, devices=[99])
    with pytest.raises(ValueError, match="toep_dist"):
        mr.create_op(cfg, dS, dphi2, dtraj, toep_dist="ring")
    with pytest.raises(ValueError, match="SEs"):
        mr.create_op(cfg, dS[:1], dphi2, dtraj)
    with pytest.raises(TypeError, match="create_op"):
        mr.AhAOp(cfg)
    with mr.create_op(cfg, dS, dphi2, dtraj, devices=np.int64(torch.cuda.current_device())) as op:
        assert op.info["devices"] == [torch.cuda.current_device()]
    with mr.create_op(cfg, dS, dphi2, dtraj, devices="all") as op:
        assert len(op.info["devices"]) == 1
    with mr.create_op(cfg, dS, dphi2, dtraj) as op:
        L0 = cfg.L
        cfg.L = L0 + 1
        try:
            assert op.config.L == L0 and op(dx).shape == dx.shape
        finally:
            cfg.L = L0
    monkeypatch.setenv("MR_AHA_TOEP_SPLIT", "bogus")
    with pytest.raises(mr.MRError, match="MR_AHA_TOEP_SPLIT"):
        mr.create_op(cfg, dS, dphi2, dtraj)
    monkeypatch.delenv("MR_AHA_TOEP_SPLIT")
    op = mr.create_op(cfg, dS, dphi2, dtraj)
    op.close()
    op.close()
    with pytest.raises(RuntimeError, match="closed"):
        op(dx)
    with pytest.raises(ValueError, match="x"):
        mr.create_op(cfg, dS, dphi2, dtraj)(dx[:1])
    K = mr.calc_toep(cfg, dphi2, dtraj)
    with pytest.raises(TypeError, match="complex64"):
        mr.aha_toeplitz(cfg, dx.to(torch.complex128), dS, K)
    assert torch.cuda.current_device() == 0


def test_no_leak(mr):
    g, traj, phi2, x, S = case("checkop_2d")
    dx, dS, dphi2, dtraj = dev(x), dev(S), dev(phi2), dev(traj)
    cfg = cfg_for(mr, g)
    with mr.create_op(cfg, dS, dphi2, dtraj) as op:
        op(dx)
    torch.cuda.synchronize()
    before = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        with mr.create_op(cfg, dS, dphi2, dtraj) as op:
            op(dx)
    torch.cuda.synchronize()
    after = torch.cuda.mem_get_info()[0]
    assert before - after <= 8 << 20


@pytest.mark.parametrize("name", ["2d_50x40_L3_c3", "2d_33x27_L1_c2_odd", "3d_25x21x15_L3_c2_odd",
                                  "stack_31x36x9_L4_c2_odd", "checkop_stack"])
def test_nufft_method(mr, golden, name):
    """method='nufft' against the exact NUDFT normal operator (the reference's
    library-vs-Python NUFFT tolerance 1e-5) and against the Toeplitz method
    (1e-4, tests/check_op.py:92-94)."""
    print(f"[nufft] {name}", flush=True)
    gd = golden(name)
    g, traj, phi2, x, S = case(name)
    cfg = cfg_for(mr, g)
    dx, dS, dphi2, dtraj = dev(x), dev(S), dev(phi2), dev(traj)
    for share in (True, False):
        with mr.create_op(cfg, dS, dphi2, dtraj, method="nufft", share_primary=share) as op:
            y = host(op(dx))
            assert rel(y, gd["AhAx_nudft"]) < 1e-5
            inf = op.info
            assert inf["toep_source"] == "none" and inf["method"] == "nufft" and inf["group"] == [1]
            with pytest.raises(mr.MRError, match="no kernel"):
                op.kernel(0)
    K = mr.calc_toep(cfg, dphi2, dtraj)
    assert rel(y, host(mr.aha_toeplitz(cfg, dx, dS, K))) < 1e-4
    with mr.create_op(cfg, dS, dphi2, dtraj, method="nufft",
                      nufft_opts={"upsampfac": 2.0, "gpu_method": 1}) as op:
        assert rel(host(op(dx)), gd["AhAx_nudft"]) < 1e-5


# fused x pass variants: pencil I/O one point per lane (x length 2 mod 4) or
# two (0 mod 4); kernel entries two per 16-byte load (16-byte aligned kernel)
# or one per 8-byte load (kernel pointer offset by 8 bytes); at x length 640
# in the 1024-thread two-coil form, the last 128 locations of an aligned
# coil pair walked per coil
FUSED_PATHS = [((126, 10), (256, 20), 16, 2, False), ((250, 8), (640, 16), 16, 2, False),
               ((160, 8), (320, 16), 16, 3, True), ((320, 12), (640, 24), 16, 2, True),
               ((96, 12), (384, 24), 8, 3, True), ((320, 12), (640, 24), 16, 1, False),
               # x length 640 in coil pairs with an odd last coil: the pair group
               # walks its last 128 locations per coil, the lone coil does not
               ((320, 12), (640, 24), 16, 3, False)]


@pytest.mark.parametrize("wide", [True, False])
@pytest.mark.parametrize("n,M,L,C,offset", FUSED_PATHS)
def test_fused_pair_paths(mr, monkeypatch, n, M, L, C, offset, wide):
    """wide=False keeps one coil per workgroup (the 16-byte entry loads then
    serve one coil's walk)."""
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[fused paths] n={n} M={M} L={L} C={C} offset={offset} wide={wide}", flush=True)
    if not wide:
        monkeypatch.setenv("MR_AHA_HIP_XBD_WIDE", "0")
    rng = np.random.default_rng(sum(n) + L + C)
    g = cv.make_geometry("2D", n, L, C, phi2_dim=M)
    kshape = (g.T,) + g.grid_shape()
    K = (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape)).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape())
         + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((C,) + g.img_shape())
         + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    if offset:
        buf = torch.empty(K.size + 1, dtype=torch.complex64, device="cuda")
        dK = buf[1:].view(kshape)
        dK.copy_(dev(K))
        assert dK.data_ptr() % 16 == 8
    else:
        dK = dev(K)
    for tri in ("lower", "upper"):
        cfg = cfg_for(mr, g, triangulation=tri)
        assert _xbd_group(mr, cfg) > 0
        ref = toeplitz.apply(g, x, S, K.astype(np.complex128), upper=(tri == "upper"))
        y = host(mr.aha_toeplitz(cfg, dev(x), dev(S), dK, None))
        assert rel(y, ref) < 1e-5, (tri, rel(y, ref))
