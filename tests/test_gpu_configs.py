"""The benchmark configurations under test, at reduced and at full size.

Reduced size (committed fixtures from the FP64 oracle, tests/golden/):
  stack_40x40x8_L16_C12_T200   the 320x320x56 stack-of-stars recipe (L=16,
                               C=12, 200-frame basis round-robin over
                               golden-angle spokes) at 40x40x8
  koosh_24_L8_C8_T100          the 160^3 kooshball recipe (L=8, C=8,
                               Fibonacci-sphere spokes, 100-frame basis)
  checked as the golden cases are: kernel vs the exact Dirichlet-sum kernel
  (1e-5), apply vs the FP64 apply on the same kernel (1e-5), end to end vs
  the exact Toeplitz result (1e-5) and the direct NUDFT (1e-4).

Full size (the benchmark's own inputs, generated on the device): properties
that do not need an FP64 apply at that size,
  * sampled kernel entries vs the exact Dirichlet sum (oracle.toeplitz.
    kernel_entries), relative 2-norm 1e-5;
  * Hermitian <u, AhA x> = <AhA u, x> and <x, AhA x> real and > 0 (1e-5);
  * the fused x pass equal to the separate-pass route (2e-6; both FP32 with
    different summation orders).

Kernel-build chunking: the components of one kernel are gridded in chunks
sized from free memory; MR_AHA_TOEP_MAX_CHUNK caps the chunk, and every cap
(one component at a time, sizes that do not divide the component count)
must give the single-chunk kernel bit for bit, masked and unmasked.
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


def cfg_for(mr, g, **kw):
    n = g.n[: g.n_spatial]
    M = g.M[: g.n_spatial]
    return mr.make_config(g.traj_type, n, g.L, phi2_dim=M, n_coils=g.n_coils, **kw)


def case(name):
    from tests.golden.make_golden import case_inputs
    return case_inputs(name)


RECIPES = ["stack_40x40x8_L16_C12_T200", "koosh_24_L8_C8_T100"]


@pytest.mark.parametrize("name", RECIPES)
@pytest.mark.parametrize("route", ["fused", "separate"])
def test_recipe_fixture_vs_oracle(mr, golden, monkeypatch, name, route):
    from oracle import toeplitz
    print(f"[recipe] {name} {route}", flush=True)
    monkeypatch.setenv("MR_AHA_HIP_XBD", "1" if route == "fused" else "0")
    gd = golden(name)
    g, traj, phi2, x, S = case(name)
    for tri in ("lower", "upper"):
        cfg = cfg_for(mr, g, triangulation=tri)
        K = mr.calc_toep(cfg, dev(phi2), dev(traj))
        Kh = host(K)
        e_k = rel(Kh.reshape(g.T, -1)[:, gd["kidx"]], gd[f"K_{tri}"])
        assert e_k < 1e-5, (tri, e_k)
        with mr.create_op(cfg, dev(S), dev(phi2), dev(traj), _phi2_toep=K) as op:
            assert (op.xbd_group > 0) == (route == "fused")
            y = host(op(dev(x)))
        e_a = rel(y, toeplitz.apply(g, x, S, Kh.astype(np.complex128), upper=(tri == "upper")))
        assert e_a < 1e-5, (tri, e_a)
        assert rel(y, gd[f"AhAx_{tri}"]) < 1e-5
        assert rel(y, gd["AhAx_nudft"]) < 1e-4


@pytest.mark.parametrize("masked", [False, True])
def test_calc_toep_chunking_bit_identical(mr, golden, monkeypatch, masked):
    """Chunk caps of 1, 3, 7 and 64 components (1088 components at L=16 over
    8 kz planes; 3 and 7 do not divide it) against one chunk."""
    name = "stack_40x40x8_L16_C12_T200"
    gd = golden(name)
    g, traj, phi2, x, S = case(name)
    dphi2, dtraj = dev(phi2), dev(traj)
    cfg = cfg_for(mr, g)
    mask = None
    if masked:
        cfg = cfg_for(mr, g, is_masked=True)
        mask = mr.calc_mask(cfg, dtraj)
    monkeypatch.delenv("MR_AHA_TOEP_MAX_CHUNK", raising=False)
    ref = mr.calc_toep(cfg, dphi2, dtraj, mask)
    if not masked:
        assert rel(host(ref).reshape(g.T, -1)[:, gd["kidx"]], gd["K_lower"]) < 1e-5
    for cap in ("1", "3", "7", "64"):
        print(f"[chunk] masked={masked} cap={cap}", flush=True)
        monkeypatch.setenv("MR_AHA_TOEP_MAX_CHUNK", cap)
        K = mr.calc_toep(cfg, dphi2, dtraj, mask)
        assert torch.equal(K, ref), cap


def test_tile_width_not_power_of_two(mr, monkeypatch):
    """An axis-pass width that is not a power of two (24 pencils) is rounded
    down to one (16): the result equals the default tiling bit for bit."""
    g, traj, phi2, x, S = case("stack_40x40x8_L16_C12_T200")
    cfg = cfg_for(mr, g)
    K = mr.calc_toep(cfg, dev(phi2), dev(traj))
    dx, dS = dev(x), dev(S)
    ref = mr.aha_toeplitz(cfg, dx, dS, K)
    for knob in ("MR_AHA_HIP_FO_TX", "MR_AHA_HIP_FY_TX", "MR_AHA_HIP_IY_TX", "MR_AHA_HIP_IO_TX",
                 "MR_AHA_HIP_FFT_TXMAX"):
        monkeypatch.setenv(knob, "24")
        assert torch.equal(mr.aha_toeplitz(cfg, dx, dS, K), ref), knob
        monkeypatch.delenv(knob)


# ---------------------------------------------------------------------------
# Full size
# ---------------------------------------------------------------------------

def _sample_entries(shape, rng, n_per_axis=8, n_comp=8):
    """(t, [kz,] ky, kx) centered indices on a few distinct in-plane (3D:
    spatial) coordinates, always including the kernel's centre."""
    T = shape[0]
    sp = shape[1:]
    picks = []
    for a, M in enumerate(sp):
        v = rng.choice(M, size=n_per_axis - 1, replace=False)
        picks.append(np.unique(np.concatenate([[M // 2], v])))
    sel = []
    if len(sp) == 3 and n_comp > 0 and shape[1] != shape[2]:
        # stack: (t, kz) pairs drawn per in-plane location
        for ky in picks[1]:
            for kx in picks[2]:
                for _ in range(n_comp):
                    sel.append((rng.integers(T), rng.integers(sp[0]), ky, kx))
    elif len(sp) == 3:
        # 3D: the centre (the largest entries, as in the 2D/stack cases, where
        # the centre row and column are always drawn) and a few spatial
        # triples, each with several components
        for t in rng.choice(T, size=n_comp, replace=False):
            sel.append((t, sp[0] // 2, sp[1] // 2, sp[2] // 2))
        for i in range(n_per_axis):
            kz, ky, kx = picks[0][i % len(picks[0])], picks[1][(3 * i) % len(picks[1])], \
                picks[2][(5 * i) % len(picks[2])]
            for t in rng.choice(T, size=n_comp, replace=False):
                sel.append((t, kz, ky, kx))
    else:
        for ky in picks[0]:
            for kx in picks[1]:
                sel.append((rng.integers(T), ky, kx))
    return np.array(sel, dtype=np.int64)


def _geometry(w):
    from oracle.conventions import make_geometry
    return make_geometry(w["traj_type"], w["n_dim"], w["L"], w["n_coils"])


def _full_size_checks(mr, monkeypatch, w, label):
    from oracle import toeplitz
    g = _geometry(w)
    cfg = mr.make_config(w["traj_type"], w["n_dim"], w["L"], n_coils=w["n_coils"])
    print(f"[full] {label}: kernel build", flush=True)
    K = mr.calc_toep(cfg, w["phi2"], w["traj"])
    torch.cuda.synchronize()
    # sampled kernel entries vs the exact Dirichlet sum
    rng = np.random.default_rng(99)
    sel = _sample_entries(tuple(K.shape), rng)
    idx = torch.from_numpy(sel).cuda()
    got = host(K[tuple(idx[:, i] for i in range(idx.shape[1]))])
    exact = toeplitz.kernel_entries(g, host(w["phi2"]), host(w["traj"]), sel)
    e_k = rel(got, exact)
    # off-centre entries are orders of magnitude below the centre; their
    # error is bounded against the largest sampled entry
    e_max = float(np.max(np.abs(got - exact)) / np.max(np.abs(exact)))
    print(f"[full] {label}: {len(sel)} sampled kernel entries, rel {e_k:.2e}, "
          f"max abs / max |K| {e_max:.2e}", flush=True)
    assert e_k < 1e-5, e_k
    assert e_max < 1e-5, e_max

    x = w["x"]
    gen = torch.Generator(device="cuda").manual_seed(7)
    u = torch.complex(torch.randn(x.shape, generator=gen, device="cuda"),
                      torch.randn(x.shape, generator=gen, device="cuda"))
    monkeypatch.setenv("MR_AHA_HIP_XBD", "1")
    with mr.create_op(cfg, w["SEs"], w["phi2"], w["traj"], _phi2_toep=K) as op:
        assert op.xbd_group > 0
        yx = op(x)
        yu = op(u)
        torch.cuda.synchronize()

    def dot(a, b):
        return complex(torch.sum(torch.conj(a).to(torch.complex128) * b.to(torch.complex128)))

    lhs, rhs = dot(u, yx), dot(yu, x)
    e_h = abs(lhs - rhs) / abs(lhs)
    q = dot(x, yx)
    print(f"[full] {label}: hermitian {e_h:.2e}, <x,AhAx> = {q.real:.4e} + {q.imag:.2e}i",
          flush=True)
    assert e_h < 1e-5
    assert q.real > 0 and abs(q.imag) < 1e-5 * q.real
    del yu, u
    torch.cuda.empty_cache()
    # fused x pass == separate passes on the same kernel
    monkeypatch.setenv("MR_AHA_HIP_XBD", "0")
    ysep = mr.aha_toeplitz(cfg, x, w["SEs"], K)
    e_r = rel(host(yx), host(ysep))
    print(f"[full] {label}: fused vs separate {e_r:.2e}", flush=True)
    assert e_r < 2e-6
    monkeypatch.setenv("MR_AHA_HIP_XBD", "1")
    del K, ysep, yx
    torch.cuda.empty_cache()


def test_stack_of_stars_full_size(mr, monkeypatch):
    """320x320x56, grid 640x640x56, L=16, C=12, 400 spokes x 640, 200 frames."""
    from mr_aha_amd import workloads
    w = workloads.stack_of_stars()
    _full_size_checks(mr, monkeypatch, w, "stack 320x320x56 L16 C12")


def test_kooshball_full_size(mr, monkeypatch):
    """160^3, grid 320^3, L=8, C=8, 20000 Fibonacci spokes x 320, 100 frames."""
    from mr_aha_amd import workloads
    w = workloads.kooshball()
    _full_size_checks(mr, monkeypatch, w, "kooshball 160^3 L8 C8")
