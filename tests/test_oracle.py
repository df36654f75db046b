"""Pins of the FP64 oracle (CPU only).

Every oracle function is checked against something that does not restate
its own formula: literal fixtures of the reference, closed forms, an
independent formulation (Toeplitz embedding vs direct NUDFT vs an explicit
dense forward model), invariants, and brute force on tiny inputs. Planted
mistakes in the oracle (transposed Gram block, wrong sign, conjugated
kernel) are shown to fail these checks.
"""

import numpy as np
import pytest

from conftest import rel
from oracle import conventions as cv
from oracle import generators as gen
from oracle import mask as omask
from oracle import nudft, toeplitz


def test_known_answers():
    assert cv.num_tri(16) == 136 and cv.num_tri(0) == 0 and cv.num_tri(1) == 1
    g = cv.make_geometry("stack", (320, 320, 56), 16, 12)
    assert g.M == (640, 640, 56)
    assert g.P * g.T == 3_119_513_600  # mr_phi2_toep_size of the 320x320x56, L=16 stack
    g1 = cv.make_geometry("2D", (128, 128), 4, 4)
    assert g1.P * g1.T == 655_360


def test_tri_pairs_literal_fixture():
    # python/mr_aha/_psf.py:104-105 of the reference
    assert cv.tri_pairs(3, False) == [(0, 0), (1, 0), (2, 0), (1, 1), (2, 1), (2, 2)]
    assert cv.tri_pairs(3, True) == [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]


@pytest.mark.parametrize("L", [1, 2, 3, 5, 8, 16, 33])
@pytest.mark.parametrize("upper", [False, True])
def test_tri_offset_closed_form(L, upper):
    pairs = cv.tri_pairs(L, upper)
    for t, (i, j) in enumerate(pairs):
        # include/mr_aha/mr_aha.h:307-309 closed forms
        expect = j * (j + 1) // 2 + i if upper else j * L - j * (j - 1) // 2 + (i - j)
        assert expect == t
        assert cv.tri_offset(i, j, L, upper) == (t, False)
        if i != j:
            assert cv.tri_offset(j, i, L, upper) == (t, True)


@pytest.mark.parametrize("M", [1, 2, 5, 8, 101, 640])
def test_centering_maps(M):
    k = np.arange(M)
    j = cv.cent2nat(k, M)
    assert sorted(j.tolist()) == list(range(M))
    assert np.array_equal(cv.nat2cent(j, M), k)
    # centered index floor(M/2) is frequency 0 -> natural 0
    assert cv.cent2nat(M // 2, M) == 0
    for N in range(1, M + 1):
        p = cv.pad_index(np.arange(N), N, M)
        assert len(set(p.tolist())) == N


def test_centering_literals():
    """Hand-worked values of the maps of src/cuda/mr_internal.cuh:119-139 and
    src/cuda/mr_kernels.cu:111-128: pad j = (x - floor(N/2)) mod M,
    centered k = (j + floor(M/2)) mod M."""
    assert cv.pad_index(0, 5, 8) == 6 and cv.pad_index(2, 5, 8) == 0 and cv.pad_index(4, 5, 8) == 2
    assert cv.pad_index(0, 4, 7) == 5 and cv.pad_index(3, 4, 7) == 1
    assert cv.nat2cent(0, 8) == 4 and cv.nat2cent(7, 8) == 3 and cv.nat2cent(0, 7) == 3
    assert cv.cent2nat(0, 8) == 4 and cv.cent2nat(3, 7) == 0 and cv.cent2nat(6, 7) == 3


def test_cpu_baseline_matches_apply():
    """The timed CPU baseline computes the oracle's apply (complex64 rounding)."""
    from oracle import cpu_baseline as cb
    for name in ("2d_50x40_L3_c3", "stack_31x36x9_L4_c2_odd", "3d_25x21x15_L3_c2_odd"):
        g, traj, phi2, x, S = gen.reference_case(name)
        for up in (False, True):
            K = toeplitz.calc_kernel(g, phi2, traj, upper=up)
            ref = toeplitz.apply(g, x, S, K, upper=up)
            y = cb.toeplitz_apply_coils(list(g.n), list(g.M), g.L, x.astype(np.complex64),
                                        S.astype(np.complex64), K.astype(np.complex64), up,
                                        range(g.n_coils), 2)
            assert rel(y, ref) < 1e-5, (name, up)


@pytest.mark.parametrize("case", ["2d_50x40_L3_c3", "2d_33x27_L1_c2_odd", "3d_24x20x16_L4_c2",
                                  "3d_25x21x15_L3_c2_odd", "stack_40x36x10_L3_c3",
                                  "stack_31x36x9_L4_c2_odd"])
def test_cpu_nufft_matches_nudft(case):
    """The CPU NUFFT A^H A (ES kernel, upsampling 2; the benchmark's CPU
    baseline) against the exact direct NUDFT normal operator: complex128 at
    eps 1e-12 to 1e-10, complex64 at eps 1e-6 (the reference's default) to
    1e-5; odd grids, 2D, 3D and stack."""
    from oracle import cpu_nufft
    g, traj, phi2, x, S = gen.reference_case(case)
    ref = nudft.aha_nudft(g, x, S, phi2, traj)
    y64 = cpu_nufft.aha_nufft(g, x, S, phi2, traj, eps=1e-12, dtype=np.complex128, threads=4)
    y32 = cpu_nufft.aha_nufft(g, x, S, phi2, traj, eps=1e-6, dtype=np.complex64, threads=4)
    assert rel(y64, ref) < 1e-10, rel(y64, ref)
    assert rel(y32, ref) < 1e-5, rel(y32, ref)


def test_cpu_nufft_planted_mistakes_fail():
    """A conjugated Gram block, a flipped trajectory, a dropped kernel
    deconvolution each move the CPU NUFFT far from the exact operator."""
    from oracle import cpu_nufft
    g, traj, phi2, x, S = gen.reference_case("stack_31x36x9_L4_c2_odd")
    ref = nudft.aha_nudft(g, x, S, phi2, traj)
    kw = dict(eps=1e-9, dtype=np.complex128, threads=2)
    assert rel(cpu_nufft.aha_nufft(g, x, S, np.conj(phi2), traj, **kw), ref) > 1e-2
    assert rel(cpu_nufft.aha_nufft(g, x, S, phi2, -traj, **kw), ref) > 1e-2
    orig = cpu_nufft.es_fourier
    try:
        cpu_nufft.es_fourier = lambda w, beta, nf, N: np.ones(N) * orig(w, beta, nf, N)[N // 2]
        assert rel(cpu_nufft.aha_nufft(g, x, S, phi2, traj, **kw), ref) > 1e-2
    finally:
        cpu_nufft.es_fourier = orig


def test_cpu_nufft_plane_sample_is_partial_sum():
    """A stack run restricted to kz planes (the baseline's timing sample) sums
    over the planes to the full result."""
    from oracle import cpu_nufft
    g, traj, phi2, x, S = gen.reference_case("stack_40x36x10_L3_c3")
    kw = dict(eps=1e-9, dtype=np.complex128, threads=2)
    full = cpu_nufft.aha_nufft(g, x, S, phi2, traj, **kw)
    nz = g.n[2]
    parts = cpu_nufft.aha_nufft(g, x, S, phi2, traj, planes=range(0, nz // 2), **kw) + \
        cpu_nufft.aha_nufft(g, x, S, phi2, traj, planes=range(nz // 2, nz), **kw)
    assert rel(parts, full) < 1e-12


@pytest.mark.parametrize("M", [4, 7, 64])
def test_dirichlet_closed_form(M):
    # D_M(t) = sum_{d=-floor(M/2)}^{ceil(M/2)-1} e^{i d t}
    #        = e^{-i t (M - 2 ceil(M/2) + 1)/2} sin(M t/2) / sin(t/2)
    rng = np.random.default_rng(0)
    x = rng.uniform(-np.pi, np.pi, 50)
    D = toeplitz.dirichlet_matrix(x, M)
    kappa = np.arange(M) - M // 2
    th = 2 * np.pi * kappa[None, :] / M - x[:, None]
    shift = (M - 2 * ((M + 1) // 2) + 1) / 2.0
    closed = np.exp(-1j * th * shift) * np.sin(M * th / 2) / np.sin(th / 2)
    assert rel(D, closed) < 1e-12


def _small(tt, n, L, C, ntr, nkx, seed):
    rng = np.random.default_rng(seed)
    g = cv.make_geometry(tt, n, L, C)
    traj = rng.uniform(-np.pi, np.pi, (g.ndim_traj, nkx, ntr)).astype(np.float32)
    phi2 = gen.make_phi2(rng, L, ntr, g.nkz if tt == "stack" else None)
    x = gen.crandn(rng, (L,) + g.img_shape())
    S = gen.crandn(rng, (C,) + g.img_shape())
    return g, traj, phi2, x, S


SMALL = [("2D", (12, 10), 3, 2, 7, 9, 1), ("2D", (9, 7), 2, 2, 5, 6, 2),
         ("3D", (6, 5, 4), 2, 2, 9, 7, 3), ("3D", (5, 5, 3), 3, 1, 6, 5, 4),
         ("stack", (8, 6, 5), 3, 2, 6, 8, 5), ("stack", (7, 6, 4), 2, 2, 5, 6, 6)]


@pytest.mark.parametrize("case", SMALL)
def test_toeplitz_equals_nudft(case):
    """Toeplitz embedding (kernel + 5 stages) == exact direct NUDFT, both triangles."""
    g, traj, phi2, x, S = _small(*case)
    ref = nudft.aha_nudft(g, x, S, phi2, traj)
    for up in (False, True):
        K = toeplitz.calc_kernel(g, phi2, traj, upper=up)
        assert rel(toeplitz.apply(g, x, S, K, upper=up), ref) < 1e-12


@pytest.mark.parametrize("case", SMALL)
def test_kernel_entries_match_full_kernel(case):
    """Sampled exact entries == the full exact kernel at every entry, with a
    sample planted at the origin (a removable singularity of the closed form)."""
    g, traj, phi2, x, S = _small(*case)
    traj = traj.copy()
    traj[:, 0, 0] = 0.0
    for up in (False, True):
        K = toeplitz.calc_kernel(g, phi2, traj, upper=up)
        idx = np.array(np.unravel_index(np.arange(K.size), K.shape)).T
        E = toeplitz.kernel_entries(g, phi2, traj, idx, upper=up, chunk=13)
        assert rel(E, K.ravel()) < 1e-12
    # a conjugated Gram block moves the entries (the check is not vacuous)
    E = toeplitz.kernel_entries(g, np.conj(phi2), traj, idx)
    assert rel(E, toeplitz.calc_kernel(g, phi2, traj).ravel()) > 1e-3


def test_planted_mistakes_fail():
    """A conjugated kernel, a transposed Gram block or a flipped NUFFT sign
    must all be caught by the Toeplitz-vs-NUDFT check."""
    g, traj, phi2, x, S = _small("2D", (10, 8), 3, 2, 6, 7, 11)
    ref = nudft.aha_nudft(g, x, S, phi2, traj)
    K = toeplitz.calc_kernel(g, phi2, traj)
    assert rel(toeplitz.apply(g, x, S, np.conj(K)), ref) > 1e-2
    phi2_t = np.ascontiguousarray(np.swapaxes(phi2, 0, 1))
    assert rel(nudft.aha_nudft(g, x, S, phi2_t, traj), ref) > 1e-2
    assert rel(nudft.aha_nudft(g, x, S, phi2, -traj), ref) > 1e-2


@pytest.mark.parametrize("tt,n", [("2D", (6, 5)), ("3D", (4, 3, 3))])
def test_dense_forward_model(tt, n):
    """Oracle NUDFT == explicit A^H A with A_c[m,(l,r)] = Phi(l,t_m) e^{-i r.x_m} S_c(r)
    (data model of include/mr_aha/mr_aha.h:48-60), rank-1 blocks Phi Phi^H."""
    rng = np.random.default_rng(3)
    L, C, ntr, nkx = 3, 2, 5, 6
    g = cv.make_geometry(tt, n, L, C)
    traj = rng.uniform(-np.pi, np.pi, (g.ndim_traj, nkx, ntr)).astype(np.float32)
    Phi = gen.orthonormal_basis(rng, 4, L)
    frames = np.arange(ntr) % 4
    p = Phi[:, frames]
    B = p[:, None] * np.conj(p[None, :])
    phi2 = np.ascontiguousarray(np.swapaxes(B, 0, 1))  # complex128 (no rounding)
    x = gen.crandn(rng, (L,) + g.img_shape())
    S = gen.crandn(rng, (C,) + g.img_shape())
    d = nudft.dense_aha(g, x, S, Phi, frames, traj)
    assert rel(nudft.aha_nudft(g, x, S, phi2, traj), d) < 1e-12
    K = toeplitz.calc_kernel(g, phi2, traj)
    assert rel(toeplitz.apply(g, x, S, K), d) < 1e-12


@pytest.mark.parametrize("tt,n", [("2D", (8, 6)), ("3D", (4, 4, 3)), ("stack", (6, 4, 3))])
def test_full_cartesian_closed_form(tt, n):
    """Sampling every Cartesian point of the n-grid with a constant Hermitian
    block B gives AhAx_l = (prod n) sum_c |S_c|^2 sum_j x_j B(j, l)."""
    rng = np.random.default_rng(4)
    L, C = 3, 2
    g = cv.make_geometry(tt, n, L, C)
    nd = g.ndim_traj
    axes = [2 * np.pi * (np.arange(g.n[a]) - g.n[a] // 2) / g.n[a] for a in range(nd)]
    grids = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([gr.ravel() for gr in grids])       # (nd, Npts)
    traj = pts[:, None, :].astype(np.float32)            # one sample per readout
    ntr = pts.shape[1]
    A = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L))
    Bm = A @ A.conj().T                                   # Hermitian, not real
    nkz = g.nkz if tt == "stack" else None
    lead = (ntr,) if nkz is None else (nkz, ntr)
    phi2 = np.broadcast_to(Bm.T.reshape(L, L, *([1] * len(lead))), (L, L) + lead).copy()
    x = gen.crandn(rng, (L,) + g.img_shape()).astype(np.complex128)
    S = gen.crandn(rng, (C,) + g.img_shape()).astype(np.complex128)
    w = np.sum(np.abs(S) ** 2, axis=0)
    xB = np.einsum("j...,jl->l...", x, Bm)
    expect = np.prod(g.n) * w[None] * xB
    K = toeplitz.calc_kernel(g, phi2, traj)
    # float32 trajectory: samples sit on the grid to ~1e-7 relative
    assert rel(toeplitz.apply(g, x, S, K), expect) < 1e-6
    assert rel(nudft.aha_nudft(g, x, S, phi2, traj), expect) < 1e-6
    # orientation matters: x B^T is a different answer
    assert rel(np.prod(g.n) * w[None] * np.einsum("j...,lj->l...", x, Bm), expect) > 1e-2


def test_hermitian_psd_and_triangles():
    g, traj, phi2, x, S = _small("3D", (6, 5, 4), 3, 2, 9, 7, 8)
    rng = np.random.default_rng(9)
    y = gen.crandn(rng, x.shape)
    K = toeplitz.calc_kernel(g, phi2, traj)
    Ku = toeplitz.calc_kernel(g, phi2, traj, upper=True)
    Ax = toeplitz.apply(g, x, S, K)
    Ay = toeplitz.apply(g, y, S, K)
    lhs = np.vdot(y, Ax)
    rhs = np.vdot(Ay, x)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    q = np.vdot(x, Ax)
    assert q.real > 0 and abs(q.imag) < 1e-10 * abs(q)
    assert rel(toeplitz.apply(g, x, S, Ku, upper=True), Ax) < 1e-12


def test_l1_unit_gram_is_psf_convolution():
    """L=1, Phi2=1, one coil with S=1: A^H A x is the convolution with the
    point-spread function, written as an explicit matrix."""
    rng = np.random.default_rng(5)
    g = cv.make_geometry("2D", (7, 6), 1, 1)
    ntr, nkx = 4, 9
    traj = rng.uniform(-np.pi, np.pi, (2, nkx, ntr)).astype(np.float32)
    phi2 = np.ones((1, 1, ntr), np.complex64)
    x = gen.crandn(rng, (1, 6, 7))
    S = np.ones((1, 6, 7), np.complex64)
    ry, rx = np.meshgrid(np.arange(6) - 3, np.arange(7) - 3, indexing="ij")
    r = np.stack([rx.ravel(), ry.ravel()])
    pts = traj.reshape(2, -1).astype(np.float64)
    E = np.exp(1j * (r.T @ pts))                      # e^{i r.x_m}
    G = E @ E.conj().T                                # sum_m e^{i (r - r').x_m}
    expect = (G @ x.ravel().astype(np.complex128)).reshape(1, 6, 7)
    K = toeplitz.calc_kernel(g, phi2, traj)
    assert rel(toeplitz.apply(g, x, S, K), expect) < 1e-12


def test_masked_kernel_is_gathered_and_full_mask_is_identity():
    g, traj, phi2, x, S = _small("2D", (10, 8), 2, 2, 6, 7, 12)
    K = toeplitz.calc_kernel(g, phi2, traj)
    m = omask.calc_mask(g, traj, 2)
    Km = toeplitz.calc_kernel(g, phi2, traj, mask=m)
    assert np.array_equal(Km, K.reshape(g.T, -1)[:, m])
    full = np.arange(g.P)
    Kf = toeplitz.calc_kernel(g, phi2, traj, mask=full)
    assert rel(toeplitz.apply(g, x, S, Kf, mask=full), toeplitz.apply(g, x, S, K)) < 1e-14


def test_mask_single_sample_box():
    """One sample on grid point 0 marks a (2r+1)^2 box centred on the centered
    index (M/2, M/2); a stack repeats it on every plane."""
    for tt, n in (("2D", (8, 8)), ("stack", (8, 8, 3))):
        g = cv.make_geometry(tt, n, 1)
        traj = np.zeros((2, 1, 1), np.float32)
        for r in (0, 1, 2):
            m = omask.calc_mask(g, traj, r)
            Mx, My = g.M[0], g.M[1]
            want = sorted((x + Mx * y) for y in range(My // 2 - r, My // 2 + r + 1)
                          for x in range(Mx // 2 - r, Mx // 2 + r + 1))
            planes = g.M[2] if tt == "stack" else 1
            want = [w + Mx * My * z for z in range(planes) for w in want]
            assert m.tolist() == want


def test_mask_brute_force():
    rng = np.random.default_rng(13)
    for tt, n in (("2D", (9, 7)), ("3D", (5, 4, 3))):
        g = cv.make_geometry(tt, n, 1)
        traj = rng.uniform(-np.pi, np.pi, (g.ndim_traj, 3, 4)).astype(np.float32)
        r = 1
        m = set(omask.calc_mask(g, traj, r).tolist())
        nd = g.ndim_traj
        M = [g.M[a] for a in range(nd)]
        hits = []
        for p in traj.reshape(nd, -1).T:
            hits.append([int(np.mod(np.rint(np.float32(p[a]) * np.float32(M[a] / omask.TAU)) + M[a] // 2, M[a]))
                         for a in range(nd)])
        want = set()
        for lin in range(int(np.prod(M))):
            c, rem = [], lin
            for a in range(nd):
                c.append(rem % M[a])
                rem //= M[a]
            for h in hits:
                if all(min((c[a] - h[a]) % M[a], (h[a] - c[a]) % M[a]) <= r for a in range(nd)):
                    want.add(lin)
                    break
        assert m == want


def test_generators_reference_draws():
    traj = gen.make_traj("2D", 5, 8)
    assert traj.shape == (2, 8, 5) and traj.dtype == np.float32
    assert np.allclose(traj[0, :, 0], gen.radial_readout(8)) and np.allclose(traj[1, :, 0], 0)
    assert np.all(traj >= -np.pi) and np.all(traj < np.pi + 1e-6)
    g, traj, phi2, x, S = gen.reference_case("2d_50x40_L3_c3")
    assert phi2.shape == (3, 3, 61) and x.shape == (3, 40, 50) and S.shape == (3, 40, 50)
    B = cv.gram_blocks(phi2)
    assert np.allclose(B, np.conj(np.swapaxes(B, -1, -2)), atol=1e-6)  # Hermitian
    assert np.all(np.linalg.eigvalsh(B.astype(np.complex128)) > -1e-5)  # PSD


@pytest.mark.parametrize("name", ["2d_33x27_L1_c2_odd", "2d_50x40_L3_c3", "checkop_2d",
                                  "stack_31x36x9_L4_c2_odd", "stack_40x40x8_L16_C12_T200",
                                  "koosh_24_L8_C8_T100"])
def test_golden_regression(golden, name):
    from tests.golden.make_golden import case_inputs, input_checksum
    gd = golden(name)
    g, traj, phi2, x, S = case_inputs(name)
    assert np.allclose(input_checksum(traj, phi2, x, S), gd["input_sum"], rtol=1e-12)
    K = toeplitz.calc_kernel(g, phi2, traj)
    assert rel(K.reshape(g.T, -1)[:, gd["kidx"]], gd["K_lower"]) < 1e-6
    y = toeplitz.apply(g, x, S, K)
    assert rel(y, gd["AhAx_lower"]) < 1e-6
    assert rel(gd["AhAx_upper"], gd["AhAx_lower"]) < 1e-6
    assert rel(gd["AhAx_nudft"], gd["AhAx_lower"]) < 1e-6
    m = omask.calc_mask(g, traj, 2)
    assert m.size == gd["mask_len"]
    assert int(np.sum(m * (1 + np.arange(m.size) % 13))) == gd["mask_sum"]
