"""GPU tests of the 16-bit kernel format (mr_op_set_kernel_format) and of CG
iterative refinement with it.

The format is restated here in numpy (`quantise`): per location and 4x4 block
of the L x L matrix the scale s = max(|Re|, |Im|) of the block's stored
entries, entries multiplied by the FP32 reciprocal 1/s and rounded to
nearest-even BF16 or FP16. The GPU apply on the 16-bit kernel must equal the
FP64 apply on that dequantised kernel to 1e-5 (the FP32 apply tolerance),
and differ from the FP32-kernel apply by no more than the format's rounding
(BF16 8e-3, FP16 2e-3 relative).
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


def _bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32)


def quantise(K, L, upper, fmt):
    """Dequantised copy of the packed kernel K (T, P) complex64."""
    from oracle import conventions as cv
    pairs = cv.tri_pairs(L, upper)
    blk = [max(r // 4, c // 4) * (max(r // 4, c // 4) + 1) // 2 + min(r // 4, c // 4)
           for (r, c) in pairs]
    nb = ((L + 3) // 4) * ((L + 3) // 4 + 1) // 2
    re = K.real.astype(np.float32)
    im = K.imag.astype(np.float32)
    mx = np.zeros((nb,) + K.shape[1:], np.float32)
    for t, b in enumerate(blk):
        mx[b] = np.maximum(mx[b], np.maximum(np.abs(re[t]), np.abs(im[t])))
    mx[~(mx > 0)] = 1.0
    out = np.empty_like(K)
    for t, b in enumerate(blk):
        inv = np.float32(1.0) / mx[b]
        a, c = re[t] * inv, im[t] * inv
        if fmt == "bf16":
            qa, qc = _bf16(a), _bf16(c)
        else:
            qa, qc = a.astype(np.float16).astype(np.float32), c.astype(np.float16).astype(np.float32)
        out[t] = (qa.astype(np.float64) * mx[b] + 1j * (qc.astype(np.float64) * mx[b]))
    return out


SHAPES = [("2D", (128, 8), 4, 3), ("2D", (160, 10), 8, 2), ("2D", (320, 6), 16, 2),
          ("stack", (320, 4, 3), 16, 2)]


@pytest.mark.parametrize("tt,n,L,C", SHAPES)
@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_kernel_format_apply(mr, tt, n, L, C, fmt):
    from oracle import conventions as cv
    from oracle import toeplitz
    print(f"[kformat] {tt} {n} L={L} {fmt}", flush=True)
    rng = np.random.default_rng(5 * L + C)
    g = cv.make_geometry(tt, n, L, C)
    kshape = (g.T,) + g.grid_shape()
    # entries spanning several decades, as a real kernel's do
    mag = np.exp(rng.uniform(-6, 0, kshape))
    K = (mag * np.exp(2j * np.pi * rng.random(kshape))).astype(np.complex64)
    x = (rng.standard_normal((L,) + g.img_shape()) + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    S = (rng.standard_normal((C,) + g.img_shape()) + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    for tri in ("lower", "upper"):
        up = tri == "upper"
        cfg = mr.make_config(g.traj_type, g.n[: g.n_spatial], L, phi2_dim=g.M[: g.n_spatial],
                             n_coils=C, triangulation=tri)
        dK = dev(K)
        with mr.create_op(cfg, dev(S), None, None, _phi2_toep=dK) as op:
            y32 = host(op(dev(x)))
            op.set_kernel_format(fmt)
            assert op.kernel_format == fmt
            y16 = host(op(dev(x)))
            Kq = quantise(K.reshape(g.T, -1), L, up, fmt).reshape(kshape)
            ref = toeplitz.apply(g, x, S, Kq, upper=up)
            assert rel(y16, ref) < 1e-5, (tri, rel(y16, ref))
            bound = 8e-3 if fmt == "bf16" else 2e-3
            d = rel(y16, y32)
            assert 0 < d < bound, (tri, d)
            op.set_kernel_format("fp32")
            assert np.array_equal(host(op(dev(x))), y32)


def test_kernel_format_unsupported_shape(mr):
    from tests.golden.make_golden import case_inputs
    g, traj, phi2, x, S = case_inputs("2d_50x40_L3_c3")
    cfg = mr.make_config(g.traj_type, g.n[: g.n_spatial], g.L, phi2_dim=g.M[: g.n_spatial],
                         n_coils=g.n_coils)
    with mr.create_op(cfg, dev(S), dev(phi2), dev(traj)) as op:
        y = host(op(dev(x)))
        with pytest.raises(Exception, match="no compiled fused x pass"):
            op.set_kernel_format("bf16")
        assert np.array_equal(host(op(dev(x))), y)  # still the FP32 path
        with pytest.raises(ValueError):
            op.set_kernel_format("fp8")


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_cg_iterative_refinement(mr, fmt):
    """CG on the 128x128 L=4 C=4 radial problem, well conditioned by lambda:
    cycles of 10 iterations on the 16-bit operator, each restarted from the
    true residual of the FP32 operator, reach the converged FP32 solution to
    1e-4, and below the error of CG on the 16-bit operator alone."""
    from tests.golden.make_golden import case_inputs
    g, traj, phi2, x, S = case_inputs("radial_128_L4_C4")
    cfg = mr.make_config(g.traj_type, g.n[: g.n_spatial], g.L, phi2_dim=g.M[: g.n_spatial],
                         n_coils=g.n_coils)
    dS, dphi2, dtraj = dev(S), dev(phi2), dev(traj)
    with mr.create_op(cfg, dS, dphi2, dtraj) as op32:
        K = op32.kernel(0)
        with mr.create_op(cfg, dS, None, None, _phi2_toep=K) as op16:
            op16.set_kernel_format(fmt)
            rng = np.random.default_rng(3)
            xt = dev((rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)).astype(np.complex64))
            Axt = op32(xt)
            lam = 0.5 * float(torch.linalg.vector_norm(Axt) / torch.linalg.vector_norm(xt))
            b = (Axt + lam * xt).contiguous()
            xref, href = op32.cg(b, lam=lam, iters=200, refresh=10)
            xmix, hmix = op16.cg(b, lam=lam, iters=60, refresh=10, op_refresh=op32)
            xlo, hlo = op16.cg(b, lam=lam, iters=60, refresh=10)
            nr = torch.linalg.vector_norm(xref)
            e_mix = float(torch.linalg.vector_norm(xmix - xref) / nr)
            e_lo = float(torch.linalg.vector_norm(xlo - xref) / nr)
            e_true = float(torch.linalg.vector_norm(xref - xt) / torch.linalg.vector_norm(xt))
            print(f"[cg-{fmt}] fp32 ref vs truth {e_true:.2e} (res {href[-1]:.1e}); refined "
                  f"{e_mix:.2e} (res {hmix[-1]:.1e}); 16-bit only {e_lo:.2e} (res {hlo[-1]:.1e})",
                  flush=True)
            assert href[-1] < 1e-5
            assert e_mix < 1e-4
            assert e_mix < e_lo
