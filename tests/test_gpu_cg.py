"""GPU tests of the CG entry point (mr_op_cg): agreement with an FP64 CG on the
oracle's apply, the reported residual against the true FP64 residual, and
HIP-graph replay against eager execution.

Tolerances (relative): 1e-3 for the residual history of the first iterations
against FP64 CG on the same FP32 kernel (the GPU vectors and operator are
FP32); 5e-2 for the final reported residual against the FP64 residual of the
returned x (at a relative residual near 1e-3 the FP32 apply's 1e-6 relative
error is a visible fraction of it); bit equality for graph vs eager.
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


def _smooth(rng, shape):
    """Complex field with only the lowest quarter of frequencies per axis."""
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = np.fft.fftn(z, axes=(-2, -1))
    ny, nx = shape[-2:]
    ky = np.abs(np.fft.fftfreq(ny))[:, None] < 0.125
    kx = np.abs(np.fft.fftfreq(nx))[None, :] < 0.125
    return np.fft.ifftn(f * (ky & kx), axes=(-2, -1))


def _problem(mr, stream=None):
    from tests.golden.make_golden import case_inputs
    g, traj, phi2, x, S = case_inputs("radial_128_L4_C4")
    cfg = mr.make_config(g.traj_type, g.n[: g.n_spatial], g.L, phi2_dim=g.M[: g.n_spatial],
                         n_coils=g.n_coils, stream=stream)
    op = mr.create_op(cfg, dev(S), dev(phi2), dev(traj))
    rng = np.random.default_rng(31)
    xt = _smooth(rng, x.shape).astype(np.complex64)
    return g, S, op, xt


def _fp64_cg(g, S, K, b, lam, iters):
    from oracle import toeplitz
    A = lambda v: toeplitz.apply(g, v, S, K) + lam * v
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rr = np.vdot(r, r).real
    bn = np.sqrt(np.vdot(b, b).real)
    hist = [np.sqrt(rr) / bn]
    for _ in range(iters):
        Ap = A(p)
        alpha = rr / np.vdot(p, Ap).real
        x = x + alpha * p
        r = r - alpha * Ap
        rn = np.vdot(r, r).real
        p = r + (rn / rr) * p
        rr = rn
        hist.append(np.sqrt(rr) / bn)
    return x, hist


def test_cg_against_fp64(mr):
    from oracle import toeplitz
    g, S, op, xt = _problem(mr)
    with op:
        K = host(op.kernel(0)).astype(np.complex128)
        dxt = dev(xt)
        Axt = op(dxt)
        lam = 1e-3 * float(torch.linalg.vector_norm(Axt) / torch.linalg.vector_norm(dxt))
        b = (Axt + lam * dxt).contiguous()
        iters = 30
        x, hist = op.cg(b, lam=lam, iters=iters, refresh=10)
        hb = host(b).astype(np.complex128)
        x64, hist64 = _fp64_cg(g, S.astype(np.complex128), K, hb, lam, 6)
        for k in range(7):
            assert abs(hist[k] - hist64[k]) <= 1e-3 * hist64[k] + 1e-7, (k, hist[k], hist64[k])
        assert hist[iters] < 0.05 * hist[0]
        # the last iteration refreshed: its residual is the true one, in FP32
        xr = host(x).astype(np.complex128)
        r = hb - toeplitz.apply(g, xr, S.astype(np.complex128), K) - lam * xr
        true = np.linalg.norm(r) / np.linalg.norm(hb)
        assert abs(hist[iters] - true) <= 5e-2 * true, (hist[iters], true)
        # the solution approaches x_true
        assert rel(xr, xt.astype(np.complex128)) < 0.5


def test_cg_graph_matches_eager(mr):
    s = torch.cuda.Stream()
    g, S, op, xt = _problem(mr, stream=s.cuda_stream)
    with op, torch.cuda.stream(s):
        dxt = dev(xt)
        b = op(dxt)
        s.synchronize()
        x1, h1 = op.cg(b, lam=1e-2, iters=12, refresh=4)
        for rep in range(3):
            # eager work and synchronisations between the replays
            y = op(x1)
            s.synchronize()
            x2, h2 = op.cg(b, lam=1e-2, iters=12, refresh=4, graph=True)
            assert torch.equal(x1, x2), rep
            assert h1 == h2, rep
            assert torch.isfinite(y).all()


def test_cg_refresh_with_second_operator(mr):
    """op_refresh = op itself is the plain residual replacement (bit-identical
    to no op_refresh); a separate but identical operator restarts each cycle
    from its true residual and converges to the same solution."""
    g, S, op, xt = _problem(mr)
    _, _, op2, _ = _problem(mr)
    with op, op2:
        b = op(dev(xt))
        lam = 0.5 * float(torch.linalg.vector_norm(b) / torch.linalg.vector_norm(dev(xt)))
        b = (b + lam * dev(xt)).contiguous()
        x1, h1 = op.cg(b, lam=lam, iters=120, refresh=10)
        x0, h0 = op.cg(b, lam=lam, iters=120, refresh=10, op_refresh=op)
        assert torch.equal(x0, x1) and h0 == h1
        x2, h2 = op.cg(b, lam=lam, iters=120, refresh=10, op_refresh=op2)
        assert h2[-1] < 1e-4 * h2[0]
        assert float(torch.linalg.vector_norm(x2 - x1) / torch.linalg.vector_norm(x1)) < 1e-4


def test_cg_argument_errors(mr):
    g, S, op, xt = _problem(mr)
    with op:
        b = op(dev(xt))
        with pytest.raises(Exception, match="n_iter"):
            op.cg(b, iters=-1)
        with pytest.raises(Exception, match="lambda"):
            op.cg(b, lam=-1.0)
        with pytest.raises(Exception, match="non-default"):
            op.cg(b, iters=2, graph=True)
