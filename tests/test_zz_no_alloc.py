"""mr_op_cg_noalloc (CG on the library's static device work array, no
allocation) against mr_op_cg, called through the C ABI.

Two calls in a row, the second on a smaller problem (the array keeps the
first call's contents), each bit-identical to mr_op_cg on the same operator
and right-hand side (same kernels, same launch shapes, same order); then the
limits: a history too long for the array and graph capture are refused with
MR_ERR_INVALID_ARG.
"""

import ctypes

import numpy as np
import pytest

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


def _op(mr, n, L, C, seed, stream):
    from oracle import conventions as cv
    g = cv.make_geometry("2D", n, L, C)
    rng = np.random.default_rng(seed)
    kshape = (g.T,) + g.grid_shape()
    # a Hermitian positive kernel: constant identity blocks plus a small random part
    K = 0.05 * (rng.standard_normal(kshape) + 1j * rng.standard_normal(kshape))
    for l in range(L):
        t = l * L - l * (l - 1) // 2
        K[t] = 1.0 + 0.05 * rng.standard_normal(kshape[1:])
    S = (rng.standard_normal((C,) + g.img_shape())
         + 1j * rng.standard_normal((C,) + g.img_shape())).astype(np.complex64)
    b = (rng.standard_normal((L,) + g.img_shape())
         + 1j * rng.standard_normal((L,) + g.img_shape())).astype(np.complex64)
    cfg = mr.make_config("2D", n, L, n_coils=C, stream=stream)
    dK = dev(K.astype(np.complex64))
    op = mr.create_op(cfg, dev(S), None, None, _phi2_toep=dK)
    return op, dev(b), dK


def _solve(mr, fn, op, b, iters, refresh):
    from mr_aha_amd import _lib
    x = torch.full_like(b, float("nan"))
    hist = (ctypes.c_double * (iters + 1))()
    rc = fn(op._handle, None, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
            ctypes.c_double(1e-2), iters, refresh, _lib.CG_ZERO_INIT, hist)
    _lib.check(rc, "cg")
    torch.cuda.synchronize()
    return x, list(hist)


def test_noalloc_matches_cg_twice(mr):
    from mr_aha_amd import _lib
    s = torch.cuda.Stream()
    for n, L, C, iters in (((96, 80), 8, 3, 25), ((48, 40), 4, 2, 12)):
        print(f"[noalloc] 2D {n} L={L} C={C} iters={iters}", flush=True)
        op, b, _keep = _op(mr, n, L, C, sum(n) + L, s)
        with op:
            x_ref, h_ref = _solve(mr, _lib.lib.mr_op_cg, op, b, iters, 5)
            x_na, h_na = _solve(mr, _lib.lib.mr_op_cg_noalloc, op, b, iters, 5)
        assert torch.isfinite(torch.view_as_real(x_na)).all()
        assert h_na == h_ref, (h_na[-3:], h_ref[-3:])
        assert torch.equal(x_na, x_ref)
        assert h_ref[-1] < h_ref[0]


def test_noalloc_limits(mr):
    from mr_aha_amd import _lib
    s = torch.cuda.Stream()
    op, b, _keep = _op(mr, (32, 32), 4, 1, 3, s)
    x = torch.empty_like(b)
    with op:
        # a residual history of 10M entries does not fit the 64 MiB array
        rc = _lib.lib.mr_op_cg_noalloc(op._handle, None, ctypes.c_void_p(b.data_ptr()),
                                       ctypes.c_void_p(x.data_ptr()), ctypes.c_double(0.0),
                                       10_000_000, 0, _lib.CG_ZERO_INIT, None)
        assert rc == -1
        assert b"scratch array" in _lib.lib.mr_last_error()
        rc = _lib.lib.mr_op_cg_noalloc(op._handle, None, ctypes.c_void_p(b.data_ptr()),
                                       ctypes.c_void_p(x.data_ptr()), ctypes.c_double(0.0), 5, 0,
                                       _lib.CG_ZERO_INIT | _lib.CG_GRAPH, None)
        assert rc == -1
        assert b"MR_CG_GRAPH" in _lib.lib.mr_last_error()
