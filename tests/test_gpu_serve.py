"""GPU test of the continuous-batching server (mr_aha_amd.serve): every
request's solution is the same when served alone and when served among
overlapping requests of other shapes (tolerance 1e-6 relative: each request
runs the same kernels in the same order on its own stream, so equality is
expected to be exact), and admission respects the slot and memory limits."""

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def serve():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from mr_aha_amd import serve
    return serve


SHAPES = [("2D", 64, 4, 3, "fp32"), ("2D", 96, 8, 4, "fp32"), ("stack", 48, 4, 2, "fp32"),
          ("2D", 128, 4, 4, "bf16"), ("2D", 80, 16, 2, "fp32"), ("stack", 40, 8, 3, "fp32")]


def _requests(serve):
    from mr_aha_amd import workloads as wl
    reqs = []
    for i, (tt, n, L, C, prec) in enumerate(SHAPES):
        if tt == "2D":
            w = wl.radial_2d(n=n, L=L, C=C, n_spokes=2 * n, nkx=2 * n, n_frames=32, seed=100 + i)
            nd = (n, n)
        else:
            w = wl.stack_of_stars(n=n, nz=8, L=L, C=C, n_spokes=n, nkx=2 * n, n_frames=32,
                                  seed=100 + i)
            nd = (n, n, 8)
        b = wl.complex_gaussian(tuple(w["x"].shape), 500 + i, "cuda")
        reqs.append(serve.Request(rid=i, traj_type=tt, n_dim=nd, L=L, n_coils=C, SEs=w["SEs"],
                                  phi2=w["phi2"], traj=w["traj"], b=b, iters=12, lam=1e-2,
                                  refresh=4, precision=prec, arrival=0.02 * i))
    return reqs


def test_same_alone_and_in_mix(serve):
    alone = serve.Server(slots=1).run(_requests(serve), realtime=False)
    mix = serve.Server(slots=3).run(_requests(serve), realtime=True)
    for a, m in zip(alone, mix):
        assert a.error is None and m.error is None, (a.error, m.error)
        d = float(torch.linalg.vector_norm(a.x - m.x) / torch.linalg.vector_norm(a.x))
        assert d < 1e-6, (a.rid, d)
        # (the residual norm of CG is not monotone; the FP64 agreement of the
        # solver is tested in test_gpu_cg.py)
        assert all(h == h and h < float("inf") for h in a.hist) and a.hist == m.hist
        assert m.t_admit >= m.arrival - 1e-3 and m.t_done > m.t_admit
    s = serve.summary(mix)
    assert s["completed"] == len(SHAPES) and s["p99_s"] >= s["p50_s"] > 0


def test_admission_limits(serve):
    reqs = _requests(serve)
    fp = [serve.footprint_bytes(r) for r in reqs]
    # a budget that holds the largest request alone but not two of them
    srv = serve.Server(budget_bytes=max(fp) + min(fp) // 2, slots=4)
    out = srv.run(reqs, realtime=False)
    assert all(r.error is None for r in out)
    # no two requests whose footprints exceed the budget overlapped
    for i, a in enumerate(out):
        for b in out[i + 1:]:
            overlap = a.t_admit < b.t_done and b.t_admit < a.t_done
            if overlap:
                assert serve.footprint_bytes(a) + serve.footprint_bytes(b) <= srv.budget
    # a request larger than the budget is refused, the others still served
    small = serve.Server(budget_bytes=min(fp), slots=2).run(_requests(serve), realtime=False)
    assert any(r.error and "exceeds the budget" in r.error for r in small)
    assert all(r.x is not None for r in small if r.error is None)


def _well_posed(serve, i, tt, n, nz, L, C, ns, prec, iters=30):
    """A request at a real serving size whose solution is known: b =
    (A^H A + lam I) x_true with lam = 0.1 lambda_max (power iteration from
    the smooth x_true), formed by the request's own FP32 operator."""
    import mr_aha_amd as mr
    from mr_aha_amd import workloads as wl
    if tt == "2D":
        w = wl.radial_2d(n=n, L=L, C=C, n_spokes=ns, nkx=2 * n, n_frames=64, seed=300 + i,
                         with_x=False)
        nd, shape = (n, n), (L, n, n)
    else:
        w = wl.stack_of_stars(n=n, nz=nz, L=L, C=C, n_spokes=ns, nkx=2 * n, n_frames=64,
                              seed=300 + i, with_x=False)
        nd, shape = (n, n, nz), (L, nz, n, n)
    z = wl.complex_gaussian(shape, 700 + i, "cuda")
    f = torch.fft.fftn(z, dim=(-2, -1))
    ky = torch.fft.fftfreq(shape[-2], device="cuda").abs()[:, None] < 0.1
    kx = torch.fft.fftfreq(shape[-1], device="cuda").abs()[None, :] < 0.1
    xt = torch.fft.ifftn(f * (ky & kx), dim=(-2, -1)).to(torch.complex64).contiguous()
    cfg = mr.make_config(tt, nd, L, n_coils=C, mem_fraction=0.25)
    with mr.create_op(cfg, w["SEs"], w["phi2"], w["traj"]) as op:
        v = xt
        for _ in range(6):
            u = op(v / torch.linalg.vector_norm(v))
            lmax = float(torch.linalg.vector_norm(u))
            v = u
        lam = 0.1 * lmax
        b = (op(xt) + lam * xt).contiguous()
    req = serve.Request(rid=i, traj_type=tt, n_dim=nd, L=L, n_coils=C, SEs=w["SEs"],
                        phi2=w["phi2"], traj=w["traj"], b=b, iters=iters, lam=lam, refresh=10,
                        precision=prec, arrival=0.05 * i)
    return req, xt


# config-5 families at their largest sizes: 2D 320^2 and stack 256^2 x 48, L16 C16
REAL = [("2D", 320, 1, 16, 16, 512, "fp32"), ("stack", 256, 48, 16, 16, 400, "fp32"),
        ("2D", 320, 1, 16, 16, 512, "bf16"), ("stack", 192, 32, 8, 12, 300, "bf16"),
        ("2D", 128, 1, 4, 8, 256, "fp32")]


def test_real_size_requests_alone_and_in_mix(serve):
    """Each request's solution equals its solo run (1e-6) and is within 1e-4
    of x_true; in the mix the requests' device work overlaps in time."""
    made = [_well_posed(serve, i, *spec) for i, spec in enumerate(REAL)]
    reqs = [r for r, _ in made]
    xts = [x for _, x in made]
    alone = {}
    for r in reqs:
        out = serve.Server(slots=1).run([r], realtime=False)[0]
        assert out.error is None, out.error
        alone[r.rid] = out.x.clone()
        print(f"[serve] request {r.rid} {r.traj_type} {r.n_dim} L={r.L} C={r.n_coils} "
              f"{r.precision}: alone {out.d_end - out.d_start:.3f} s", flush=True)
    made = [_well_posed(serve, i, *spec) for i, spec in enumerate(REAL)]
    mix = serve.Server(slots=4).run([r for r, _ in made], realtime=True)
    for m in mix:
        assert m.error is None, m.error
        d = float(torch.linalg.vector_norm(m.x - alone[m.rid]) / torch.linalg.vector_norm(m.x))
        e = float(torch.linalg.vector_norm(m.x - xts[m.rid]) / torch.linalg.vector_norm(xts[m.rid]))
        print(f"[serve] request {m.rid}: mix vs alone {d:.1e}, error vs x_true {e:.1e}", flush=True)
        assert d < 1e-6, (m.rid, d)
        assert e < 1e-4, (m.rid, e)
    ov = serve.device_overlap(mix)
    print(f"[serve] device time with two or more requests at work: {ov:.0%}", flush=True)
    assert ov > 0.0
