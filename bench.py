#!/usr/bin/env python3
"""Benchmark: Toeplitz A^H A applies/s on one MI355X.

Default workload (one step = one A^H A apply, inputs resident in HBM):
stack-of-stars 320x320x56 image, 640x640x56 Toeplitz grid, L=16, C=12,
400 golden-angle spokes x 640 samples on every kz plane, frames round-robin
over 200, synthetic data, kernel built once with mr_calc_toep.

Prints one JSON line. `roofline` is for the dominant kernel of the apply:
its algorithmic bytes per launch divided by its average launch time, the
latter measured with HIP events on the library's stream over the timed
region (MR_AHA_HIP_PROFILE). `cpu_baseline` times the oracle's CPU Toeplitz
apply (complex64, pocketfft threads) on one of the C coils of the same
workload and scales to applies/s.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# read + write bandwidth of the library's streaming copy on one MI355X
# (python bench.py --calibrate-copy; profiles/stream_copy_ceiling.md)
COPY_CEILING_GBS = 5240.3


def stage_bytes(cfg, nc, xbd, precision="fp32"):
    """Algorithmic HBM bytes of one launch of each profiled stage for a group
    of nc coils: every array a stage must read or write, counted once (the
    kernel once per launch, however many coil groups of a row share it).

    xbd > 0 (fused x pass): F3 = outer-axis FFT (S x products, zero-pad),
    F2 = y-FFT, BD = fused x-FFT + block multiply + inverse x-FFT in place,
    I2 = inverse y-FFT, I3 = inverse outer FFT with conj(S) and the coil sum;
    in 2D the outer axis is y (F2 and I2 only).
    xbd == 0, stack-of-stars (z transformed on the image side): F3 slot = Z1
    (S x products, z-FFT), I3 slot = Z2 (inverse z-FFT, conj(S), coil sum).
    """
    Nx, Ny, Nz = cfg.n_dim
    Mx, My, Mz = cfg.phi2_dim
    L = cfg.L
    V = L * nc
    T = L * (L + 1) // 2
    N = Nx * Ny * Nz
    P = Mx * My * Mz
    c8 = 8
    kcol = cfg.num_masked if cfg.is_masked else P
    nblk = ((L + 3) // 4) * ((L + 3) // 4 + 1) // 2
    kern = T * kcol * c8 if precision == "fp32" else (T + nblk) * kcol * 4
    xs = N * L * c8 + N * nc * c8       # x and the maps, once
    outs = N * nc * c8 + N * L * c8     # maps and AhAx
    three = cfg.n_spatial == 3
    if xbd:
        if not three:
            b = Nx * My * V * c8
            return {"F2": xs + b, "BD": 2 * b + kern, "I2": b + outs}
        a = Nx * Ny * Mz * V * c8
        b = Nx * My * Mz * V * c8
        return {"F3": xs + a, "F2": a + b, "BD": 2 * b + kern, "I2": b + a, "I3": a + outs}
    img_v = N * V * c8                  # image-size volumes
    rows_x = Mx * Ny * Nz * V * c8      # x-transformed image rows
    rows_xy = Mx * My * Nz * V * c8     # x,y-transformed planes
    full = P * V * c8
    zfirst = cfg.traj_type == 2 and Mz == Nz
    if zfirst:
        return {"F3": xs + img_v, "F1": img_v + rows_x, "F2": rows_x + rows_xy,
                "BD": 2 * full + kern, "I2": rows_xy + rows_x, "I1": rows_x + img_v,
                "I3": img_v + outs}
    return {
        "F1": xs + rows_x,
        "F2": rows_x + rows_xy,
        "F3": (rows_xy + full) if three else 0,
        "BD": 2 * full + kern,
        "I3": (full + rows_xy) if three else 0,
        "I2": rows_xy + rows_x,
        "I1": rows_x + xs,
    }


def apply_model_bytes(cfg):
    """Stage-model algorithmic bytes of one apply: T P 8 + 4 C L P 8 + (2L + 2C) N 8."""
    Nx, Ny, Nz = cfg.n_dim
    Mx, My, Mz = cfg.phi2_dim
    L, C = cfg.L, cfg.n_coils
    T = L * (L + 1) // 2
    N = Nx * Ny * Nz
    P = Mx * My * Mz
    return T * P * 8 + 4 * C * L * P * 8 + (2 * L + 2 * C) * N * 8


def calibrate_copy(nbytes=1 << 30, reps=4):
    """Device copies of a known size.

    * torch copy_ of `nbytes` `reps` times: the reference point for scaling
      the rocprofv3 FETCH_SIZE / WRITE_SIZE readings of the library's kernels;
    * the library's streaming copy (mr_stream_copy: 16-byte non-temporal
      loads and stores) of 8 GiB, timed with HIP events: the attainable HBM
      bandwidth (bytes read + written per second)."""
    import torch
    from mr_aha_amd import _lib
    x = torch.empty(nbytes // 8, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    big = 8 << 30
    a = torch.empty(big // 8, dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    s = torch.cuda.current_stream()
    for _ in range(2):
        _lib.check(_lib.lib.mr_stream_copy(a.data_ptr(), b.data_ptr(), big, s.cuda_stream),
                   "mr_stream_copy")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    n = 10
    e0.record()
    for _ in range(n):
        _lib.check(_lib.lib.mr_stream_copy(a.data_ptr(), b.data_ptr(), big, s.cuda_stream),
                   "mr_stream_copy")
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": reps,
                      "stream_copy_bytes": big, "stream_copy_ms": round(ms, 3),
                      "stream_copy_gbs": round(2 * big / (ms * 1e6), 1)}), flush=True)


def profile_record(kernel, workload, precision="fp32"):
    """The committed rocprofv3 record of `kernel` on `workload` with the
    kernel stored in `precision` (profiles/hbm_traffic.json), or {}."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return {}
    for rec in json.loads(path.read_text()).get("records", []):
        if (rec.get("kernel") == kernel and rec.get("workload") == workload
                and rec.get("precision", "fp32") == precision):
            return rec
    return {}


def measured_traffic(kernel, workload, precision="fp32"):
    """HBM bytes per launch from the committed record, or None."""
    rec = profile_record(kernel, workload, precision)
    return int(rec["hbm_bytes_per_launch"]) if "hbm_bytes_per_launch" in rec else None


def host_threads() -> int:
    """Host threads for the CPU legs: the process's CPU affinity, capped by
    OMP_NUM_THREADS when the environment sets it (the share of the host this
    process is given)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, int(n))


def rel_err(a, b) -> float:
    import numpy as np
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_leg(w, coil_outputs, threads):
    """CPU baseline and FP64 oracle of the workload, one coil.

    Timed: the reference's CPU path, the NUFFT A^H A (ES kernel, eps 1e-6,
    complex64; oracle/cpu_nufft.py) restricted to coil 0, all ranks and kz
    planes, scaled by the coil count to applies/s. Oracle: the same operator
    in complex128 at eps 1e-9 for coil 0, against which the GPU's one-coil
    outputs (`coil_outputs`: name -> array) are stated."""
    import numpy as np
    from oracle import cpu_nufft
    from oracle.conventions import make_geometry
    g = make_geometry(w["traj_type"], w["n_dim"], w["L"], w["n_coils"])
    x = w["x"].cpu().numpy()
    S = w["SEs"][:1].cpu().numpy()
    phi2 = w["phi2"].cpu().numpy()
    traj = w["traj"].cpu().numpy()
    cpu_nufft.lib()
    t0 = time.perf_counter()
    y32 = cpu_nufft.aha_nufft(g, x, S, phi2, traj, eps=1e-6, dtype=np.complex64,
                              threads=threads)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    y64 = cpu_nufft.aha_nufft(g, x, S, phi2, traj, eps=1e-9, dtype=np.complex128,
                              threads=threads)
    dt64 = time.perf_counter() - t1
    C = w["n_coils"]
    base = {"value": 1.0 / (dt * C), "unit": "A^H A applies/s", "cores": threads, "kind": "port",
            "sample": f"NUFFT A^H A (method=\"nufft\", ES kernel, eps 1e-6, complex64) of coil "
                      f"0 of {C}, all ranks and planes: {dt:.1f} s on {threads} threads, "
                      f"scaled x{C}",
            "error_vs_fp64": rel_err(y32, y64)}
    errs = {k: rel_err(v, y64) for k, v in coil_outputs.items()}
    return base, errs, {"fp64_oracle": "NUFFT A^H A, complex128, eps 1e-9, coil 0",
                        "fp64_oracle_s": round(dt64, 1)}


def smooth_field(shape, seed, device, frac=0.1):
    """Complex Gaussian field low-passed to |f| < frac per in-plane axis."""
    import torch
    from mr_aha_amd import workloads
    z = workloads.complex_gaussian(shape, seed, device)
    f = torch.fft.fftn(z, dim=(-2, -1))
    ky = torch.fft.fftfreq(shape[-2], device=device).abs()[:, None] < frac
    kx = torch.fft.fftfreq(shape[-1], device=device).abs()[None, :] < frac
    return torch.fft.ifftn(f * (ky & kx), dim=(-2, -1)).to(torch.complex64).contiguous()


def power_lmax(op, shape, device, iters=20, v=None):
    """Largest eigenvalue of A^H A by power iteration on the device (from v,
    default a complex Gaussian)."""
    import torch
    from mr_aha_amd import workloads
    if v is None:
        v = workloads.complex_gaussian(shape, 4242, device)
    lam = 0.0
    for _ in range(iters):
        v = v / torch.linalg.vector_norm(v)
        u = op(v)
        lam = float(torch.linalg.vector_norm(u))
        v = u
    return lam


def cg_rhs_fp64(w, xt, lam, threads):
    """Oracle side of the CG leg: b = (A^H A + lam I) x_true with the FP64
    NUFFT operator (complex128, eps 1e-12), so that x_true is the FP64
    solution the GPU solves are measured against."""
    import numpy as np
    from oracle import cpu_nufft
    from oracle.conventions import make_geometry
    g = make_geometry(w["traj_type"], w["n_dim"], w["L"], w["n_coils"])
    x = xt.cpu().numpy()
    y = cpu_nufft.aha_nufft(g, x, w["SEs"].cpu().numpy(), w["phi2"].cpu().numpy(),
                            w["traj"].cpu().numpy(), eps=1e-12, dtype=np.complex128,
                            threads=threads)
    return (y + lam * x).astype(np.complex64)


def cg_leg(args, threads):
    """Config 3: CG on (A^H A + lam I) x = b for the 128x128 L=4 C=4 radial
    inputs, lam = 1e-3 lambda_max(A^H A), b formed from a smooth x_true by the
    FP64 NUFFT operator (so x_true is the FP64 solution). FP32 kernel against
    the BF16 kernel refined by the FP32 operator (restart from its true
    residual every 10 iterations). Each solve runs the fewest iterations (a
    multiple of 10, at most 400) whose error against x_true is <= 1e-4; that
    solve is then timed (HIP-graph-replayed iterations)."""
    import torch
    import mr_aha_amd as mr
    from mr_aha_amd import workloads
    w = workloads.radial_2d(device="cuda")
    s = torch.cuda.Stream()
    cfg = mr.make_config(w["traj_type"], w["n_dim"], w["L"], n_coils=w["n_coils"], stream=s)
    out = {}
    with torch.cuda.stream(s):
        K = mr.calc_toep(cfg, w["phi2"], w["traj"])
        op32 = mr.create_op(cfg, w["SEs"], None, None, _phi2_toep=K)
        op16 = mr.create_op(cfg, w["SEs"], None, None, _phi2_toep=K)
        op16.set_kernel_format("bf16")
        xt = smooth_field(tuple(w["x"].shape), 77, "cuda")
        lmax = power_lmax(op32, tuple(w["x"].shape), "cuda")
        lam = 1e-3 * lmax
        if threads:
            b = torch.from_numpy(cg_rhs_fp64(w, xt, lam, threads)).cuda()
            bsrc = "FP64 NUFFT operator (oracle), x_true the FP64 solution"
        else:
            b = (op32(xt) + lam * xt).contiguous()
            bsrc = "FP32 Toeplitz operator (no oracle run), x_true its solution"
        nx = float(torch.linalg.vector_norm(xt))
        for name, op, ref in (("fp32", op32, None), ("bf16", op16, op32)):
            its, err = None, None
            for k in range(10, 401, 10):
                x, h = op.cg(b, lam=lam, iters=k, refresh=10, op_refresh=ref, graph=True)
                err = float(torch.linalg.vector_norm(x - xt)) / nx
                if err <= 1e-4:
                    its = k
                    break
            k = its or 400
            for _ in range(args.warmup):
                op.cg(b, lam=lam, iters=k, refresh=10, op_refresh=ref, graph=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                x, h = op.cg(b, lam=lam, iters=k, refresh=10, op_refresh=ref, graph=True)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            err = float(torch.linalg.vector_norm(x - xt)) / nx
            out[name] = {"iterations": k, "reached_1e-4": its is not None,
                         "error_vs_fp64": err, "final_rel_residual": h[-1],
                         "recon_per_s": round(1.0 / dt, 2), "ms_per_recon": round(dt * 1e3, 3)}
        op16.close()
        op32.close()
    out["speedup_mixed_vs_fp32"] = round(out["bf16"]["recon_per_s"] / out["fp32"]["recon_per_s"], 3)
    out["workload"] = ("2D radial 128x128, grid 256x256, L=4, C=4, 64x256 spokes; "
                       f"lambda = 1e-3 lambda_max = {lam:.4g}; b from the {bsrc}; "
                       "refresh every 10 iterations")
    return out


def serve_requests(n_req, rate, seed=2026):
    """Config 5 request mix: Poisson arrivals; 2D radial {128,192,256,320}^2 or
    stack {192,256}^2 x {32,48}; L {4,8,16}; C {8,12,16}; 256-512 spokes of
    2N; 20-50 CG iterations (multiples of 10, so that each solve ends on a
    true-residual refresh)."""
    import random
    rnd = random.Random(seed)
    specs = []
    t = 0.0
    for i in range(n_req):
        t += rnd.expovariate(rate)
        if rnd.random() < 0.5:
            n = rnd.choice([128, 192, 256, 320])
            spec = ("2D", (n, n))
        else:
            n = rnd.choice([192, 256])
            spec = ("stack", (n, n, rnd.choice([32, 48])))
        specs.append(dict(i=i, tt=spec[0], nd=spec[1], L=rnd.choice([4, 8, 16]),
                          C=rnd.choice([8, 12, 16]), ns=rnd.randint(256, 512),
                          iters=10 * rnd.randint(2, 5), arrival=t))
    return specs


def serve_leg(args):
    """Config 5 through mr_aha_amd.serve: every request is a well-posed CG
    problem, b = (A^H A + lam I) x_true with lam = 0.1 lambda_max (so its
    20-50 iterations reach 1e-4), formed before the run by the request's own
    FP32 operator; errors are against x_true. FP32 kernels, then 16-bit
    kernels refined against FP32 operators."""
    import torch
    import mr_aha_amd as mr
    from mr_aha_amd import serve, workloads
    specs = serve_requests(args.requests, args.rate)
    prepared = []
    t_prep = time.perf_counter()
    for sp in specs:
        n = sp["nd"][0]
        if sp["tt"] == "2D":
            w = workloads.radial_2d(n=n, L=sp["L"], C=sp["C"], n_spokes=sp["ns"], nkx=2 * n,
                                    n_frames=64, seed=900 + sp["i"], with_x=False)
            shape = (sp["L"], n, n)
        else:
            w = workloads.stack_of_stars(n=n, nz=sp["nd"][2], L=sp["L"], C=sp["C"],
                                         n_spokes=sp["ns"], nkx=2 * n, n_frames=64,
                                         seed=900 + sp["i"], with_x=False)
            shape = (sp["L"], sp["nd"][2], n, n)
        xt = smooth_field(shape, 1900 + sp["i"], "cuda")
        cfg = mr.make_config(w["traj_type"], w["n_dim"], w["L"], n_coils=w["n_coils"],
                             mem_fraction=0.25)
        with mr.create_op(cfg, w["SEs"], w["phi2"], w["traj"]) as op:
            lam = 0.1 * power_lmax(op, shape, "cuda", iters=6, v=xt)
            b = (op(xt) + lam * xt).contiguous()
        prepared.append((sp, w, xt, b, lam))
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t_prep
    out = {}
    for prec in ("fp32", "bf16"):
        reqs = []
        for sp, w, xt, b, lam in prepared:
            reqs.append(serve.Request(rid=sp["i"], traj_type=sp["tt"], n_dim=sp["nd"], L=sp["L"],
                                      n_coils=sp["C"], SEs=w["SEs"], phi2=w["phi2"],
                                      traj=w["traj"], b=b, iters=sp["iters"], lam=lam,
                                      refresh=10, precision=prec, arrival=sp["arrival"]))
        torch.cuda.synchronize()
        done = serve.Server(slots=args.slots).run(reqs, realtime=True)
        summ = serve.summary(done)
        errs = []
        for r, (sp, w, xt, b, lam) in zip(sorted(done, key=lambda r: r.rid), prepared):
            if r.error is None:
                errs.append(float(torch.linalg.vector_norm(r.x - xt) /
                                  torch.linalg.vector_norm(xt)))
        summ["errors"] = sorted({r.error for r in done if r.error})[:3]
        summ["max_error_vs_x_true"] = max(errs) if errs else None
        summ["median_error_vs_x_true"] = sorted(errs)[len(errs) // 2] if errs else None
        summ["all_within_1e-4"] = bool(errs) and len(errs) == len(done) and max(errs) <= 1e-4
        out[prec] = summ
        print(f"[serve] {prec}: {summ}", file=sys.stderr, flush=True)
        del reqs, done
        torch.cuda.empty_cache()
    out["workload"] = (f"{args.requests} Poisson requests at {args.rate}/s, {args.slots} slots: "
                       "2D radial {128,192,256,320}^2 or stack {192,256}^2 x {32,48}; L {4,8,16}; "
                       "C {8,12,16}; 256-512 spokes of 2N; 20-50 CG iterations; "
                       "lam = 0.1 lambda_max; b formed before the run "
                       f"({t_prep:.0f} s, not timed)")
    return out


def time_applies(op, x, out, steps, warmup):
    """Warm-up, then `steps` applies timed with HIP events on the device and
    per-stage events on the library's stream (MR_AHA_HIP_PROFILE)."""
    import torch
    from mr_aha_amd import _lib
    for _ in range(warmup):
        op.apply(x, out)
    torch.cuda.synchronize()
    os.environ["MR_AHA_HIP_PROFILE"] = "1"
    ms = (ctypes.c_double * 8)()
    calls = (ctypes.c_int * 8)()
    _lib.check(_lib.lib.mr_op_profile(op._handle, ms, calls, 8), "mr_op_profile")  # reset
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    t_wall = time.perf_counter()
    for _ in range(steps):
        op.apply(x, out)
    e1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    os.environ.pop("MR_AHA_HIP_PROFILE")
    _lib.check(_lib.lib.mr_op_profile(op._handle, ms, calls, 8), "mr_op_profile")
    return e0.elapsed_time(e1) / steps, t_wall * 1000 / steps, ms, calls


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="stack", choices=["stack", "radial", "koosh", "2d72"],
                    help="workload of the applies/s (stack: the headline; 2d72: the reference's "
                         "large-rank shape, random kernel; the other legs run with stack only)")
    ap.add_argument("--masked", action="store_true",
                    help="2d72: the masked operator (mask of the 754 x 320 radial trajectory)")
    ap.add_argument("--requests", type=int, default=200, help="serving leg: number of requests")
    ap.add_argument("--rate", type=float, default=4.0, help="serving leg: arrivals per second")
    ap.add_argument("--slots", type=int, default=4, help="serving leg: concurrent requests")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / FP64 oracle")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 16-bit kernel leg")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG leg (config 3)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg (config 5)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in tile rules instead of timing the candidates "
                         "(mr_op_autotune) before the warm-up")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py measures one GPU: the multi-GPU coil split is not part of this build "
              "(every run here is single-GPU); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import mr_aha_amd as mr
    from mr_aha_amd import _lib, workloads

    torch.cuda.set_device(0)
    dev = "cuda:0"
    if args.config != "stack":
        # the CG and serving legs belong to the headline run; 2d72 (random
        # kernel, L=72) has no 16-bit form and no trajectory data to check
        args.no_cg = args.no_serve = True
        if args.config == "2d72":
            args.no_mixed = args.no_cpu = True
        if args.config == "koosh":
            # the 3D CPU NUFFT of 6.4M samples (and its FP64 oracle at w = 10)
            # takes minutes per coil: no CPU leg
            args.no_cpu = True
    threads = 0 if args.no_cpu else host_threads()
    t_gen = time.perf_counter()
    K_given, mask = None, None
    if args.config == "2d72":
        # src/tests/bench_large_rank.py's shape: 320^2 image, 640^2 grid, L=72,
        # 12 coils, 754 golden-angle spokes x 320 samples (for the mask); the
        # apply's time does not depend on the kernel's values, so it is random
        n, L, C = 320, 72, 12
        g = torch.Generator(device=dev).manual_seed(72)
        traj = workloads.radial_traj(754, n, dev)
        w = dict(traj_type="2D", n_dim=(n, n), L=L, n_coils=C, traj=traj, phi2=None,
                 SEs=workloads.complex_gaussian((C, n, n), 2, dev),
                 x=workloads.complex_gaussian((L, n, n), 1, dev))
        workload = ("2d72: 2D 320x320, grid 640x640, L=72, C=12, random kernel" +
                    (", masked (754x320 radial spokes)" if args.masked else ""))
    elif args.config == "stack":
        w = workloads.stack_of_stars(device=dev)
        workload = "stack-of-stars 320x320x56, grid 640x640x56, L=16, C=12, 400x640 spokes"
    elif args.config == "radial":
        w = workloads.radial_2d(device=dev)
        workload = "2D radial 128x128, grid 256x256, L=4, C=4, 64x256 spokes"
    else:
        w = workloads.kooshball(device=dev)
        workload = "3D kooshball 160^3, grid 320^3, L=8, C=8, 20000x320 spokes"
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    cfg = mr.make_config(w["traj_type"], w["n_dim"], w["L"], n_coils=w["n_coils"],
                         mem_fraction=0.9)
    print(f"[bench] {workload}: inputs generated in {t_gen:.1f} s", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    if args.config == "2d72":
        if args.masked:
            cfg = mr.make_config("2D", w["n_dim"], w["L"], n_coils=w["n_coils"], is_masked=True,
                                 mem_fraction=0.9)
            mask = mr.calc_mask(cfg, w["traj"])
            cfg = mr.make_config("2D", w["n_dim"], w["L"], n_coils=w["n_coils"], is_masked=True,
                                 num_masked=int(mask.numel()), mem_fraction=0.9)
        T = w["L"] * (w["L"] + 1) // 2
        cols = mask.numel() if mask is not None else 4 * w["n_dim"][0] * w["n_dim"][1]
        K_given = torch.complex(torch.randn((T, cols), generator=g, device=dev),
                                torch.randn((T, cols), generator=g, device=dev))
        op = mr.create_op(cfg, w["SEs"], None, None, mask, _phi2_toep=K_given)
    else:
        op = mr.create_op(cfg, w["SEs"], w["phi2"], w["traj"])
    torch.cuda.synchronize()
    t_create = time.perf_counter() - t0
    info = op.info
    print(f"[bench] operator created in {t_create:.1f} s (kernel {info['toep_ms'] / 1e3:.1f} s), "
          f"group {info['group'][0]}", file=sys.stderr, flush=True)

    x = w["x"]
    out = torch.empty_like(x)
    tiles = None
    if not args.no_autotune:
        t_tune = time.perf_counter()
        tiles = op.autotune(x, reps=2)
        print(f"[bench] autotuned in {time.perf_counter() - t_tune:.1f} s: {tiles}",
              file=sys.stderr, flush=True)
    ms_per_step, wall_ms, ms, calls = time_applies(op, x, out, args.steps, args.warmup)
    value = 1000.0 / ms_per_step

    group = info["group"][0]
    xbd = int(_lib.lib.mr_op_xbd_group(op._handle))
    xbd_form = int(_lib.lib.mr_op_xbd_form(op._handle))
    t_in_use = op.tiles
    sb = stage_bytes(cfg, group, xbd, "fp32")
    stages = {}
    for i, name in enumerate(_lib.PROFILE_STAGES):
        if calls[i]:
            avg = ms[i] / calls[i]
            nb = sb.get(name, 0)
            stages[name] = {"avg_ms": round(avg, 4), "launches": int(calls[i]),
                            "gbs": round(nb / (avg * 1e6), 1) if nb else None}
    dom = max((s for s in stages if s in sb), key=lambda s: stages[s]["avg_ms"])
    ct_axis = {256, 320, 384, 512, 640, 56}
    if dom == "BD":
        dom_kernel = ({2: "mr_k_xbd_pd"}.get(xbd_form, "mr_k_xbd_ct") if xbd
                      else ("mr_k_bd" if cfg.L <= 16 else "mr_k_bdt"))
    else:
        dom_kernel = "mr_fft_axis_ct" if set(cfg.phi2_dim[1:]) <= ct_axis else "mr_fft_axis_fast"
    dom_ms = stages[dom]["avg_ms"]
    achieved = sb[dom] / (dom_ms * 1e6)
    B = apply_model_bytes(cfg)
    print(f"[bench] FP32 kernel: {value:.2f} applies/s, stages {stages}", file=sys.stderr,
          flush=True)

    # mixed precision: the same operator reading the BF16 kernel (per-4x4-block scales)
    mixed = None
    if not args.no_mixed:
        op.set_kernel_format("bf16")
        mms, _, mmsa, mcalls = time_applies(op, x, out, args.steps, args.warmup)
        op.set_kernel_format("fp32")
        mixed = {"kernel": "bf16 entries, fp32 per-4x4-block scales", "applies_per_s":
                 round(1000.0 / mms, 3), "ms_per_apply": round(mms, 3),
                 "speedup_vs_fp32": round(ms_per_step / mms, 3),
                 "fused_pass_ms": round(mmsa[3] / mcalls[3], 3) if mcalls[3] else None}
        print(f"[bench] BF16 kernel: {mixed}", file=sys.stderr, flush=True)

    # one-coil outputs of the same kernel for the error statements against the
    # FP64 oracle (coil 0: the CPU leg restricts itself to it)
    cpu, errs, oracle_info = None, None, None
    K = op.kernel(0)
    op.close()
    if threads:
        cfg1 = mr.make_config(w["traj_type"], w["n_dim"], w["L"], n_coils=1, mem_fraction=0.9)
        outs = {}
        with mr.create_op(cfg1, w["SEs"][:1].contiguous(), None, None, _phi2_toep=K) as op1:
            outs["fp32"] = op1(x).cpu().numpy()
            if not args.no_mixed:
                op1.set_kernel_format("bf16")
                outs["bf16"] = op1(x).cpu().numpy()
        del K
        torch.cuda.empty_cache()
        cpu, errs, oracle_info = cpu_leg(w, outs, threads)
        print(f"[bench] CPU leg: {cpu}; GPU one-coil errors vs FP64 oracle {errs}",
              file=sys.stderr, flush=True)
        if mixed is not None:
            mixed["error_vs_fp64"] = errs.get("bf16")
    else:
        del K
    del w, x, out
    torch.cuda.empty_cache()

    cg = None
    if not args.no_cg:
        cg = cg_leg(args, threads)
        print(f"[bench] CG leg: {cg}", file=sys.stderr, flush=True)
    srv = None
    if not args.no_serve and args.requests > 0:
        srv = serve_leg(args)

    line = {
        "metric": "A^H A applies/s @320x320x56 L16 C12 (mixed vs FP32); p50/p99 recon latency",
        "value": round(value, 3), "unit": "A^H A applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 complex)", "data": "synthetic",
        "config": {"workload": workload, "precision": "fp32 kernel", "coil_group": group,
                   "pipeline": (f"fused x pass, {xbd} coil(s) per workgroup" if xbd
                                else "separate FFT and block passes"),
                   "kernel_build_s": round(info["toep_ms"] / 1e3, 2),
                   "tiles": t_in_use if tiles is not None else f"built-in rules {t_in_use}"},
        "error_vs_fp64": errs.get("fp32") if errs else None,
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "stage": dom,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom_kernel, workload, "fp32"),
                     "algorithmic_bytes_per_launch": sb[dom],
                     # the byte model's fraction; the kernel's measured limiter
                     # (rocprofv3 counters, committed record) when there is one
                     "limiter": profile_record(dom_kernel, workload).get("limiter"),
                     "attainable_copy_gbs": COPY_CEILING_GBS},
        "apply_model": {"bytes": B, "gbs": round(B / (ms_per_step * 1e6), 1),
                        "frac": round(B / (ms_per_step * 1e6) / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "wall_ms_per_step": round(wall_ms, 3),
        "mixed": mixed,
        "cpu_baseline": cpu,
        "oracle": oracle_info,
        "cg": cg,
        "serve": srv,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
