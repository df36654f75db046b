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
    """Plain device copy of a known size, the reference point for scaling the
    rocprofv3 FETCH_SIZE / WRITE_SIZE readings of the library's kernels."""
    import torch
    x = torch.empty(nbytes // 8, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": reps}),
          flush=True)


def measured_traffic(kernel, workload, precision="fp32"):
    """HBM bytes per launch of `kernel` on `workload` with the kernel stored
    in `precision`, from the committed rocprofv3 record
    (profiles/hbm_traffic.json), or None."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return None
    for rec in json.loads(path.read_text()).get("records", []):
        if (rec.get("kernel") == kernel and rec.get("workload") == workload
                and rec.get("precision", "fp32") == precision):
            return int(rec["hbm_bytes_per_launch"])
    return None


def cpu_baseline(w, K, cfg, budget_s=30.0):
    """Oracle CPU Toeplitz apply on one coil of the workload (complex64)."""
    import numpy as np
    from oracle import cpu_baseline as cb
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(int(threads), 16))
    x = w["x"].cpu().numpy()
    S = w["SEs"][:1].cpu().numpy()
    Kh = K.cpu().numpy()
    n = list(cfg.n_dim)
    M = list(cfg.phi2_dim)
    t0 = time.perf_counter()
    cb.toeplitz_apply_coils(n, M, cfg.L, x, S, Kh, cfg.triangulation == 1, [0], threads)
    dt = time.perf_counter() - t0
    del Kh
    C = cfg.n_coils
    return {"value": 1.0 / (dt * C), "unit": "A^H A applies/s", "cores": threads, "kind": "port",
            "sample": f"1 of {C} coils of one apply (all {cfg.L} ranks, full grid), "
                      f"{dt:.1f} s, scaled x{C}"}


def bench_cg(args):
    """CG reconstructions of the 128x128 L=4 C=4 radial problem: 40 iterations,
    true residual every 10, FP32 operator vs 16-bit kernel refined against it,
    HIP-graph-replayed iterations. Error of each against the FP32 solution
    converged to a 1e-6 relative residual (200 iterations)."""
    import torch
    import mr_aha_amd as mr
    from mr_aha_amd import workloads
    w = workloads.radial_2d(device="cuda")
    s = torch.cuda.Stream()
    cfg = mr.make_config(w["traj_type"], w["n_dim"], w["L"], n_coils=w["n_coils"], stream=s)
    res = {}
    with torch.cuda.stream(s):
        K = mr.calc_toep(cfg, w["phi2"], w["traj"])
        op32 = mr.create_op(cfg, w["SEs"], None, None, _phi2_toep=K)
        op16 = mr.create_op(cfg, w["SEs"], None, None, _phi2_toep=K)
        op16.set_kernel_format("bf16")
        # smooth object: low-pass complex Gaussian coefficients
        xt = workloads.complex_gaussian(tuple(w["x"].shape), 77, "cuda")
        f = torch.fft.fftn(xt, dim=(-2, -1))
        ky = torch.fft.fftfreq(xt.shape[-2], device="cuda").abs()[:, None] < 0.1
        kx = torch.fft.fftfreq(xt.shape[-1], device="cuda").abs()[None, :] < 0.1
        xt = torch.fft.ifftn(f * (ky & kx), dim=(-2, -1)).to(torch.complex64).contiguous()
        Axt = op32(xt)
        lam = 1e-3 * float(torch.linalg.vector_norm(Axt) / torch.linalg.vector_norm(xt))
        b = (Axt + lam * xt).contiguous()
        xref, href = op32.cg(b, lam=lam, iters=200, refresh=10, graph=True)
        for name, op, ref in (("fp32", op32, None), ("bf16", op16, op32)):
            for _ in range(args.warmup):
                op.cg(b, lam=lam, iters=40, refresh=10, op_refresh=ref, graph=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                x, h = op.cg(b, lam=lam, iters=40, refresh=10, op_refresh=ref, graph=True)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            err = float(torch.linalg.vector_norm(x - xref) / torch.linalg.vector_norm(xref))
            res[name] = {"recon_per_s": round(1.0 / dt, 3), "ms_per_recon": round(dt * 1e3, 3),
                         "final_residual": h[-1], "error_vs_converged_fp32": err}
        op16.close()
        op32.close()
    line = {"metric": "CG reconstructions/s, 128x128 L4 C4 radial, 40 iterations (FP32 vs bf16 "
                      "kernel refined against FP32)",
            "value": res["bf16"]["recon_per_s"], "unit": "reconstructions/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "dtype": "c64 vectors, bf16 kernel, fp64 CG scalars", "data": "synthetic",
            "config": {"workload": "2D radial 128x128, grid 256x256, L=4, C=4, 64x256 spokes; "
                                   "CG 40 iterations, refresh 10, lambda 1e-3 |A|"},
            "fp32": res["fp32"], "bf16": res["bf16"],
            "speedup_mixed_vs_fp32": round(res["bf16"]["recon_per_s"] / res["fp32"]["recon_per_s"], 3),
            "reference_residual": href[-1]}
    print(json.dumps(line), flush=True)


def bench_serve(args):
    """Poisson mix of CG reconstruction requests (shapes of the serving
    configuration) through mr_aha_amd.serve, FP32 and then 16-bit kernels."""
    import random
    import torch
    from mr_aha_amd import serve, workloads
    rnd = random.Random(2026)
    specs = []
    t = 0.0
    for i in range(args.requests):
        t += rnd.expovariate(4.0)               # 4 requests/s on average
        if rnd.random() < 0.5:
            n = rnd.choice([128, 192, 256, 320])
            spec = ("2D", (n, n))
        else:
            n = rnd.choice([192, 256])
            spec = ("stack", (n, n, rnd.choice([32, 48])))
        L = rnd.choice([4, 8, 16])
        C = rnd.choice([8, 12, 16])
        ns = rnd.randint(256, 512)
        specs.append((i, spec, L, C, ns, rnd.randint(20, 50), t))

    def make(prec):
        reqs = []
        for i, (tt, nd), L, C, ns, it, arr in specs:
            n = nd[0]
            if tt == "2D":
                w = workloads.radial_2d(n=n, L=L, C=C, n_spokes=ns, nkx=2 * n, n_frames=64,
                                        seed=900 + i)
            else:
                w = workloads.stack_of_stars(n=n, nz=nd[2], L=L, C=C, n_spokes=ns, nkx=2 * n,
                                             n_frames=64, seed=900 + i)
            b = workloads.complex_gaussian(tuple(w["x"].shape), 1900 + i, "cuda")
            reqs.append(serve.Request(rid=i, traj_type=tt, n_dim=nd, L=L, n_coils=C,
                                      SEs=w["SEs"], phi2=w["phi2"], traj=w["traj"], b=b,
                                      iters=it, lam=1e-2, refresh=10, precision=prec,
                                      arrival=arr))
        return reqs

    out = {}
    for prec in ("fp32", "bf16"):
        reqs = make(prec)
        torch.cuda.synchronize()
        done = serve.Server(slots=4).run(reqs, realtime=True)
        out[prec] = serve.summary(done)
        out[prec]["errors"] = sorted({r.error for r in done if r.error})[:3]
        res = sorted(r.hist[-1] for r in done if r.error is None)
        out[prec]["median_final_residual"] = res[len(res) // 2] if res else None
        print(f"[serve] {prec}: {out[prec]}", file=sys.stderr, flush=True)
        del reqs, done
        torch.cuda.empty_cache()
    line = {"metric": "serving: CG reconstruction requests/s and p50/p99 latency (Poisson mix)",
            "value": round(out["fp32"]["requests_per_s"], 3), "unit": "requests/s", "n_gpus": 1,
            "higher_is_better": True, "data": "synthetic",
            "config": {"workload": f"{args.requests} Poisson requests at 4/s: 2D radial "
                                   "{128,192,256,320}^2 or stack {192,256}^2 x {32,48}; L {4,8,16}; "
                                   "C {8,12,16}; 256-512 spokes of 2N; 20-50 CG iterations; 4 slots"},
            "fp32": out["fp32"], "bf16": out["bf16"]}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="stack", choices=["stack", "radial", "koosh", "cg", "serve"],
                    help="stack/radial/koosh: applies/s; cg: CG reconstructions/s, FP32 vs "
                         "16-bit kernel; serve: request mix, p50/p99 latency")
    ap.add_argument("--requests", type=int, default=24, help="serve: number of requests")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in tile rules instead of timing the candidates "
                         "(mr_op_autotune) before the warm-up")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16"],
                    help="kernel storage read by the applies (16-bit: per-4x4-block scales)")
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
    if args.config == "cg":
        return bench_cg(args)
    if args.config == "serve":
        return bench_serve(args)
    t_gen = time.perf_counter()
    if args.config == "stack":
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
    op = mr.create_op(cfg, w["SEs"], w["phi2"], w["traj"])
    torch.cuda.synchronize()
    t_create = time.perf_counter() - t0
    info = op.info
    print(f"[bench] operator created in {t_create:.1f} s (kernel {info['toep_ms'] / 1e3:.1f} s), "
          f"group {info['group'][0]}", file=sys.stderr, flush=True)

    if args.precision != "fp32":
        op.set_kernel_format(args.precision)
    x = w["x"]
    out = torch.empty_like(x)
    tiles = None
    if not args.no_autotune:
        t_tune = time.perf_counter()
        tiles = op.autotune(x, reps=2)
        print(f"[bench] autotuned in {time.perf_counter() - t_tune:.1f} s: {tiles}",
              file=sys.stderr, flush=True)
    for _ in range(args.warmup):
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
    for _ in range(args.steps):
        op.apply(x, out)
    e1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    os.environ.pop("MR_AHA_HIP_PROFILE")
    _lib.check(_lib.lib.mr_op_profile(op._handle, ms, calls, 8), "mr_op_profile")
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    value = 1000.0 / ms_per_step

    group = info["group"][0]
    xbd = int(_lib.lib.mr_op_xbd_group(op._handle))
    sb = stage_bytes(cfg, group, xbd, args.precision)
    stages = {}
    for i, name in enumerate(_lib.PROFILE_STAGES):
        if calls[i]:
            avg = ms[i] / calls[i]
            nb = sb.get(name, 0)
            stages[name] = {"avg_ms": round(avg, 4), "launches": int(calls[i]),
                            "gbs": round(nb / (avg * 1e6), 1) if nb else None}
    dom = max((s for s in stages if s in sb), key=lambda s: stages[s]["avg_ms"])
    dom_kernel = ({"BD": "mr_k_xbd_ct" if xbd else "mr_k_bd"}.get(dom, "mr_fft_axis_fast"))
    dom_ms = stages[dom]["avg_ms"]
    achieved = sb[dom] / (dom_ms * 1e6)
    B = apply_model_bytes(cfg)

    cpu = None
    if not args.no_cpu:
        K = op.kernel(0)
        cpu = cpu_baseline(w, K, cfg)
        del K
    op.close()

    line = {
        "metric": "A^H A applies/s @320x320x56 L16 C12 (mixed vs FP32); p50/p99 recon latency",
        "value": round(value, 3), "unit": "A^H A applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 complex)" if args.precision == "fp32"
                 else f"c64 vectors, {args.precision} kernel",
        "data": "synthetic",
        "config": {"workload": workload, "precision": f"{args.precision} kernel",
                   "coil_group": group,
                   "pipeline": (f"fused x pass, {xbd} coil(s) per workgroup" if xbd
                                else "separate FFT and block passes"),
                   "kernel_build_s": round(info["toep_ms"] / 1e3, 2),
                   "tiles": tiles if tiles is not None else "built-in rules"},
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "stage": dom,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom_kernel, workload, args.precision),
                     "algorithmic_bytes_per_launch": sb[dom]},
        "apply_model": {"bytes": B, "gbs": round(B / (ms_per_step * 1e6), 1),
                        "frac": round(B / (ms_per_step * 1e6) / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "wall_ms_per_step": round(t_wall * 1000 / args.steps, 3),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
