"""Reconstruction serving on one GPU: continuous batching of CG requests.

Requests of mixed matrix size, rank and coil count arrive over time. Each
needs an operator (its Toeplitz kernel built on the GPU), a CG solve on the
normal equations and its memory until it finishes. The server:

* admits pending requests in arrival order while a slot is free and their
  HBM footprint (kernel, 16-bit copy, plan work buffers, the kernel build's
  gridding chunks, CG vectors, maps; each an upper bound of what the library
  allocates for them) fits the remaining budget, so requests that arrive
  while others run start as soon as memory and a slot free up;
* runs every admitted request on its own non-blocking HIP stream from a
  worker thread (the library calls release the GIL), so kernel builds,
  applies and CG iterations of different requests overlap on the device;
* solves with mr_op_cg: FP64 scalars on the device, HIP-graph-replayed
  iterations, and for a 16-bit kernel ("bf16"/"fp16") cycles restarted from
  the true residual of an FP32 operator sharing the same kernel memory;
* records per request the arrival, admission and completion times, from
  which latency percentiles and throughput follow.

Nothing here has a counterpart in the reference, whose callers run one
reconstruction at a time on top of op(x) (README.md:70-73 of the reference).
"""

from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import torch

from . import _hiprt, api


@dataclass
class Request:
    rid: int
    traj_type: str
    n_dim: tuple
    L: int
    n_coils: int
    SEs: torch.Tensor
    phi2: torch.Tensor
    traj: torch.Tensor
    b: torch.Tensor
    iters: int = 30
    lam: float = 0.0
    refresh: int = 10
    precision: str = "fp32"
    arrival: float = 0.0          # seconds after the start of the run
    # results
    x: Optional[torch.Tensor] = None
    hist: List[float] = field(default_factory=list)
    t_admit: float = 0.0
    t_done: float = 0.0
    # device time of the request's first and last work on its stream (HIP
    # events), seconds after the run's reference event
    d_start: float = 0.0
    d_end: float = 0.0
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


# Device memory the kernel build of one request may take for its gridding
# chunks. calc_toep sizes its chunk as at most half of mem_fraction times the
# free memory, so a build configuration with mem_fraction = 2 BUILD_BYTES /
# total memory keeps the chunk within BUILD_BYTES whatever is free (or one
# component, when a single component is larger; build_bytes counts that).
BUILD_BYTES = 1 << 30


def _build_component_bytes(req: Request) -> int:
    """Upper estimate of one gridded component: the 2x upsampled fine grid
    (rounded up to whole bins) and the natural-order grid, complex128."""
    nsp = 2 if req.traj_type == "2D" else 3
    n = list(req.n_dim) + [1] * (3 - nsp)
    M = [2 * v for v in n[:nsp]] + [1] * (3 - nsp)
    if req.traj_type == "stack":
        M[2] = n[2]
    tr = 3 if req.traj_type == "3D" else 2
    fine = 1
    for a in range(tr):
        fine *= 2 * M[a] + 64
    return (fine + M[0] * M[1] * (M[2] if tr == 3 else 1)) * 16


def build_bytes(req: Request) -> int:
    return max(BUILD_BYTES, _build_component_bytes(req))


def footprint_bytes(req: Request) -> int:
    """Upper estimate of the request's device memory while it runs."""
    nsp = 2 if req.traj_type == "2D" else 3
    n = list(req.n_dim) + [1] * (3 - nsp)
    M = [2 * v for v in n[:nsp]] + [1] * (3 - nsp)
    if req.traj_type == "stack":
        M[2] = n[2]
    N = n[0] * n[1] * n[2]
    P = M[0] * M[1] * M[2]
    L, C = req.L, req.n_coils
    T = L * (L + 1) // 2
    kern = T * P * 8
    if req.precision != "fp32":
        nb = ((L + 3) // 4) * ((L + 3) // 4 + 1) // 2
        kern += (T + nb) * P * 4
    work = (n[0] * n[1] * M[2] + n[0] * M[1] * M[2]) * L * C * 8   # fused-pass buffers
    if req.precision != "fp32":
        work *= 2   # the 16-bit operator and its FP32 refresh operator: two plans
    vecs = 5 * N * L * 8 + N * C * 8
    return int(kern + work + build_bytes(req) + vecs)


class Server:
    """Continuous-batching CG server (see the module docstring)."""

    def __init__(self, budget_bytes: Optional[int] = None, slots: int = 4, graph: bool = True,
                 device: str = "cuda"):
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            budget_bytes = int(0.8 * free)
        self.budget = int(budget_bytes)
        self.slots = int(slots)
        self.graph = bool(graph)
        self.device = device
        self._lock = threading.Condition()
        self._used = 0
        self._running = 0

    def _solve(self, req: Request, t0: float, ref_ev) -> None:
        # a non-blocking stream: the graph capture of one request must not be
        # joined by legacy-stream work that another request's thread issues
        handle = _hiprt.stream_create_nonblocking()
        stream = torch.cuda.ExternalStream(handle)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        try:
            with torch.cuda.stream(stream):
                cfg = api.make_config(req.traj_type, req.n_dim, req.L, n_coils=req.n_coils,
                                      stream=stream, mem_fraction=0.25)
                # the kernel build's chunk buffers within build_bytes (see BUILD_BYTES)
                _, total = torch.cuda.mem_get_info()
                cfg_build = api.make_config(req.traj_type, req.n_dim, req.L, n_coils=req.n_coils,
                                            stream=stream,
                                            mem_fraction=min(0.9, 2.0 * BUILD_BYTES / total))
                K = api.calc_toep(cfg_build, req.phi2, req.traj)
                op = api.create_op(cfg, req.SEs, None, None, _phi2_toep=K)
                op_ref = None
                try:
                    if req.precision != "fp32":
                        op_ref = op
                        op = api.create_op(cfg, req.SEs, None, None, _phi2_toep=K)
                        op.set_kernel_format(req.precision)
                    x, hist = op.cg(req.b, lam=req.lam, iters=req.iters, refresh=req.refresh,
                                    op_refresh=op_ref, graph=self.graph)
                    stream.synchronize()
                    req.x, req.hist = x, hist
                finally:
                    op.close()
                    if op_ref is not None:
                        op_ref.close()
                    del K
        except Exception as e:  # reported per request
            req.error = f"{type(e).__name__}: {e}"
        finally:
            ev1.record(stream)
            stream.synchronize()
            req.d_start = ref_ev.elapsed_time(ev0) / 1e3
            req.d_end = ref_ev.elapsed_time(ev1) / 1e3
            _hiprt.stream_destroy(handle)
        req.t_done = time.perf_counter() - t0

    def run(self, requests: List[Request], realtime: bool = True) -> List[Request]:
        """Serve `requests` (arrival times in seconds; realtime=False admits
        them all at t = 0) and return them with their results and times."""
        torch.cuda.synchronize()   # inputs written on other streams are complete
        order = sorted(requests, key=lambda r: r.arrival)
        for r in order:
            if footprint_bytes(r) > self.budget:
                r.error = f"footprint {footprint_bytes(r)} B exceeds the budget {self.budget} B"
        pending = [r for r in order if r.error is None]
        ref_ev = torch.cuda.Event(enable_timing=True)
        ref_ev.record()
        t0 = time.perf_counter()
        done = threading.Semaphore(0)

        def finish(req):
            # the request's torch-allocated arrays (its kernel above all) go
            # back to the device, not to torch's cache: the library allocates
            # its plans, 16-bit kernels and CG vectors with hipMalloc, which
            # cannot use cached blocks, and the budget assumes they are free
            torch.cuda.empty_cache()
            with self._lock:
                self._used -= footprint_bytes(req)
                self._running -= 1
                self._lock.notify_all()
            done.release()

        with ThreadPoolExecutor(max_workers=self.slots) as pool:
            admitted = 0
            while pending:
                now = time.perf_counter() - t0
                head = pending[0]
                if realtime and head.arrival > now:
                    time.sleep(min(head.arrival - now, 0.01))
                    continue
                need = footprint_bytes(head)
                with self._lock:
                    while self._running >= self.slots or self._used + need > self.budget:
                        self._lock.wait(timeout=0.05)
                    self._used += need
                    self._running += 1
                pending.pop(0)
                head.t_admit = time.perf_counter() - t0
                if not realtime:
                    head.arrival = 0.0
                fut = pool.submit(self._solve, head, t0, ref_ev)
                fut.add_done_callback(lambda f, r=head: finish(r))
                admitted += 1
            for _ in range(admitted):
                done.acquire()
        return order


def percentile(vals, q):
    v = sorted(vals)
    if not v:
        return float("nan")
    k = (len(v) - 1) * q / 100.0
    lo = int(k)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def device_overlap(reqs: List[Request]) -> float:
    """Fraction of the device-busy time (the union of the requests' device
    intervals) during which two or more requests had work on the device."""
    ev = sorted([(r.d_start, 1) for r in reqs] + [(r.d_end, -1) for r in reqs])
    busy = multi = 0.0
    active = 0
    last = ev[0][0] if ev else 0.0
    for t, d in ev:
        if active >= 1:
            busy += t - last
        if active >= 2:
            multi += t - last
        active += d
        last = t
    return multi / busy if busy > 0 else 0.0


def summary(reqs: List[Request]) -> dict:
    ok = [r for r in reqs if r.error is None]
    lat = [r.latency for r in ok]
    span = max(r.t_done for r in ok) - min(r.arrival for r in ok) if ok else float("nan")
    return dict(requests=len(reqs), completed=len(ok), failed=len(reqs) - len(ok),
                requests_per_s=len(ok) / span if ok and span > 0 else float("nan"),
                p50_s=percentile(lat, 50), p99_s=percentile(lat, 99),
                mean_s=sum(lat) / len(lat) if lat else float("nan"),
                device_overlap=device_overlap(ok))
