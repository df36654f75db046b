"""Generate the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no stored vectors (its tests/make_reference.py writes
data/ref_mlx/*.npz from a CUDA run, and data/ is not in the repository), so
the fixtures here are produced by the project's oracle on the reference's own
case list and seeds (tests/make_reference.py:130-146), on the uniform-random
problems of tests/check_op.py:82-89 (2D and stack, seed-fixed), and on the
2D radial benchmark shape. Inputs are regenerated from their seeds by
``oracle.generators`` at test time; each file stores a checksum of them.

Stored per case (complex64 unless noted):
  kidx                    int64 sampled flat kernel indices (row-major)
  K_lower, K_upper        exact phi2_toep (T, len(kidx)) at those indices
  AhAx_lower, AhAx_upper  Toeplitz A^H A with the exact kernel
  AhAx_nudft              exact direct-NUDFT normal operator
  mask_sum, mask_len      int64 checksum of the radius-2 neighbourhood mask
  input_sum               float64 checksum of traj, phi2, x, SEs

Run: PYTHONPATH=. python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import generators as gen  # noqa: E402
from oracle import mask as omask  # noqa: E402
from oracle import nudft, toeplitz  # noqa: E402

OUT = Path(__file__).resolve().parent
NKIDX = 1024


def input_checksum(traj, phi2, x, SEs) -> np.ndarray:
    return np.array([float(np.sum(np.abs(a.astype(np.complex128)) * (1 + np.arange(a.size) % 7).reshape(a.shape)))
                     for a in (traj, phi2, x, SEs)])


def case_inputs(name: str):
    if name.startswith("checkop_"):
        key = name[len("checkop_"):]
        rng = np.random.default_rng(7)
        return gen.check_op_problem(key, rng)
    if name == "radial_128_L4_C4":
        g, traj, phi2, x, SEs, _ = gen.radial_workload()
        return g, traj, phi2, x, SEs
    if name in RECIPE_CASES:
        g, traj, phi2, x, SEs, _ = RECIPE_CASES[name]()
        return g, traj, phi2, x, SEs
    return gen.reference_case(name)


#: The benchmark recipes at reduced size: stack-of-stars with L=16, C=12, a
#: 200-frame orthonormal basis assigned round-robin over golden-angle spokes
#: (the 320x320x56 workload's recipe at 40x40x8, 48 spokes x 80 samples), and
#: the Fibonacci-sphere kooshball with L=8, C=8, a 100-frame basis (the 160^3
#: workload's recipe at 24^3, 400 spokes x 48 samples).
RECIPE_CASES = {
    "stack_40x40x8_L16_C12_T200": lambda: gen.stack_workload(
        n=40, nz=8, L=16, C=12, n_spokes=48, n_kx=80, n_frames=200, seed=20261205),
    "koosh_24_L8_C8_T100": lambda: gen.kooshball_workload(
        n=24, L=8, C=8, n_spokes=400, n_kx=48, n_frames=100, seed=20261206),
}

CASES = [c[0] for c in gen.REFERENCE_CASES] + ["checkop_2d", "checkop_stack",
                                               "radial_128_L4_C4"] + list(RECIPE_CASES)


def make(name: str) -> Path:
    g, traj, phi2, x, SEs = case_inputs(name)
    out = {}
    rng = np.random.default_rng(12345)
    kidx = np.sort(rng.choice(g.P, size=min(NKIDX, g.P), replace=False))
    for tri, up in (("lower", False), ("upper", True)):
        K = toeplitz.calc_kernel(g, phi2, traj, upper=up)
        out[f"K_{tri}"] = K.reshape(g.T, -1)[:, kidx].astype(np.complex64)
        out[f"AhAx_{tri}"] = toeplitz.apply(g, x, SEs, K, upper=up).astype(np.complex64)
    out["AhAx_nudft"] = nudft.aha_nudft(g, x, SEs, phi2, traj).astype(np.complex64)
    m = omask.calc_mask(g, traj, 2)
    out["mask_sum"] = np.int64(int(np.sum(m * (1 + np.arange(m.size) % 13))))
    out["mask_len"] = np.int64(m.size)
    out["kidx"] = kidx.astype(np.int64)
    out["input_sum"] = input_checksum(traj, phi2, x, SEs)
    path = OUT / f"{name}.npz"
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    names = sys.argv[1:] or CASES
    for n in names:
        p = make(n)
        print(f"wrote {p.name} ({p.stat().st_size / 1e3:.0f} kB)", flush=True)
