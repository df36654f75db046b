"""Build the oracle's C++ spreader for the CPU NUFFT baseline (g++, OpenMP).

Test infrastructure: output in oracle/_build/ (git-ignored, travels with the
tree). Two variants: x86-64-v3 (AVX2/FMA, chosen at load time when the host
has them) and the baseline x86-64 instruction set.
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"


def build(verbose: bool = False):
    OUT.mkdir(exist_ok=True)
    src = HERE / "cpu_nufft.cpp"
    outs = []
    for name, arch in (("libmr_cpu_nufft_v3.so", ["-march=x86-64-v3"]), ("libmr_cpu_nufft.so", [])):
        lib = OUT / name
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", *arch, str(src), "-o",
               str(lib)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stderr}")
        outs.append(lib)
        if verbose:
            print(f"built {lib}")
    return outs


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
