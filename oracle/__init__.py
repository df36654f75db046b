"""CPU oracle for the Toeplitz A^H A path — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy FP64 (and, for the CPU baseline, plain C),
what the reference library computes on its ``method="toeplitz"`` path:

* ``conventions``  index maps, triangle packing, centering
  (include/mr_aha/mr_aha.h:19-92, 299-317; src/cuda/mr_internal.cuh:119-181)
* ``generators``   seeded inputs restating tests/make_reference.py:160-236,
  tests/check_op.py:125-172 and the benchmark configurations
* ``toeplitz``     the exact kernel ``phi2_toep`` by direct Dirichlet sums
  (src/cuda/mr_calc_toep.cu:10-24, 171-200, 537-618) and the five-stage
  apply (src/cuda/mr_aha.cu:694-790, src/cuda/mr_kernels.cu:73-176)
* ``nudft``        the direct (exact) non-uniform DFT normal operator that
  ``method="nufft"`` approximates (src/cuda/mr_nufft_op.cu:613-700,
  python/mr_aha/nufft.py:473-552), and a dense brute-force forward model
* ``cpu_nufft``    the FINUFFT-equivalent CPU A^H A used as the timed CPU
  baseline (exponential-of-semicircle spreading, C + OpenMP)

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as a checker or a timed baseline.
The product (``mr_aha_amd`` and ``libmr_aha_c.so``) never imports or links it.

Parity status of each function is stated in its docstring ("pinned by ...").
"""
