"""CPU oracle for the Toeplitz A^H A path — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy FP64 (complex64 for the timed CPU baseline),
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
* ``mask``         the neighbourhood mask of the trajectory, integer-exact
  (src/cuda/mr_calc_mask.cu:1-438)
* ``cpu_baseline`` the five-stage Toeplitz apply in complex64 with threaded
  pocketfft (scipy.fft), timed as the CPU baseline of the benchmark

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as a checker or a timed baseline.
The product (``mr_aha_amd`` and ``libmr_aha_c.so``) never imports or links it.

What pins each function (tests/test_oracle.py unless named), chosen so that a
plausible slip (a conjugated or transposed block, a flipped sign, an
off-by-one centering, a dropped scale) fails at least one of them:

  conventions.num_tri, tri_pairs, tri_offset
      known answers (num_tri(16) = 136, phi2_toep size of the 320x320x56 L16
      shape), the literal L = 3 packing fixture of python/mr_aha/_psf.py:104-105,
      the closed-form offsets of include/mr_aha/mr_aha.h:307-309
  conventions.cent2nat, nat2cent, pad_index
      round trips and hand-worked values for even and odd M
  conventions.hermitian_block, gram_blocks
      upper == lower apply; the explicit forward model A_c[m,(l,r)] =
      Phi(l,t_m) e^{-i r.x_m} S_c(r) (dense brute force); a transposed block
      is planted and must fail
  toeplitz.dirichlet_matrix
      closed form sin(M u/2)/sin(u/2) with the centered phase at
      non-singular points
  toeplitz.strengths, calc_kernel, apply
      Toeplitz embedding == exact direct NUDFT normal operator (both
      triangles, 2D/3D/stack, odd grids); full Cartesian sampling closed form
      (prod n_tr) sum_c |S_c|^2 (x Phi2); L = 1 unit Gram == PSF convolution;
      Hermitian and PSD; masked kernel == unmasked gathered, full mask ==
      identity; golden regression of the committed fixtures
  nudft.aha_nudft, dense_aha
      against each other (1e-12) and the closed form above; planted
      conjugated Gram and flipped trajectory sign must fail
  mask.calc_mask
      one sample marks the closed-form (2r+1)^d box; brute force over grid
      points on random trajectories
  generators
      shapes, ranges and Hermitian PSD Gram blocks of the seeded draws
  cpu_baseline.toeplitz_apply_coils
      equals toeplitz.apply (complex64 tolerance)

Parity unpinned: none of the functions above. What no check here can pin is
the reference's own FP32 output (its cufinufft/cuFFT rounding): the reference
cannot run in this environment, so parity with it is to tolerance, through
these exact FP64 restatements.
"""
