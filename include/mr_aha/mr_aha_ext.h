/*
 * mr_aha_ext.h — entry points of libmr_aha_c that have no counterpart in the
 * reference's C ABI. They serve measurement and the MI355X-specific
 * features; the drop-in ABI is mr_aha.h alone.
 */
#ifndef MR_AHA_MR_AHA_EXT_H
#define MR_AHA_MR_AHA_EXT_H

#include "mr_aha/mr_aha.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Stages timed when MR_AHA_HIP_PROFILE=1 is set at execute time. */
#define MR_PROFILE_STAGES 8  /* F1 F2 F3 BD I3 I2 I1 other */

/* GPU milliseconds and launch counts per stage accumulated by the plan's
 * executions since the previous call, then reset. Synchronises with the
 * last recorded stage. No counterpart in the reference (its
 * MR_AHA_CUDA_PROFILE prints totals to stderr, src/cuda/mr_aha.cu:516-623);
 * for benchmarks that need per-kernel durations. */
MR_EXPORT int mr_aha_plan_profile(mr_aha_plan plan, double *ms, int *calls, int n);
/* Same for the plan inside a Toeplitz operator. */
MR_EXPORT int mr_op_profile(mr_op op, double *ms, int *calls, int n);

/* Apply pipeline chosen by a plan: the number of coils per workgroup of the
 * fused x pass (forward x-FFT, block multiply and inverse x-FFT of one
 * k-space row in one kernel), or 0 when the plan runs separate FFT and
 * block passes. With the fused pass the stages timed as F3/F2/BD/I2/I3 are
 * the outer-axis FFT, the y-FFT, the fused x pass, the inverse y-FFT and the
 * inverse outer-axis FFT (2D: F2 = y-FFT with the sensitivity product,
 * I2 = inverse y-FFT with the coil sum). Negative on error. No counterpart
 * in the reference. */
MR_EXPORT int mr_aha_plan_xbd_group(mr_aha_plan plan);
MR_EXPORT int mr_op_xbd_group(mr_op op);

/* Form of the fused x pass a plan launches: 0 = separate passes (no fused
 * pass), 1 = one workgroup per (k-space row, coil group), 2 = the persistent
 * form (one workgroup per CU walking its items, the next row's image loaded
 * into LDS while the current one finishes), 3 = the block step on the matrix
 * cores (for a 16-byte aligned kernel). Negative on error. No counterpart in
 * the reference. */
MR_EXPORT int mr_aha_plan_xbd_form(mr_aha_plan plan);
MR_EXPORT int mr_op_xbd_form(mr_op op);

/* Autotuner of the fused route (no counterpart in the reference, whose cuFFT
 * plans fix their own tiling). For each axis pass of the apply (outer-axis
 * FFT, y-FFT, inverse y-FFT, inverse outer-axis FFT; 2D: the first and last)
 * the tile widths 1, 2, ..., 64 pencils per workgroup are timed on the
 * caller's data, `reps` applies each (1..100), and the fastest is kept for
 * later executes; then the two-coil form of the fused x pass is timed
 * against one coil per workgroup. Results do not depend on the choice. The
 * arrays are those of mr_aha_plan_execute / mr_op_execute; d_AhAx receives
 * the last trial's output. Synchronous. Clears the per-stage profile. A plan
 * on separate FFT passes is left unchanged. MR_AHA_HIP_{FO,FY,IY,IO}_TX
 * environment settings still override the tuned widths. */
MR_EXPORT int mr_aha_plan_autotune(mr_aha_plan plan, const mr_complex *d_x,
                                   const mr_complex *d_SEs, const mr_complex *d_phi2_toep,
                                   const int64_t *d_mask, mr_complex *d_AhAx, int reps);
MR_EXPORT int mr_op_autotune(mr_op op, const mr_complex *d_x, mr_complex *d_AhAx, int reps);

/* Tiling in use: tiles[0..3] = pencils per workgroup requested for the
 * outer-axis, y, inverse y and inverse outer-axis passes (0: the built-in
 * rule), tiles[4] = 1 when the fused x pass runs two-coil 1024-thread
 * workgroups, tiles[5] = 1 when its block step runs on the matrix cores (for
 * a 16-byte aligned kernel). Writes min(n, 6) entries. No counterpart in the
 * reference. */
MR_EXPORT int mr_aha_plan_tiles(mr_aha_plan plan, int *tiles, int n);
MR_EXPORT int mr_op_tiles(mr_op op, int *tiles, int n);

/* Conjugate gradients on (A^H A + lambda I) x = b with operator `op`
 * (FP32 vectors; FP64 inner products and step lengths, kept on the device).
 * Every `refresh` iterations (0 = never) the recursive residual is replaced
 * by the true residual b - (A^H A + lambda I) x, computed with `op_refresh`
 * (same shape, device and stream; e.g. the FP32 operator when `op` holds a
 * 16-bit kernel) or with `op` when op_refresh is NULL. A refresh through a
 * different operator also restarts the search directions, so each cycle of
 * `refresh` iterations solves with `op` for the correction of the residual
 * of `op_refresh` (iterative refinement). d_x holds the initial
 * guess (or is zeroed with MR_CG_ZERO_INIT) and receives the solution.
 * res_hist (host, optional) receives ||r_k|| / ||b|| for k = 0..n_iter, r_k
 * being the residual the iteration carries. MR_CG_GRAPH captures one
 * ordinary and one refresh iteration as HIP graphs and replays them (needs a
 * non-default cfg.stream). Synchronises with the stream before returning.
 * No counterpart in the reference: its callers run CG on top of
 * mr_op_execute (reference README.md:70-73). */
/* Kernel storage format of a Toeplitz operator's applies. MR_KERNEL_BF16 and
 * MR_KERNEL_FP16 quantise the operator's FP32 kernel once (each Hermitian
 * entry as two 16-bit values scaled by the largest component of its 4x4
 * block of the L x L matrix, one FP32 scale per block and location) and the
 * fused x pass reads that instead: half the kernel bytes per apply, at the
 * precision of the format (relative entry error 2^-9 BF16, 2^-11 FP16). The
 * FP32 kernel is kept: MR_KERNEL_FP32 switches back. Available where the
 * fused x pass is compiled for the shape (x grid 256, 320, 384, 512 or 640
 * with L = 4, 8 or 16; unmasked); MR_ERR_UNSUPPORTED otherwise. Pair with
 * mr_op_cg's op_refresh = an FP32 operator for iterative refinement. No
 * counterpart in the reference. */
#define MR_KERNEL_FP32 0
#define MR_KERNEL_BF16 1
#define MR_KERNEL_FP16 2
MR_EXPORT int mr_op_set_kernel_format(mr_op op, int format);

#define MR_CG_ZERO_INIT 1
#define MR_CG_GRAPH 2
MR_EXPORT int mr_op_cg(mr_op op, mr_op op_refresh, const mr_complex *d_b, mr_complex *d_x,
                       double lambda, int n_iter, int refresh, int flags, double *res_hist);

/* mr_op_cg without any allocation, for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, kernels and results as mr_op_cg; its work vectors (3 n complex
 * values), reduction partials and residual history live in a static 64 MiB
 * device array, so n and n_iter are limited to
 * 3 * ceil256(8 n) + 8192 + 256 + ceil256(8 (n_iter + 2)) <= 64 MiB
 * (about 2.7 M unknowns); larger problems fail with MR_ERR_INVALID_ARG.
 * MR_CG_GRAPH is rejected (graph instantiation allocates). Without res_hist
 * the call returns with the solve queued on the operator's stream. Every
 * call on one device shares that device's array: calls that overlap on one
 * GPU (other streams, other threads) race. No counterpart in the
 * reference. */
MR_EXPORT int mr_op_cg_noalloc(mr_op op, mr_op op_refresh, const mr_complex *d_b,
                               mr_complex *d_x, double lambda, int n_iter, int refresh,
                               int flags, double *res_hist);

/* Streaming copy of `bytes` (a multiple of 16, 16-byte aligned buffers) on
 * `stream`: 16-byte non-temporal loads and stores, the attainable HBM
 * bandwidth that the passes' achieved bandwidth is compared with (benchmark
 * calibration). No counterpart in the reference. */
MR_EXPORT int mr_stream_copy(const void *d_src, void *d_dst, int64_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif
