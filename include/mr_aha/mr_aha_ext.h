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

#ifdef __cplusplus
}
#endif

#endif
