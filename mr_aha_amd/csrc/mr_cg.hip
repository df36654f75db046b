// Conjugate gradients on the normal equations (A^H A + lambda I) x = b with a
// Toeplitz (or NUFFT) operator: the solver callers of the reference run on
// top of op(x) (README.md:70-73 of the reference). Not part of the
// reference's library.
//
// Precision: vectors and the operator are FP32. The inner products and the
// step lengths are FP64 (per-block FP64 partial sums, one FP64 reduction),
// and every `refresh` iterations the recursive residual is replaced by the
// true residual b - (A^H A + lambda I) x, computed with `op_refresh` (an
// operator of higher precision, for iterative refinement of a mixed-precision
// operator) or `op` itself (residual replacement, van der Vorst & Ye 2000).
//
// All scalars stay on the device: the step lengths are formed by the kernel
// that finishes each reduction, so an iteration issues no host
// synchronisation and can be captured in a HIP graph (graph = 1: one graph
// for an ordinary iteration, one for a refresh iteration, each replayed).

#include "mr_hip.hpp"
#include "mr_aha/mr_aha_ext.h"

#include <cmath>
#include <mutex>

int mr_op_stream_len(mr_op op, hipStream_t *s, long long *n, int *device);

namespace {

constexpr int CG_THREADS = 256;
constexpr int CG_BLOCKS = 512;  // partial sums per reduction

// scalar slots
enum { S_RR, S_PAP, S_ALPHA, S_BETA, S_BB, S_N };

__device__ __forceinline__ void block_sum2(double &sx, double &sy, double2 *out)
{
    __shared__ double bx[CG_THREADS], by[CG_THREADS];
    bx[threadIdx.x] = sx;
    by[threadIdx.x] = sy;
    __syncthreads();
    for (int h = CG_THREADS / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            bx[threadIdx.x] += bx[threadIdx.x + h];
            by[threadIdx.x] += by[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = make_double2(bx[0], by[0]);
}

// part[block] = sum conj(a) b over the block's grid-stride share (FP64)
__global__ void __launch_bounds__(CG_THREADS)
k_dot(const float2 *__restrict__ a, const float2 *__restrict__ b, long long n, double2 *part)
{
    double sx = 0.0, sy = 0.0;
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS) {
        const float2 u = a[i], v = b[i];
        sx += (double)u.x * v.x + (double)u.y * v.y;
        sy += (double)u.x * v.y - (double)u.y * v.x;
    }
    block_sum2(sx, sy, part + blockIdx.x);
}

// Ap <- Ap + lambda p; part = sum conj(p) Ap
__global__ void __launch_bounds__(CG_THREADS)
k_lam_dot(const float2 *__restrict__ p, float2 *__restrict__ Ap, float lam, long long n,
          double2 *part)
{
    double sx = 0.0, sy = 0.0;
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS) {
        const float2 u = p[i];
        float2 v = Ap[i];
        v.x = __builtin_fmaf(lam, u.x, v.x);
        v.y = __builtin_fmaf(lam, u.y, v.y);
        Ap[i] = v;
        sx += (double)u.x * v.x + (double)u.y * v.y;
        sy += (double)u.x * v.y - (double)u.y * v.x;
    }
    block_sum2(sx, sy, part + blockIdx.x);
}

// x <- x + alpha p; r <- r - alpha Ap (UPDATE_R) ; part = |r|^2
template <bool UPDATE_R>
__global__ void __launch_bounds__(CG_THREADS)
k_step(float2 *__restrict__ x, float2 *__restrict__ r, const float2 *__restrict__ p,
       const float2 *__restrict__ Ap, const double *scal, long long n, double2 *part)
{
    const float alpha = (float)scal[S_ALPHA];
    double sx = 0.0;
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS) {
        const float2 u = p[i];
        float2 xi = x[i];
        xi.x = __builtin_fmaf(alpha, u.x, xi.x);
        xi.y = __builtin_fmaf(alpha, u.y, xi.y);
        x[i] = xi;
        if (UPDATE_R) {
            const float2 w = Ap[i];
            float2 ri = r[i];
            ri.x = __builtin_fmaf(-alpha, w.x, ri.x);
            ri.y = __builtin_fmaf(-alpha, w.y, ri.y);
            r[i] = ri;
            sx += (double)ri.x * ri.x + (double)ri.y * ri.y;
        }
    }
    double sy = 0.0;
    block_sum2(sx, sy, part + blockIdx.x);
}

// r <- b - y - lambda x; part = |r|^2 (true residual)
__global__ void __launch_bounds__(CG_THREADS)
k_resid(const float2 *__restrict__ b, const float2 *__restrict__ y, const float2 *__restrict__ x,
        float lam, float2 *__restrict__ r, long long n, double2 *part)
{
    double sx = 0.0, sy = 0.0;
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS) {
        const float2 bi = b[i], yi = y[i], xi = x[i];
        float2 ri;
        ri.x = __builtin_fmaf(-lam, xi.x, bi.x - yi.x);
        ri.y = __builtin_fmaf(-lam, xi.y, bi.y - yi.y);
        r[i] = ri;
        sx += (double)ri.x * ri.x + (double)ri.y * ri.y;
    }
    block_sum2(sx, sy, part + blockIdx.x);
}

// p <- r + beta p
__global__ void __launch_bounds__(CG_THREADS)
k_pupd(const float2 *__restrict__ r, float2 *__restrict__ p, const double *scal, long long n)
{
    const float beta = (float)scal[S_BETA];
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS) {
        const float2 ri = r[i];
        float2 pi = p[i];
        pi.x = __builtin_fmaf(beta, pi.x, ri.x);
        pi.y = __builtin_fmaf(beta, pi.y, ri.y);
        p[i] = pi;
    }
}

__global__ void __launch_bounds__(CG_THREADS)
k_copy2(const float2 *__restrict__ src, float2 *__restrict__ d0, float2 *__restrict__ d1,
        long long n)
{
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS) {
        const float2 v = src[i];
        d0[i] = v;
        if (d1) d1[i] = v;
    }
}

__global__ void __launch_bounds__(CG_THREADS) k_zero(float2 *__restrict__ x, long long n)
{
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * CG_THREADS)
        x[i] = make_float2(0.f, 0.f);
}

// Finish a reduction (single block) and form the scalars of `mode`:
//   0: bb  = sum                           (|b|^2)
//   1: rr  = sum, hist[k] = rr              (initial residual)
//   2: pAp = Re sum, alpha = rr / pAp
//   3: beta = sum / rr, rr = sum, hist[k] = rr
//   4: beta = 0 (restart), rr = sum, hist[k] = rr
__global__ void __launch_bounds__(CG_THREADS)
k_finish(const double2 *__restrict__ part, int nparts, double *scal, double *hist, int k, int mode)
{
    double sx = 0.0, sy = 0.0;
    for (int i = threadIdx.x; i < nparts; i += CG_THREADS) {
        sx += part[i].x;
        sy += part[i].y;
    }
    __shared__ double2 tot;
    block_sum2(sx, sy, &tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double s = tot.x;
    if (mode == 0) {
        scal[S_BB] = s;
    } else if (mode == 1) {
        scal[S_RR] = s;
        hist[k] = s;
    } else if (mode == 2) {
        scal[S_PAP] = s;
        scal[S_ALPHA] = (s > 0.0) ? scal[S_RR] / s : 0.0;
    } else {
        const double rr = scal[S_RR];
        scal[S_BETA] = (mode == 3 && rr > 0.0) ? s / rr : 0.0;
        scal[S_RR] = s;
        hist[k] = s;
    }
}

struct cg_ws {
    float2 *r = nullptr, *p = nullptr, *Ap = nullptr;
    double2 *part = nullptr;
    double *scal = nullptr, *hist = nullptr;
    void free_all()
    {
        (void)hipFree(r);
        (void)hipFree(p);
        (void)hipFree(Ap);
        (void)hipFree(part);
        (void)hipFree(scal);
        (void)hipFree(hist);
    }
};

unsigned grid_for(long long n)
{
    long long g = (n + CG_THREADS - 1) / CG_THREADS;
    return (unsigned)(g < CG_BLOCKS ? (g < 1 ? 1 : g) : CG_BLOCKS);
}

// One iteration's work after the apply of p (Ap holds A^H A p):
// refresh = false: recursive residual; true: x update, then the true residual.
// A refresh through a different operator restarts the directions (beta = 0):
// each cycle then solves for the correction of the refreshed residual with
// `op`, i.e. iterative refinement of op's precision against opr's.
int cg_iter(mr_op op, mr_op opr, hipStream_t s, const cg_ws &w, const float2 *b, float2 *x,
            float lam, long long n, int k, bool refresh)
{
    const unsigned G = grid_for(n);
    MR_TRY(mr_op_execute(op, reinterpret_cast<const mr_complex *>(w.p),
                         reinterpret_cast<mr_complex *>(w.Ap)));
    hipLaunchKernelGGL(k_lam_dot, dim3(G), dim3(CG_THREADS), 0, s, w.p, w.Ap, lam, n, w.part);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(CG_THREADS), 0, s, w.part, (int)G, w.scal, w.hist,
                       k, 2);
    if (!refresh) {
        hipLaunchKernelGGL(k_step<true>, dim3(G), dim3(CG_THREADS), 0, s, x, w.r, w.p, w.Ap,
                           w.scal, n, w.part);
    } else {
        hipLaunchKernelGGL(k_step<false>, dim3(G), dim3(CG_THREADS), 0, s, x, w.r, w.p, w.Ap,
                           w.scal, n, w.part);
        MR_TRY(mr_op_execute(opr, reinterpret_cast<const mr_complex *>(x),
                             reinterpret_cast<mr_complex *>(w.Ap)));
        hipLaunchKernelGGL(k_resid, dim3(G), dim3(CG_THREADS), 0, s, b, w.Ap, x, lam, w.r, n,
                           w.part);
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(CG_THREADS), 0, s, w.part, (int)G, w.scal, w.hist,
                       k + 1, (refresh && opr != op) ? 4 : 3);
    hipLaunchKernelGGL(k_pupd, dim3(G), dim3(CG_THREADS), 0, s, w.r, w.p, w.scal, n);
    MR_LAUNCH_CHECK("cg iteration");
    return MR_SUCCESS;
}

int cg_capture(mr_op op, mr_op opr, hipStream_t s, const cg_ws &w, const float2 *b, float2 *x,
               float lam, long long n, int k, bool refresh, hipGraphExec_t *exec)
{
    MR_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = cg_iter(op, opr, s, w, b, x, lam, n, k, refresh);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != MR_SUCCESS) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) MR_FAIL(MR_ERR_CUDA, "CG graph capture failed: %s", hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess)
        MR_FAIL(MR_ERR_CUDA, "CG graph instantiation failed: %s", hipGetErrorString(e2));
    return MR_SUCCESS;
}

// Checks shared by both entry points; fills the stream, length and device.
int cg_args(const char *fn, mr_op op, mr_op opr, const mr_complex *d_b, const mr_complex *d_x,
            double lambda, int n_iter, int refresh, int flags, hipStream_t *s, long long *n,
            int *dev)
{
    MR_REQUIRE(op != nullptr, "%s: op is NULL", fn);
    MR_REQUIRE(d_b != nullptr && d_x != nullptr, "%s: d_b and d_x are required", fn);
    MR_REQUIRE((const void *)d_b != (const void *)d_x, "%s: d_x must not alias d_b", fn);
    MR_REQUIRE(n_iter >= 0, "%s: n_iter = %d must be >= 0", fn, n_iter);
    MR_REQUIRE(refresh >= 0, "%s: refresh = %d must be >= 0", fn, refresh);
    MR_REQUIRE(lambda >= 0.0, "%s: lambda = %g must be >= 0", fn, lambda);
    hipStream_t s2 = nullptr;
    long long n2 = 0;
    int dev2 = 0;
    MR_TRY(mr_op_stream_len(op, s, n, dev));
    MR_TRY(mr_op_stream_len(opr, &s2, &n2, &dev2));
    MR_REQUIRE(*n == n2 && *dev == dev2 && *s == s2,
               "%s: op_refresh must have the same shape, device and stream as op", fn);
    if (flags & MR_CG_GRAPH) {
        MR_REQUIRE(*s != nullptr, "%s: graph capture needs a non-default cfg.stream", fn);
        MR_REQUIRE(mr_env_ll("MR_AHA_HIP_PROFILE", 0) == 0,
                   "%s: graph capture is incompatible with MR_AHA_HIP_PROFILE", fn);
    }
    return MR_SUCCESS;
}

// The solve on work vectors w (allocated by the caller of this function).
// Runs on the operator's device, which must be current.
// sync_end: wait for the stream before returning (the work vectors are freed
// afterwards); otherwise a call without res_hist returns with the solve
// queued on the stream.
int cg_run(mr_op op, mr_op opr, hipStream_t s, const cg_ws &w, const mr_complex *d_b,
           mr_complex *d_x, double lambda, int n_iter, int refresh, int flags, long long n,
           double *res_hist, bool sync_end)
{
    const bool graph = (flags & MR_CG_GRAPH) != 0;
    const float2 *b = reinterpret_cast<const float2 *>(d_b);
    float2 *x = reinterpret_cast<float2 *>(d_x);
    const float lam = (float)lambda;
    const unsigned G = grid_for(n);
    hipGraphExec_t gx[2] = {nullptr, nullptr};
    int rc = MR_SUCCESS;
    auto fail = [&](int code) {
        rc = code;
        return code;
    };
    do {
        // |b|^2; r0 = b - (A^H A + lambda) x0, or b when x0 = 0
        hipLaunchKernelGGL(k_dot, dim3(G), dim3(CG_THREADS), 0, s, b, b, n, w.part);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(CG_THREADS), 0, s, w.part, (int)G, w.scal,
                           w.hist, 0, 0);
        if (flags & MR_CG_ZERO_INIT) {
            hipLaunchKernelGGL(k_zero, dim3(G), dim3(CG_THREADS), 0, s, x, n);
            hipLaunchKernelGGL(k_copy2, dim3(G), dim3(CG_THREADS), 0, s, b, w.r, w.p, n);
            hipLaunchKernelGGL(k_dot, dim3(G), dim3(CG_THREADS), 0, s, w.r, w.r, n, w.part);
        } else {
            if (mr_op_execute(opr, d_x, reinterpret_cast<mr_complex *>(w.Ap)) != MR_SUCCESS) {
                fail(MR_ERR_CUDA);
                break;
            }
            hipLaunchKernelGGL(k_resid, dim3(G), dim3(CG_THREADS), 0, s, b, w.Ap, x, lam, w.r, n,
                               w.part);
            hipLaunchKernelGGL(k_copy2, dim3(G), dim3(CG_THREADS), 0, s, w.r, w.p,
                               (float2 *)nullptr, n);
        }
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(CG_THREADS), 0, s, w.part, (int)G, w.scal,
                           w.hist, 0, 1);
        if (hipGetLastError() != hipSuccess) {
            mr_set_error("mr_op_cg: initial kernels failed to launch");
            fail(MR_ERR_CUDA);
            break;
        }
        for (int k = 0; k < n_iter && rc == MR_SUCCESS; ++k) {
            const bool ref = refresh > 0 && ((k + 1) % refresh) == 0;
            if (!graph) {
                if (cg_iter(op, opr, s, w, b, x, lam, n, k, ref) != MR_SUCCESS) fail(MR_ERR_CUDA);
                continue;
            }
            // the iteration index only selects the history slot: the two
            // kinds of iteration are captured once, writing |r|^2 to the
            // spare slot hist[n_iter + 1], which is copied to hist[k + 1]
            // after each replay
            hipGraphExec_t &e = gx[ref ? 1 : 0];
            if (!e && cg_capture(op, opr, s, w, b, x, lam, n, n_iter, ref, &e) != MR_SUCCESS) {
                fail(MR_ERR_CUDA);
                break;
            }
            if (hipGraphLaunch(e, s) != hipSuccess) {
                mr_set_error("mr_op_cg: graph launch failed");
                fail(MR_ERR_CUDA);
                break;
            }
            if (hipMemcpyAsync(w.hist + k + 1, w.hist + n_iter + 1, sizeof(double),
                               hipMemcpyDeviceToDevice, s) != hipSuccess) {
                fail(MR_ERR_CUDA);
                break;
            }
        }
        if (rc != MR_SUCCESS) break;
        if (res_hist) {
            // squared residual norms straight into the caller's array, then
            // scaled in place by |b|
            double bb = 0.0;
            if (hipMemcpyAsync(res_hist, w.hist, (size_t)(n_iter + 1) * sizeof(double),
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(&bb, w.scal + S_BB, sizeof(double), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                mr_set_error("mr_op_cg: reading the residual history failed");
                fail(MR_ERR_CUDA);
                break;
            }
            if (!(bb > 0.0)) bb = 1.0;
            for (int k = 0; k <= n_iter; ++k) res_hist[k] = std::sqrt(res_hist[k] / bb);
        }
    } while (false);
    if (gx[0]) (void)hipGraphExecDestroy(gx[0]);
    if (gx[1]) (void)hipGraphExecDestroy(gx[1]);
    if (rc == MR_SUCCESS && !res_hist && sync_end) {
        if (hipStreamSynchronize(s) != hipSuccess) {
            mr_set_error("mr_op_cg: stream synchronisation failed");
            rc = MR_ERR_CUDA;
        }
    } else if (rc != MR_SUCCESS) {
        (void)hipStreamSynchronize(s);
    }
    return rc;
}

// Work memory of mr_op_cg_noalloc: r, p, Ap (n complex values each), the
// reduction partials, the scalars and the residual history, carved from one
// static device array (each device the library runs on holds its own copy
// until the process ends). The host may not use the array's name as a
// device address: its address on each device comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t MR_CG_SCRATCH_BYTES = size_t(64) << 20;
constexpr int MR_CG_MAX_DEVICES = 64;
__device__ __align__(256) unsigned char scratch[MR_CG_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_ptr[MR_CG_MAX_DEVICES] = {};

int scratch_on_device(int dev, unsigned char **out)
{
    MR_REQUIRE(dev >= 0 && dev < MR_CG_MAX_DEVICES,
               "mr_op_cg_noalloc: device %d is outside 0..%d", dev, MR_CG_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            MR_FAIL(MR_ERR_CUDA, "mr_op_cg_noalloc: hipGetSymbolAddress of the scratch array "
                                 "failed on device %d: %s", dev, hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return MR_SUCCESS;
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" MR_EXPORT int mr_op_cg(mr_op op, mr_op op_refresh, const mr_complex *d_b,
                                  mr_complex *d_x, double lambda, int n_iter, int refresh,
                                  int flags, double *res_hist)
{
    mr_clear_error();
    mr_op opr = op_refresh ? op_refresh : op;
    hipStream_t s = nullptr;
    long long n = 0;
    int dev = 0;
    MR_TRY(cg_args("mr_op_cg", op, opr, d_b, d_x, lambda, n_iter, refresh, flags, &s, &n, &dev));
    int prev = 0;
    MR_HIP_CHECK(hipGetDevice(&prev));
    if (prev != dev) MR_HIP_CHECK(hipSetDevice(dev));
    cg_ws w;
    int rc = MR_SUCCESS;
    const size_t vb = (size_t)n * sizeof(float2);
    if (hipMalloc((void **)&w.r, vb) != hipSuccess || hipMalloc((void **)&w.p, vb) != hipSuccess ||
        hipMalloc((void **)&w.Ap, vb) != hipSuccess ||
        hipMalloc((void **)&w.part, CG_BLOCKS * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&w.scal, S_N * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&w.hist, (size_t)(n_iter + 2) * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        mr_set_error("mr_op_cg: failed to allocate %zu bytes of work vectors", 3 * vb);
        rc = MR_ERR_ALLOC;
    } else {
        rc = cg_run(op, opr, s, w, d_b, d_x, lambda, n_iter, refresh, flags, n, res_hist, true);
    }
    w.free_all();
    if (prev != dev) (void)hipSetDevice(prev);
    return rc;
}

extern "C" MR_EXPORT int mr_op_cg_noalloc(mr_op op, mr_op op_refresh, const mr_complex *d_b,
                                          mr_complex *d_x, double lambda, int n_iter,
                                          int refresh, int flags, double *res_hist)
{
    mr_clear_error();
    mr_op opr = op_refresh ? op_refresh : op;
    hipStream_t s = nullptr;
    long long n = 0;
    int dev = 0;
    MR_TRY(cg_args("mr_op_cg_noalloc", op, opr, d_b, d_x, lambda, n_iter, refresh, flags, &s, &n,
                   &dev));
    // graph instantiation allocates: this entry point runs its iterations eagerly
    MR_REQUIRE((flags & MR_CG_GRAPH) == 0,
               "mr_op_cg_noalloc: MR_CG_GRAPH is not available without allocation");
    const size_t vb = round256((size_t)n * sizeof(float2));
    const size_t need = 3 * vb + round256(CG_BLOCKS * sizeof(double2)) +
                        round256(S_N * sizeof(double)) +
                        round256((size_t)(n_iter + 2) * sizeof(double));
    MR_REQUIRE(need <= MR_CG_SCRATCH_BYTES,
               "mr_op_cg_noalloc: %lld unknowns and %d iterations need %zu bytes of work "
               "memory, above the %zu-byte scratch array (use mr_op_cg)",
               n, n_iter, need, MR_CG_SCRATCH_BYTES);
    int prev = 0;
    MR_HIP_CHECK(hipGetDevice(&prev));
    if (prev != dev) MR_HIP_CHECK(hipSetDevice(dev));
    unsigned char *base = nullptr;
    int rc = scratch_on_device(dev, &base);
    if (rc == MR_SUCCESS) {
        cg_ws w;
        w.r = reinterpret_cast<float2 *>(base);
        w.p = reinterpret_cast<float2 *>(base + vb);
        w.Ap = reinterpret_cast<float2 *>(base + 2 * vb);
        unsigned char *q = base + 3 * vb;
        w.part = reinterpret_cast<double2 *>(q);
        q += round256(CG_BLOCKS * sizeof(double2));
        w.scal = reinterpret_cast<double *>(q);
        q += round256(S_N * sizeof(double));
        w.hist = reinterpret_cast<double *>(q);
        rc = cg_run(op, opr, s, w, d_b, d_x, lambda, n_iter, refresh, flags, n, res_hist,
                    false);
    }
    if (prev != dev) (void)hipSetDevice(prev);
    return rc;
}
