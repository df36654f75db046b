// HIP-side helpers shared by the device translation units.
#pragma once

#include <cstdlib>

#include "mr_internal.hpp"
#include "mr_fft.hpp"

#include <hip/hip_runtime.h>

#define MR_HIP_CHECK(call)                                                     \
    do {                                                                       \
        hipError_t mr_e_ = (call);                                             \
        if (mr_e_ != hipSuccess)                                               \
            MR_FAIL(MR_ERR_CUDA, "%s failed at %s:%d: %s", #call, __FILE__,    \
                    __LINE__, hipGetErrorString(mr_e_));                       \
    } while (0)

#define MR_LAUNCH_CHECK(name)                                                  \
    do {                                                                       \
        hipError_t mr_e_ = hipGetLastError();                                  \
        if (mr_e_ != hipSuccess)                                               \
            MR_FAIL(MR_ERR_CUDA, "kernel %s failed to launch at %s:%d: %s",    \
                    name, __FILE__, __LINE__, hipGetErrorString(mr_e_));       \
    } while (0)

// Most workgroups of `threads` work-items one dispatch may hold: the dispatch
// packet counts work-items in 32 bits, and a larger grid would be silently
// truncated. Kept a multiple of 1024 workgroups. MR_AHA_HIP_MAX_BLOCKS lowers
// it (tests run the chunked launches at small sizes with it).
inline long long mr_hw_max_blocks(int threads)
{
    return (0xFFFFFFFFLL / threads) & ~1023LL;
}
inline long long mr_max_blocks(int threads)
{
    const long long hw = mr_hw_max_blocks(threads);
    const char *e = getenv("MR_AHA_HIP_MAX_BLOCKS");
    const long long v = e ? atoll(e) : 0;
    return (v > 0 && v < hw) ? v : hw;
}

// Fail (instead of launching a truncated grid) when blocks x threads
// work-items do not fit one dispatch; for launches that are not chunked.
#define MR_GRID_CHECK(blocks, threads, name)                                   \
    do {                                                                       \
        if ((long long)(blocks) > mr_hw_max_blocks(threads))                   \
            MR_FAIL(MR_ERR_UNSUPPORTED,                                        \
                    "%s: %lld workgroups of %d exceed one dispatch (2^32 "     \
                    "work-items)", name, (long long)(blocks), (int)(threads)); \
    } while (0)

// Largest power of two <= v (1 for v < 1).
inline int mr_pow2_floor(int v)
{
    int p = 1;
    while (p <= v / 2) p *= 2;
    return p;
}

// Make cfg's device current (gpu_device_id < 0 keeps the caller's); *prev
// receives the device to restore or -1.
int mr_push_device(const mr_config *cfg, int *prev);
void mr_pop_device(int prev);

// Budget = clamp(mem_fraction) * free bytes of the current device.
int mr_mem_budget(const mr_config *cfg, size_t *budget);

// Launch one axis pass of the FFT engine on `stream`: the register-staged
// single-buffer kernel for smooth FP32 lengths (MR_AHA_HIP_FFT_FAST=0 disables it),
// the ping-pong kernel otherwise (any length, FP64). TX is a power of two.
template <typename R, bool INV, class Load, class Store>
int mr_fft_pass(const mr_fft_len<R> &f, const Load &ld, const Store &st, long long npencil,
                size_t acc_per_pencil, int want_tx, hipStream_t stream, const char *name)
{
    using V = typename cvec<R>::T;
    if (npencil <= 0) return MR_SUCCESS;
    if (want_tx > MR_FFT_MAXTX) want_tx = MR_FFT_MAXTX;
    // the kernels take 2^lgTX pencils per workgroup: a width that is not a
    // power of two is rounded down here, before LDS and the grid are sized
    want_tx = mr_pow2_floor(want_tx);
    static const bool fast_on = mr_env_ll("MR_AHA_HIP_FFT_FAST", 1) != 0;
    bool smooth = sizeof(R) == 4 && fast_on;
    for (int i = 0; i < f.nst && smooth; ++i) smooth = mr_fft_unrolled(f.radix[i]);
    int TX;
    size_t lds;
    if (smooth && f.n <= MR_FFT_EMAX) {
        const int ns = mr_fft_ns(f.n);
        auto need = [&](int tx) {
            return ((size_t)f.n + (size_t)tx * ns + (size_t)tx * acc_per_pencil) * sizeof(V);
        };
        TX = want_tx;
        while (TX > 1 && ((long long)TX * f.n > MR_FFT_EMAX || need(TX) > 64 * 1024)) TX /= 2;
        while (TX > 1 && TX / 2 >= npencil) TX /= 2;
        lds = need(TX);
    } else {
        smooth = false;
        mr_fft_launch l = mr_fft_pick<R>(f.n, npencil, acc_per_pencil, want_tx);
        TX = l.TX;
        lds = l.lds;
    }
    int lg = 0;
    while ((1 << lg) < TX) ++lg;
    const long long blocks = (npencil + TX - 1) / TX;
    // a dispatch counts its work-items in 32 bits: larger grids are launched
    // in chunks of whole tiles (pbase: first pencil of the chunk)
    const int threads = smooth ? MR_FFT_FAST_THREADS : 256;
    const long long maxb = mr_max_blocks(threads);
    for (long long b0 = 0; b0 < blocks; b0 += maxb) {
        const long long nb = blocks - b0 < maxb ? blocks - b0 : maxb;
        if (smooth)
            hipLaunchKernelGGL((mr_fft_axis_fast<R, INV, Load, Store>), dim3((unsigned)nb),
                               dim3(threads), lds, stream, f, ld, st, npencil, lg, b0 * TX);
        else
            hipLaunchKernelGGL((mr_fft_axis_kernel<R, INV, Load, Store>), dim3((unsigned)nb),
                               dim3(threads), lds, stream, f, ld, st, npencil, lg, b0 * TX);
        MR_LAUNCH_CHECK(name);
    }
    return MR_SUCCESS;
}

// Generic axis functors over an array viewed as [I, len, O] (inner stride 1):
// pencil p = i + I * o, element e along the axis at offset i + I*(e + len*o).

// Full-length load.
template <typename V, bool FAST> struct mr_ld_axis {
    static constexpr bool kElemFast = FAST;
    using PB = long long;
    const V *in;
    long long I;
    int n;
    __device__ int nsub() const { return 1; }
    __device__ PB pencil(long long p, int) const
    {
        const long long o = p / I, i = p - o * I;
        return i + I * (long long)n * o;
    }
    __device__ V get(const PB &b, int e) const { return in[b + I * e]; }
};

// Zero-padded load: natural element e of a padded axis of length M reads
// source index s = (e + c) mod M of a compact axis of length Ns when s < Ns
// (the centering scatter of src/cuda/mr_kernels.cu:111-128), else 0.
template <typename V, bool FAST> struct mr_ld_pad {
    static constexpr bool kElemFast = FAST;
    using PB = long long;
    const V *in;
    long long I;
    int M, Ns, c;
    __device__ int nsub() const { return 1; }
    __device__ PB pencil(long long p, int) const
    {
        const long long o = p / I, i = p - o * I;
        return i + I * (long long)Ns * o;
    }
    __device__ V get(const PB &b, int e) const
    {
        int s = e + c;
        if (s >= M) s -= M;
        return (s < Ns) ? in[b + I * s] : cmake<V>(0, 0);
    }
};

// Full-length store to [I, n, O].
template <typename V, bool FAST> struct mr_st_axis {
    static constexpr bool kElemFast = FAST;
    using PB = long long;
    V *out;
    long long I;
    int n;
    __device__ PB pencil(long long p, int) const
    {
        const long long o = p / I, i = p - o * I;
        return i + I * (long long)n * o;
    }
    __device__ void put(const PB &b, int, int e, int, V v, V *) const { out[b + I * e] = v; }
    __device__ void flush(long long, int, V *) const {}
};

// Cropping store: natural element e lands at compact index s = (e + c) mod M
// when s < Ns; other elements are discarded (crop of mr_kernels.cu:158-165).
template <typename V, bool FAST> struct mr_st_crop {
    static constexpr bool kElemFast = FAST;
    using PB = long long;
    V *out;
    long long I;
    int M, Ns, c;
    __device__ PB pencil(long long p, int) const
    {
        const long long o = p / I, i = p - o * I;
        return i + I * (long long)Ns * o;
    }
    __device__ void put(const PB &b, int, int e, int, V v, V *) const
    {
        int s = e + c;
        if (s >= M) s -= M;
        if (s < Ns) out[b + I * s] = v;
    }
    __device__ void flush(long long, int, V *) const {}
};

// Kernel (centered) linear index of natural-order grid location loc.
__device__ __forceinline__ long long mr_nat_to_cent_lin(long long loc, const mr_dims &d)
{
    const int j0 = (int)(loc % d.m[0]);
    const long long q = loc / d.m[0];
    const int j1 = (int)(q % d.m[1]);
    const int j2 = (int)(q / d.m[1]);
    int k0 = j0 + d.c_m[0]; if (k0 >= d.m[0]) k0 -= d.m[0];
    int k1 = j1 + d.c_m[1]; if (k1 >= d.m[1]) k1 -= d.m[1];
    int k2 = j2 + d.c_m[2]; if (k2 >= d.m[2]) k2 -= d.m[2];
    return k0 + (long long)d.m[0] * (k1 + (long long)d.m[1] * k2);
}
