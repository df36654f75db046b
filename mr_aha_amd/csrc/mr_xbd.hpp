// Device helpers of the fused x pass (forward x-FFT, block multiply and
// inverse x-FFT of one k-space row in one workgroup), shared by the kernel
// for any smooth x length and the kernels specialised at compile time for
// the common lengths.
#pragma once

#include "mr_hip.hpp"

#define MR_XBD_THREADS 512
#define MR_XBD_EMAX 16384

// Raw buffer access (gfx9 resource word 3 = 0x00020000): the base address and
// the byte range sit in scalar registers, a lane supplies a 32-bit byte
// offset, and reads past the range return zero. A wavefront-uniform base
// (a kernel column, an image pencil) then costs no vector instructions for
// addressing, where a 64-bit pointer per lane costs two or three per access.
typedef unsigned int mr_u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t mr_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}

__device__ __forceinline__ float2 mr_bload(__amdgpu_buffer_rsrc_t r, unsigned off,
                                          unsigned soff = 0)
{
    const mr_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0);
    return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}

__device__ __forceinline__ void mr_bstore(__amdgpu_buffer_rsrc_t r, unsigned off, float2 v)
{
    mr_u32x2 u;
    u.x = __float_as_uint(v.x);
    u.y = __float_as_uint(v.y);
    __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)off, 0, 0);
}

// Block multiply of one location, in place on the L values a[0], a[ms], ...
// (LDS). Packed entry t of the location is read at byte offset boff of
// kernel column t (ker + t * ks). One resource serves a walked column: its
// entries differ by the scalar offset j*ks*8 (the setup guarantees every
// offset stays below 2^32, and the range is left open because the range
// check includes the scalar offset). Columns are
// walked so that output c is final once column c is done (lower: ascending
// c, since column c holds rows r >= c and the mirrors of earlier columns have
// already reached y_c; upper: descending c, by symmetry), and v_c is not
// read by any later column: y_c is stored over v_c at once. The inputs v_r
// are read from LDS where they are used (their slots are only overwritten
// once dead), so the registers hold the L partial outputs and two columns of
// kernel entries: the next column's loads are issued before the current one
// is multiplied.
template <int L, bool UPPER>
__device__ __forceinline__ void mr_bd_walk_g(float2 *a, int ms, const float2 *ker, long long ks,
                                             unsigned boff)
{
    float2 y[L], kk[2][L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = make_float2(0.f, 0.f);
    // packed offset and length of the i-th column walked
    auto col = [](int i, int &c, int &t0, int &len) {
        c = UPPER ? L - 1 - i : i;
        t0 = UPPER ? c * (c + 1) / 2 : c * L - c * (c - 1) / 2;
        len = UPPER ? c + 1 : L - c;
    };
    {
        int c, t0, len;
        col(0, c, t0, len);
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(ker + ks * t0, 0xFFFFFFFFu);
#pragma unroll
        for (int j = 0; j < L; ++j)
            if (j < len) kk[0][j] = mr_bload(rs, boff, (unsigned)(ks * 8 * j));
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
        int c, t0, len;
        col(i, c, t0, len);
        if (i + 1 < L) {
            int c1, t1, len1;
            col(i + 1, c1, t1, len1);
            const __amdgpu_buffer_rsrc_t rs = mr_rsrc(ker + ks * t1, 0xFFFFFFFFu);
#pragma unroll
            for (int j = 0; j < L; ++j)
                if (j < len1) kk[(i + 1) & 1][j] = mr_bload(rs, boff, (unsigned)(ks * 8 * j));
        }
        asm volatile("" ::: "memory");
        const int r0 = UPPER ? 0 : c;
        const float2 vc = a[c * ms];
#pragma unroll
        for (int j = 0; j < L; ++j) {
            if (j < len) {
                const int r = r0 + j;
                const float2 e = kk[i & 1][j];
                const float2 vr = (r == c) ? vc : a[r * ms];
                y[c] = cmac(y[c], e, vr);
                if (r != c) y[r] = cmacc(y[r], e, vc);
            }
        }
        a[c * ms] = y[c];
    }
}


// Host side of the specialised kernels: the kernel for x length n, rank L
// and mask mode, or nullptr when n is not one of the specialised lengths.
// *gmax receives the most coils one workgroup may take.
typedef void (*mr_xbd_ct_fn)(float2 *, const float2 *, const int *, long long, mr_dims,
                             const float2 *, long long, int, int, int);
mr_xbd_ct_fn mr_xbd_ct_pick(int n, int L, bool masked, int *gmax);
