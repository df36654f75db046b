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


// Block multiply of one location for two coils at once (inputs a0[l*ms] and
// a1[l*ms], in place), so that each packed kernel entry is loaded once for
// both: the vector-memory instructions of the kernel stream, which bound the
// fused pass, are halved per coil. Walk order and in-place semantics as
// mr_bd_walk_g. The entries stream through a ring of D registers (entry
// w + D is requested as entry w is consumed) so the two coils' outputs fit
// in registers beside them; (pi, pj) is the walk index and row of the next
// entry to request, advanced as the loop is unrolled, so every offset is a
// constant.
template <int L, bool UPPER, int D>
__device__ __forceinline__ void mr_bd_walk_pair(float2 *a0, float2 *a1, int ms, const float2 *ker,
                                                long long ks, unsigned boff)
{
    constexpr int T = L * (L + 1) / 2;
    float2 y0[L], y1[L], kk[D];
    int pi = 0, pj = 0;
    auto ld = [&]() {
        const int c = UPPER ? L - 1 - pi : pi;
        const int t0 = UPPER ? c * (c + 1) / 2 : c * L - c * (c - 1) / 2;
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(ker + ks * t0, 0xFFFFFFFFu);
        const float2 e = mr_bload(rs, boff, (unsigned)(ks * 8 * pj));
        if (++pj == L - pi) {
            ++pi;
            pj = 0;
        }
        return e;
    };
#pragma unroll
    for (int w = 0; w < D; ++w)
        if (w < T) kk[w] = ld();
#pragma unroll
    for (int l = 0; l < L; ++l) {
        y0[l] = make_float2(0.f, 0.f);
        y1[l] = make_float2(0.f, 0.f);
    }
    int w = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const int c = UPPER ? L - 1 - i : i;
        const float2 vc0 = a0[c * ms], vc1 = a1[c * ms];
#pragma unroll
        for (int j = 0; j < L - i; ++j, ++w) {
            const float2 e = kk[w % D];
            if (w + D < T) kk[w % D] = ld();
            const int r = (UPPER ? 0 : c) + j;
            if (r == c) {
                y0[c] = cmac(y0[c], e, vc0);
                y1[c] = cmac(y1[c], e, vc1);
            } else {
                y0[c] = cmac(y0[c], e, a0[r * ms]);
                y0[r] = cmacc(y0[r], e, vc0);
                y1[c] = cmac(y1[c], e, a1[r * ms]);
                y1[r] = cmacc(y1[r], e, vc1);
            }
            // entry w is consumed before the next request: at most D entries
            // are held in registers
            asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[r].x), "+v"(y0[r].y),
                         "+v"(y1[c].x), "+v"(y1[c].y), "+v"(y1[r].x), "+v"(y1[r].y)::"memory");
        }
        a0[c * ms] = y0[c];
        a1[c * ms] = y1[c];
    }
}

// mr_bd_walk_pair with the kernel entries fetched two per 16-byte load. The
// wavefront holds 32 pairs of adjacent locations (lane l < 32: the even one,
// lane l + 32: the odd one). For a column's rows j, j+1 the low half loads row
// j's entries of both locations of its pair and the high half row j+1's
// (boff_pair: the pair's even-location offset, plus one column for the high
// half); v_permlane32_swap then leaves every lane with both entries of its own
// location. An odd column length ends with one 8-byte load at boff_own. The
// loads, which bound the fused pass, drop from T to about T/2 per location.
// Needs the pair's entries adjacent and 16-byte aligned (an even kernel index
// for the even location).
//
// Load unit u covers rows j, j+1 (or the last row j) of walk index i; units
// are consumed in walk order through a ring of D, and the walk is unrolled
// by template recursion over u so that every index is a constant.
template <int L> struct mr_x2_units {
    static constexpr int count()
    {
        int n = 0;
        for (int i = 0; i < L; ++i) n += (L - i + 1) / 2;
        return n;
    }
    static constexpr int walk(int u)
    {
        int i = 0;
        while (u >= (L - i + 1) / 2) u -= (L - i + 1) / 2, ++i;
        return i;
    }
    static constexpr int row(int u)
    {
        int i = 0;
        while (u >= (L - i + 1) / 2) u -= (L - i + 1) / 2, ++i;
        return 2 * u;
    }
};

typedef unsigned int mr_u32x4 __attribute__((ext_vector_type(4)));

template <int L, bool UPPER, int U>
__device__ __forceinline__ mr_u32x4 mr_x2_load(const float2 *ker, long long ks, unsigned bp,
                                               unsigned bo)
{
    constexpr int i = mr_x2_units<L>::walk(U), j = mr_x2_units<L>::row(U);
    constexpr int c = UPPER ? L - 1 - i : i, len = L - i;
    constexpr int t0 = UPPER ? c * (c + 1) / 2 : c * L - c * (c - 1) / 2;
    const __amdgpu_buffer_rsrc_t rs = mr_rsrc(ker + ks * t0, 0xFFFFFFFFu);
    if constexpr (j + 1 < len) {
        return __builtin_amdgcn_raw_buffer_load_b128(rs, (int)bp, (int)(ks * 8 * j), 0);
    } else {
        const mr_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)bo, (int)(ks * 8 * j), 0);
        return mr_u32x4{v.x, v.y, 0u, 0u};
    }
}

template <int L, bool UPPER, int D, int U, int NC>
__device__ __forceinline__ void mr_x2_step(float2 (&y0)[L], float2 (&y1)[L], mr_u32x4 (&kq)[D],
                                           float2 *a0, float2 *a1, int ms, const float2 *ker,
                                           long long ks, unsigned bp, unsigned bo, float2 &vc0,
                                           float2 &vc1)
{
    constexpr int NU = mr_x2_units<L>::count();
    if constexpr (U < NU) {
        constexpr int i = mr_x2_units<L>::walk(U), j = mr_x2_units<L>::row(U);
        constexpr int c = UPPER ? L - 1 - i : i, len = L - i;
        if constexpr (j == 0) {
            vc0 = a0[c * ms];
            if constexpr (NC == 2) vc1 = a1[c * ms];
        }
        const mr_u32x4 q = kq[U % D];
        if constexpr (U + D < NU) kq[U % D] = mr_x2_load<L, UPPER, U + D>(ker, ks, bp, bo);
        float2 e[2];
        if constexpr (j + 1 < len) {
            // lanes 0-31 hold row j of the pair, 32-63 row j+1: after the
            // swap every lane has rows j and j+1 of its own location
            unsigned qx = q.x, qy = q.y, qz = q.z, qw = q.w;
            asm volatile("" : "+v"(qx), "+v"(qy), "+v"(qz), "+v"(qw));
            const auto sx = __builtin_amdgcn_permlane32_swap(qx, qz, false, false);
            const auto sy = __builtin_amdgcn_permlane32_swap(qy, qw, false, false);
            e[0] = make_float2(__uint_as_float(sx[0]), __uint_as_float(sy[0]));
            e[1] = make_float2(__uint_as_float(sx[1]), __uint_as_float(sy[1]));
        } else {
            e[0] = make_float2(__uint_as_float(q.x), __uint_as_float(q.y));
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (j + h < len) {
                const int r = (UPPER ? 0 : c) + j + h;
                if (r == c) {
                    y0[c] = cmac(y0[c], e[h], vc0);
                    if constexpr (NC == 2) y1[c] = cmac(y1[c], e[h], vc1);
                } else {
                    y0[c] = cmac(y0[c], e[h], a0[r * ms]);
                    y0[r] = cmacc(y0[r], e[h], vc0);
                    if constexpr (NC == 2) {
                        y1[c] = cmac(y1[c], e[h], a1[r * ms]);
                        y1[r] = cmacc(y1[r], e[h], vc1);
                    }
                }
                // consumed before the next unit is requested: at most D units
                // are held in registers
                if constexpr (NC == 2)
                    asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[r].x), "+v"(y0[r].y),
                                 "+v"(y1[c].x), "+v"(y1[c].y), "+v"(y1[r].x), "+v"(y1[r].y)::"memory");
                else
                    asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[r].x), "+v"(y0[r].y)::"memory");
            }
        }
        if constexpr (j + 2 >= len) {
            a0[c * ms] = y0[c];
            if constexpr (NC == 2) a1[c * ms] = y1[c];
        }
        mr_x2_step<L, UPPER, D, U + 1, NC>(y0, y1, kq, a0, a1, ms, ker, ks, bp, bo, vc0, vc1);
    }
}

template <int L, bool UPPER, int D, int U = 0>
__device__ __forceinline__ void mr_x2_prefetch(mr_u32x4 (&kq)[D], const float2 *ker, long long ks,
                                               unsigned bp, unsigned bo)
{
    if constexpr (U < D && U < mr_x2_units<L>::count()) {
        kq[U] = mr_x2_load<L, UPPER, U>(ker, ks, bp, bo);
        mr_x2_prefetch<L, UPPER, D, U + 1>(kq, ker, ks, bp, bo);
    }
}

// NC = 2: coils a0 and a1 share every load; NC = 1: a0 only (a1 unused).
template <int L, bool UPPER, int D, int NC = 2>
__device__ __forceinline__ void mr_bd_walk_pair_x2(float2 *a0, float2 *a1, int ms, const float2 *ker,
                                                   long long ks, unsigned boff_pair,
                                                   unsigned boff_own)
{
    float2 y0[L], y1[L], vc0, vc1 = make_float2(0.f, 0.f);
    mr_u32x4 kq[D];
    mr_x2_prefetch<L, UPPER, D>(kq, ker, ks, boff_pair, boff_own);
#pragma unroll
    for (int l = 0; l < L; ++l) {
        y0[l] = make_float2(0.f, 0.f);
        y1[l] = make_float2(0.f, 0.f);
    }
    mr_x2_step<L, UPPER, D, 0, NC>(y0, y1, kq, a0, a1, ms, ker, ks, boff_pair, boff_own, vc0, vc1);
}

// mr_bd_walk_pair_x2 for the lower triangle with L even, two coils, columns
// walked in pairs (c, c + 1): each input v_r below both is read from LDS once
// for the two columns instead of once per column (72 instead of 136 reads per
// coil at L = 16). The loads per column pair, in order: column c rows c, c+1
// (16-byte pair load), column c+1 row c+1 (8-byte own-location load), then
// for r = c+2, c+4, ..: column c rows r, r+1 and column c+1 rows r, r+1.
// Every output accumulates its terms in the order of the column walk, so the
// results equal mr_bd_walk_pair_x2's bit for bit.
template <int L> struct mr_x2c_units {
    static_assert(L % 2 == 0, "column pairs");
    // units of the column pair starting at column c
    static constexpr int per(int c) { return 2 + (L - c - 2); }
    static constexpr int count()
    {
        int n = 0;
        for (int c = 0; c < L; c += 2) n += per(c);
        return n;
    }
    // unit u: column (c or c+1), first row, and whether it is the single load
    struct unit { int col, row; bool single; };
    static constexpr unit at(int u)
    {
        int c = 0;
        while (u >= per(c)) u -= per(c), c += 2;
        if (u == 0) return unit{c, c, false};
        if (u == 1) return unit{c + 1, c + 1, true};
        const int k = (u - 2) / 2, b = (u - 2) % 2;
        return unit{c + b, c + 2 + 2 * k, false};
    }
};

template <int L, int U>
__device__ __forceinline__ mr_u32x4 mr_x2c_load(const float2 *ker, long long ks, unsigned bp,
                                                unsigned bo)
{
    constexpr auto un = mr_x2c_units<L>::at(U);
    constexpr int c = un.col, j = un.row - un.col;
    constexpr int t0 = c * L - c * (c - 1) / 2;
    const __amdgpu_buffer_rsrc_t rs = mr_rsrc(ker + ks * t0, 0xFFFFFFFFu);
    if constexpr (!un.single) {
        return __builtin_amdgcn_raw_buffer_load_b128(rs, (int)bp, (int)(ks * 8 * j), 0);
    } else {
        const mr_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)bo, (int)(ks * 8 * j), 0);
        return mr_u32x4{v.x, v.y, 0u, 0u};
    }
}

template <int L, int D, int U = 0>
__device__ __forceinline__ void mr_x2c_prefetch(mr_u32x4 (&kq)[D], const float2 *ker, long long ks,
                                               unsigned bp, unsigned bo)
{
    if constexpr (U < D && U < mr_x2c_units<L>::count()) {
        kq[U] = mr_x2c_load<L, U>(ker, ks, bp, bo);
        mr_x2c_prefetch<L, D, U + 1>(kq, ker, ks, bp, bo);
    }
}

// the entries of rows r, r+1 of this lane's location from a pair load
__device__ __forceinline__ void mr_x2_rows(const mr_u32x4 &q, float2 (&e)[2])
{
    unsigned qx = q.x, qy = q.y, qz = q.z, qw = q.w;
    asm volatile("" : "+v"(qx), "+v"(qy), "+v"(qz), "+v"(qw));
    const auto sx = __builtin_amdgcn_permlane32_swap(qx, qz, false, false);
    const auto sy = __builtin_amdgcn_permlane32_swap(qy, qw, false, false);
    e[0] = make_float2(__uint_as_float(sx[0]), __uint_as_float(sy[0]));
    e[1] = make_float2(__uint_as_float(sx[1]), __uint_as_float(sy[1]));
}

// Step over unit U (the first of a pair of units, or the leading pair of a
// column pair, which are consumed together). NC = 2: coils a0 and a1 share
// every load; NC = 1: a0 only.
template <int L, int D, int U, int NC>
__device__ __forceinline__ void mr_x2c_step(float2 (&y0)[L], float2 (&y1)[L], mr_u32x4 (&kq)[D],
                                           float2 *a0, float2 *a1, int ms, const float2 *ker,
                                           long long ks, unsigned bp, unsigned bo, float2 &vc0,
                                           float2 &vc1, float2 &vd0, float2 &vd1)
{
    constexpr int NU = mr_x2c_units<L>::count();
    if constexpr (U < NU) {
        constexpr auto ua = mr_x2c_units<L>::at(U);
        constexpr auto ub = mr_x2c_units<L>::at(U + 1);
        const mr_u32x4 qa = kq[U % D], qb = kq[(U + 1) % D];
        float2 ea[2], eb[2];
        mr_x2_rows(qa, ea);
        if constexpr (ub.single) eb[0] = make_float2(__uint_as_float(qb.x), __uint_as_float(qb.y));
        else mr_x2_rows(qb, eb);
        if constexpr (U + D < NU) kq[U % D] = mr_x2c_load<L, U + D>(ker, ks, bp, bo);
        if constexpr (U + 1 + D < NU) kq[(U + 1) % D] = mr_x2c_load<L, U + 1 + D>(ker, ks, bp, bo);
        if constexpr (ub.single) {
            // column pair (c, c+1) opens: rows c, c+1 of column c, row c+1 of
            // column c+1 (its diagonal)
            constexpr int c = ua.col;
            vc0 = a0[c * ms];
            vd0 = a0[(c + 1) * ms];
            if constexpr (NC == 2) {
                vc1 = a1[c * ms];
                vd1 = a1[(c + 1) * ms];
            }
            y0[c] = cmac(y0[c], ea[0], vc0);
            if constexpr (NC == 2) y1[c] = cmac(y1[c], ea[0], vc1);
            y0[c] = cmac(y0[c], ea[1], vd0);
            y0[c + 1] = cmacc(y0[c + 1], ea[1], vc0);
            if constexpr (NC == 2) {
                y1[c] = cmac(y1[c], ea[1], vd1);
                y1[c + 1] = cmacc(y1[c + 1], ea[1], vc1);
            }
            y0[c + 1] = cmac(y0[c + 1], eb[0], vd0);
            if constexpr (NC == 2) y1[c + 1] = cmac(y1[c + 1], eb[0], vd1);
            if constexpr (NC == 2)
                asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[c + 1].x), "+v"(y0[c + 1].y),
                             "+v"(y1[c].x), "+v"(y1[c].y), "+v"(y1[c + 1].x), "+v"(y1[c + 1].y)::"memory");
            else
                asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[c + 1].x),
                             "+v"(y0[c + 1].y)::"memory");
        } else {
            constexpr int c = ua.col, r = ua.row;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float2 w0 = a0[(r + h) * ms];
                // column c's terms, then column c+1's, as the column walk orders them
                y0[c] = cmac(y0[c], ea[h], w0);
                y0[r + h] = cmacc(y0[r + h], ea[h], vc0);
                if constexpr (NC == 2) {
                    const float2 w1 = a1[(r + h) * ms];
                    y1[c] = cmac(y1[c], ea[h], w1);
                    y1[r + h] = cmacc(y1[r + h], ea[h], vc1);
                    y0[c + 1] = cmac(y0[c + 1], eb[h], w0);
                    y0[r + h] = cmacc(y0[r + h], eb[h], vd0);
                    y1[c + 1] = cmac(y1[c + 1], eb[h], w1);
                    y1[r + h] = cmacc(y1[r + h], eb[h], vd1);
                    asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[r + h].x),
                                 "+v"(y0[r + h].y), "+v"(y1[c].x), "+v"(y1[c].y), "+v"(y1[r + h].x),
                                 "+v"(y1[r + h].y), "+v"(y0[c + 1].x), "+v"(y0[c + 1].y),
                                 "+v"(y1[c + 1].x), "+v"(y1[c + 1].y)::"memory");
                } else {
                    y0[c + 1] = cmac(y0[c + 1], eb[h], w0);
                    y0[r + h] = cmacc(y0[r + h], eb[h], vd0);
                    asm volatile("" : "+v"(y0[c].x), "+v"(y0[c].y), "+v"(y0[r + h].x),
                                 "+v"(y0[r + h].y), "+v"(y0[c + 1].x), "+v"(y0[c + 1].y)::"memory");
                }
            }
        }
        // the column pair is done: outputs c and c+1 are final
        constexpr auto un = mr_x2c_units<L>::at(U + 2 < NU ? U + 2 : 0);
        if constexpr (U + 2 >= NU || un.row == un.col) {
            constexpr int c = ua.col;
            a0[c * ms] = y0[c];
            a0[(c + 1) * ms] = y0[c + 1];
            if constexpr (NC == 2) {
                a1[c * ms] = y1[c];
                a1[(c + 1) * ms] = y1[c + 1];
            }
        }
        mr_x2c_step<L, D, U + 2, NC>(y0, y1, kq, a0, a1, ms, ker, ks, bp, bo, vc0, vc1, vd0, vd1);
    }
}

template <int L, int D, int NC = 2>
__device__ __forceinline__ void mr_bd_walk_pair_x2c(float2 *a0, float2 *a1, int ms,
                                                    const float2 *ker, long long ks,
                                                    unsigned boff_pair, unsigned boff_own)
{
    float2 y0[L], y1[L], vc0, vc1 = make_float2(0.f, 0.f), vd0, vd1 = make_float2(0.f, 0.f);
    mr_u32x4 kq[D];
    mr_x2c_prefetch<L, D>(kq, ker, ks, boff_pair, boff_own);
#pragma unroll
    for (int l = 0; l < L; ++l) {
        y0[l] = make_float2(0.f, 0.f);
        y1[l] = make_float2(0.f, 0.f);
    }
    mr_x2c_step<L, D, 0, NC>(y0, y1, kq, a0, a1, ms, ker, ks, boff_pair, boff_own, vc0, vc1, vd0,
                             vd1);
}

// ---------------------------------------------------------------------------
// Block multiply on the matrix cores (v_mfma_f32_4x4x1_16b_f32)
// ---------------------------------------------------------------------------
// One wavefront multiplies 32 adjacent locations e0 .. e0+31 for a pair of
// coils (or one coil). The instruction holds 16 independent 4x4 blocks; block
// b is one location, so each block carries its own k-point's matrix:
//   D_b[m][n] += A_b[m] * B_b[n]     (K = 1 per instruction, lane 4b+m / 4b+n)
// Rows m are (coil, part) = (c0 re, c0 im, c1 re, c1 im) of the inputs v_k;
// columns n are four outputs i = 4 ct + n. With A(k, i) = P + iQ the complex
// row-vector product y_i = sum_k v_k A(k, i) is, per coil c,
//   Re y_i = sum_k Re v_k P - Im v_k Q,   Im y_i = sum_k Re v_k Q + Im v_k P,
// so two accumulators per column tile (v P and v Q, over k = 0..L-1) give
// both coils' outputs in every lane's own registers: D1[2c] - D2[2c+1] and
// D2[2c] + D1[2c+1].
//
// Lane 4b + n loads A(k, 4 ct + n) of locations e0 + 2b and e0 + 2b + 1 with
// one 16-byte load (adjacent locations have adjacent kernel entries), so a
// wavefront runs two sets of blocks, the even and the odd locations. A(k, i)
// comes from the stored column i ("column form", k >= i lower / k <= i upper)
// or, conjugated, from the stored column k ("row form"). The column tiles run
// in two halves (fewer accumulators live); outputs of the first half are
// held until the second half has read every input, then all are stored in
// place.
typedef float mr_f4 __attribute__((ext_vector_type(4)));
#ifndef MR_MF_RING
#define MR_MF_RING 4
#endif

template <int L, bool UPPER> struct mr_mf_units {
    static constexpr int NCT = L / 4;                 // column tiles of 4 outputs
    static constexpr int H = NCT >= 2 ? 2 : 1;        // halves
    static constexpr int CPH = NCT / H;               // tiles per half
    static constexpr int NU = H * L * CPH;            // 16-byte loads per lane
    static constexpr int half(int u) { return u / (L * CPH); }
    static constexpr int k(int u) { return (u / CPH) % L; }
    static constexpr int cc(int u) { return u % CPH; }
    static constexpr int ct(int u) { return half(u) * CPH + cc(u); }
    // packed offset of stored column i's first row (column form: t = cs + k)
    static constexpr int cs(int i) { return UPPER ? i * (i + 1) / 2 : i * L - i * (i - 1) / 2 - i; }
    // row form: t = rs(k) + i
    static constexpr int rs(int kk) { return UPPER ? kk * (kk + 1) / 2 : kk * L - kk * (kk - 1) / 2 - kk; }
    // 0: every lane of unit u reads the column form, 1: every lane the row
    // form, 2: lane n chooses (the unit crosses the diagonal)
    static constexpr int form(int u)
    {
        const int kk = k(u), i0 = 4 * ct(u);
        if (!UPPER) return kk >= i0 + 3 ? 0 : (kk < i0 ? 1 : 2);
        return kk <= i0 ? 0 : (kk > i0 + 3 ? 1 : 2);
    }
};

template <int L, bool UPPER, int U>
__device__ __forceinline__ mr_f4 mr_mf_load(const float2 *const (&colp)[L / 4], const float2 *rowp,
                                            long long ks, int n)
{
    using M = mr_mf_units<L, UPPER>;
    constexpr int kk = M::k(U), ct = M::ct(U), f = M::form(U);
    const float2 *pc = colp[ct] + (long long)kk * ks;
    const float2 *pr = rowp + (long long)(M::rs(kk) + 4 * ct) * ks;
    const float2 *p;
    if constexpr (f == 0) {
        p = pc;
    } else if constexpr (f == 1) {
        p = pr;
    } else {
        const bool col = UPPER ? (n >= kk - 4 * ct) : (n <= kk - 4 * ct);
        p = col ? pc : pr;
    }
    return *reinterpret_cast<const mr_f4 *>(p);
}

template <int L, bool UPPER, int D, int U = 0>
__device__ __forceinline__ void mr_mf_prefetch(mr_f4 (&ring)[D], const float2 *const (&colp)[L / 4],
                                               const float2 *rowp, long long ks, int n)
{
    if constexpr (U < D && U < mr_mf_units<L, UPPER>::NU) {
        ring[U] = mr_mf_load<L, UPPER, U>(colp, rowp, ks, n);
        mr_mf_prefetch<L, UPPER, D, U + 1>(ring, colp, rowp, ks, n);
    }
}

template <int L, bool UPPER, int D, int NC, int U>
__device__ __forceinline__ void mr_mf_step(mr_f4 (&ring)[D], mr_f4 (&acc)[2][L / 4][2],
                                           float2 (&keep)[2][L / 4][2], const float2 *const (&colp)[L / 4],
                                           const float2 *rowp, long long ks, int n, const float *vb,
                                           const float *vb2, int ms, float (&av)[2])
{
    using M = mr_mf_units<L, UPPER>;
    if constexpr (U < M::NU) {
        constexpr int kk = M::k(U), cc = M::cc(U), ct = M::ct(U), f = M::form(U);
        if constexpr (cc == 0) {
            // inputs v_k of this lane's row: (coil n >> 1, part n & 1) at the
            // two locations of the lane's block; a second coil absent: zero.
            // Rows from L/2 on are addressed from a second base (vb2), which
            // keeps the LDS offsets immediate.
            constexpr int kh = kk >= L / 2 ? kk - L / 2 : kk;
            const float *v = kk >= L / 2 ? vb2 : vb;
            av[0] = v[2 * kh * ms];
            av[1] = v[2 * kh * ms + 2];
            if constexpr (NC == 1) {
                if (n >= 2) av[0] = av[1] = 0.f;
            }
        }
        const mr_f4 q = ring[U % D];
        if constexpr (U + D < M::NU) ring[U % D] = mr_mf_load<L, UPPER, U + D>(colp, rowp, ks, n);
        float q0 = q.y, q1 = q.w;
        if constexpr (f == 1) {
            q0 = -q0;
            q1 = -q1;
        } else if constexpr (f == 2) {
            const bool col = UPPER ? (n >= kk - 4 * ct) : (n <= kk - 4 * ct);
            if (!col) {
                q0 = -q0;
                q1 = -q1;
            }
        }
        acc[0][cc][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[0], q.x, acc[0][cc][0], 0, 0, 0);
        acc[0][cc][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[0], q0, acc[0][cc][1], 0, 0, 0);
        acc[1][cc][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[1], q.z, acc[1][cc][0], 0, 0, 0);
        acc[1][cc][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[1], q1, acc[1][cc][1], 0, 0, 0);
        if constexpr (kk == L - 1 && cc == M::CPH - 1) {
            // end of a half: outputs of its column tiles, y = D1 - D2 / D2 + D1
            constexpr int h = M::half(U);
#pragma unroll
            for (int c2 = 0; c2 < M::CPH; ++c2) {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const mr_f4 d1 = acc[s][c2][0], d2 = acc[s][c2][1];
                        keep[s][h * M::CPH + c2][c] =
                            make_float2(d1[2 * c] - d2[2 * c + 1], d2[2 * c] + d1[2 * c + 1]);
                    }
                    acc[s][c2][0] = mr_f4{0.f, 0.f, 0.f, 0.f};
                    acc[s][c2][1] = mr_f4{0.f, 0.f, 0.f, 0.f};
                }
            }
        }
        mr_mf_step<L, UPPER, D, NC, U + 1>(ring, acc, keep, colp, rowp, ks, n, vb, vb2, ms, av);
    }
}

// ap: the pair's first coil (pencil l at ap + l * ms, location e at + e); the
// second coil at ap + L * ms. kb: kernel index of location e0 + 2b (even).
template <int L, bool UPPER, int NC>
__device__ __forceinline__ void mr_bd_mfma32(float2 *ap, int ms, const float2 *ker, long long ks,
                                             long long kb, int e0, int lane)
{
    static_assert(L % 4 == 0, "MFMA block multiply: L a multiple of 4");
    using M = mr_mf_units<L, UPPER>;
    constexpr int D = M::NU < MR_MF_RING ? M::NU : MR_MF_RING;
    // opaque per call: the per-lane kernel bases below are recomputed for each
    // 32-location task instead of being hoisted out of the caller's loop for
    // every instantiation at once (which spilled)
    int n = lane & 3, b = lane >> 2;
    asm volatile("" : "+v"(n), "+v"(b));
    const int e = e0 + 2 * b;
    const float2 *kbase = ker + kb;
    const float2 *colp[M::NCT];
#pragma unroll
    for (int ct = 0; ct < M::NCT; ++ct) colp[ct] = kbase + (long long)M::cs(4 * ct + n) * ks;
    const float2 *rowp = kbase + (long long)n * ks;
    mr_f4 ring[D];
    mr_mf_prefetch<L, UPPER, D>(ring, colp, rowp, ks, n);
    const float *vb =
        reinterpret_cast<const float *>(ap + (long long)((n >> 1) & (NC - 1)) * L * ms + e) + (n & 1);
    const float *vb2 = vb + 2 * (L / 2) * ms;
    asm volatile("" : "+v"(vb2));
    mr_f4 acc[2][M::NCT][2];
    float2 keep[2][M::NCT][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int c = 0; c < M::NCT; ++c) acc[s][c][0] = acc[s][c][1] = mr_f4{0.f, 0.f, 0.f, 0.f};
    float av[2];
    mr_mf_step<L, UPPER, D, NC, 0>(ring, acc, keep, colp, rowp, ks, n, vb, vb2, ms, av);
    // every input has been read: store both locations of each output at once
#pragma unroll
    for (int ct = 0; ct < M::NCT; ++ct)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const float2 y0 = keep[0][ct][c], y1 = keep[1][ct][c];
            *reinterpret_cast<mr_f4 *>(ap + (long long)(c * L + 4 * ct + n) * ms + e) =
                mr_f4{y0.x, y0.y, y1.x, y1.y};
        }
}

// ---------------------------------------------------------------------------
// 16-bit kernel format (BF16 or FP16 entries with FP32 scales per 4x4 block)
// ---------------------------------------------------------------------------
// Entry t of location k is stored as one 32-bit word q[t * kcol + k] (real
// part in the low half, imaginary part in the high half), both divided by the
// scale of its 4x4 block of the L x L matrix: sc[bi * kcol + k] = the largest
// |real| or |imaginary| part of the block's stored entries (1 for an all-zero
// block), bi = h (h + 1) / 2 + l for block row/column indices h >= l.
#define MR_KF_FP32 0
#define MR_KF_BF16 1
#define MR_KF_FP16 2

__host__ __device__ constexpr int mr_kf_nblk(int L)
{
    return ((L + 3) / 4) * ((L + 3) / 4 + 1) / 2;
}

__device__ __forceinline__ int mr_kf_blk(int r, int c)
{
    const int a = r >> 2, b = c >> 2;
    const int h = a > b ? a : b, l = a > b ? b : a;
    return h * (h + 1) / 2 + l;
}

template <int KF> __device__ __forceinline__ float2 mr_kf_decode(unsigned u)
{
    if constexpr (KF == MR_KF_BF16) {
        return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xFFFF0000u));
    } else {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 h = __builtin_bit_cast(h2, u);
        return make_float2((float)h.x, (float)h.y);
    }
}

__device__ __forceinline__ unsigned mr_bf16_rne(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

// Block multiply of one location from the 16-bit kernel: as mr_bd_walk_g,
// with the entries of each 4-row run of a column (one block) accumulated
// unscaled and the block's scale applied once: y_c += s * sum_r q_rc v_r,
// y_r += conj(q_rc) (s v_c).
template <int L, bool UPPER, int KF>
__device__ __forceinline__ void mr_bd_walk_q(float2 *a, int ms, const unsigned *q, const float *sc,
                                             long long ks, unsigned bq, unsigned bs)
{
    float2 y[L];
    unsigned kk[2][L];
#pragma unroll
    for (int l = 0; l < L; ++l) y[l] = make_float2(0.f, 0.f);
    auto col = [](int i, int &c, int &t0, int &len) {
        c = UPPER ? L - 1 - i : i;
        t0 = UPPER ? c * (c + 1) / 2 : c * L - c * (c - 1) / 2;
        len = UPPER ? c + 1 : L - c;
    };
    auto ldw = [&](const unsigned *base, unsigned off, unsigned soff) {
        const __amdgpu_buffer_rsrc_t r = mr_rsrc(base, 0xFFFFFFFFu);
        return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0);
    };
    {
        int c, t0, len;
        col(0, c, t0, len);
#pragma unroll
        for (int j = 0; j < L; ++j)
            if (j < len) kk[0][j] = ldw(q + ks * t0, bq, (unsigned)(ks * 4 * j));
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
        int c, t0, len;
        col(i, c, t0, len);
        if (i + 1 < L) {
            int c1, t1, len1;
            col(i + 1, c1, t1, len1);
#pragma unroll
            for (int j = 0; j < L; ++j)
                if (j < len1) kk[(i + 1) & 1][j] = ldw(q + ks * t1, bq, (unsigned)(ks * 4 * j));
        }
        // scales of the blocks this column crosses (rows r0 .. r0 + len - 1)
        const int r0 = UPPER ? 0 : c;
        float s[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int rb = (r0 >> 2) + b;
            if (rb * 4 < r0 + len) {
                const int bi = (rb > (c >> 2)) ? rb * (rb + 1) / 2 + (c >> 2)
                                               : (c >> 2) * ((c >> 2) + 1) / 2 + rb;
                s[b] = __uint_as_float(ldw((const unsigned *)sc, bs, (unsigned)(ks * 4 * bi)));
            }
        }
        asm volatile("" ::: "memory");
        const float2 vc = a[c * ms];
        float2 part = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < L; ++j) {
            if (j < len) {
                const int r = r0 + j;
                const int b = (r >> 2) - (r0 >> 2);
                const float2 e = mr_kf_decode<KF>(kk[i & 1][j]);
                const float2 vr = (r == c) ? vc : a[r * ms];
                part = cmac(part, e, vr);
                if (r != c) {
                    const float2 vcs = make_float2(s[b] * vc.x, s[b] * vc.y);
                    y[r] = cmacc(y[r], e, vcs);
                }
                // end of this block's run of rows: apply its scale to y_c
                if (j == len - 1 || ((r + 1) & 3) == 0) {
                    y[c].x = __builtin_fmaf(s[b], part.x, y[c].x);
                    y[c].y = __builtin_fmaf(s[b], part.y, y[c].y);
                    part = make_float2(0.f, 0.f);
                }
            }
        }
        a[c * ms] = y[c];
    }
}

// Host side of the specialised kernels: the kernel for x length n, rank L,
// mask mode and kernel format kf (MR_KF_*), or nullptr when there is none.
// *gmax receives the most coils one workgroup may take.
typedef void (*mr_xbd_ct_fn)(float2 *, const float2 *, const int *, long long, mr_dims,
                             const float2 *, long long, int, int, int, const unsigned *,
                             const float *);
// mf: the variant whose block step runs on the matrix cores (FP32 kernel,
// unmasked; the caller checks the kernel's 16-byte alignment per launch).
mr_xbd_ct_fn mr_xbd_ct_pick(int n, int L, bool masked, int kf, int *gmax, bool mf = false);
// The workgroup of 1024 threads and two coils (160 KB of LDS, one workgroup
// per CU) for x length n, rank L and kernel format kf, or nullptr.
mr_xbd_ct_fn mr_xbd_ct_pick_wide(int n, int L, bool masked, int kf, bool mf = false);
// Persistent two-coil 1024-thread form (x length 512 or 640, L = 16, FP32
// kernel), for an image filling the centred half of the x grid with 16-byte
// aligned rows; launched with one workgroup per CU. Null for other shapes.
mr_xbd_ct_fn mr_xbd_ct_pick_persistent(int n, int L, bool masked, int kf);

// Fill the 16-bit format q [T, kcol], sc [NB, kcol] from an FP32 kernel.
int mr_kf_quantize(int kf, const float2 *ker, long long kcol, int L, int upper, unsigned *q,
                   float *sc, hipStream_t stream);
