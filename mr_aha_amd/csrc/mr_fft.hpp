// Project FFT engine for gfx950: batched 1-D complex transforms along one
// axis of a strided array, one workgroup per tile of pencils staged in LDS.
//
// Algorithm: self-sorting (Stockham) mixed-radix DIT. Stage s with radix R
// and Ns = product of earlier radices maps, for every butterfly j < n/R,
//   v_r = A[j + r n/R] * w^(k r n/(Ns R)),  k = j mod Ns,  w = e^{-+2 pi i/n}
//   V   = DFT_R(v)
//   B[(j div Ns) Ns R + k + q Ns] = V_q
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC 2008). Radices 2, 3, 4, 5, 7, 8 have unrolled
// butterflies, and 6 and 10 prime-factor ones (no internal twiddles, so
// e.g. 640 = 8 * 8 * 10 takes three stages); any other prime factor uses a
// per-output dot product over
// the n-th-root table, so every length works (prime grids of the test
// fixtures) while smooth lengths take the fast butterflies.
//
// The transform is unnormalised; FWD uses e^{-2 pi i jk/n}, INV e^{+...}.
// Data never leaves the workgroup between stages: a pencil is loaded once
// through a Load functor and written once through a Store functor, which
// is where padding, sensitivity products, crops and coil sums are fused.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#define MR_FFT_MAXST 24

template <typename R> struct cvec;
template <> struct cvec<float>  { using T = float2; };
template <> struct cvec<double> { using T = double2; };

template <typename V> __device__ __forceinline__ V cmake(decltype(V::x) a, decltype(V::x) b)
{
    V r; r.x = a; r.y = b; return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmake<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmake<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b)
{
    return cmake<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
template <typename V> __device__ __forceinline__ V cmulc(V a, V b)
{
    return cmake<V>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return cmake<V>(a.x, -a.y); }
// y + a * b and y + conj(a) * b as four fused multiply-adds (a complex
// product followed by an add would take six instructions: the compiler may
// not reassociate the add into the product)
__device__ __forceinline__ float2 cmac(float2 y, float2 a, float2 b)
{
    y.x = __builtin_fmaf(a.x, b.x, y.x);
    y.x = __builtin_fmaf(-a.y, b.y, y.x);
    y.y = __builtin_fmaf(a.x, b.y, y.y);
    y.y = __builtin_fmaf(a.y, b.x, y.y);
    return y;
}
__device__ __forceinline__ float2 cmacc(float2 y, float2 a, float2 b)
{
    y.x = __builtin_fmaf(a.x, b.x, y.x);
    y.x = __builtin_fmaf(a.y, b.y, y.x);
    y.y = __builtin_fmaf(a.x, b.y, y.y);
    y.y = __builtin_fmaf(-a.y, b.x, y.y);
    return y;
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename V> __device__ __forceinline__ V cmul_mi(V a)
{
    return INV ? cmake<V>(-a.y, a.x) : cmake<V>(a.y, -a.x);
}

// Division by a run-time constant d as a multiply-high and a shift
// (round-up multiplier m = ceil(2^(31+l)/d), l = ceil(log2 d)); exact for
// dividends below 2^31. Built on the host, used in index maps on the device.
struct mr_fdiv {
    unsigned d, mul, shr;
};
inline mr_fdiv mr_fdiv_make(unsigned d)
{
    mr_fdiv f;
    f.d = d;
    if (d <= 1) {
        f.mul = 0;
        f.shr = 0;
        return f;
    }
    unsigned l = 0;
    while ((1u << l) < d) ++l;
    const unsigned p = 31 + l;
    f.mul = (unsigned)(((1ull << p) + d - 1) / d);
    f.shr = p - 32;
    return f;
}
__host__ __device__ __forceinline__ unsigned mr_fdiv_q(const mr_fdiv &f, unsigned n)
{
#ifdef __HIP_DEVICE_COMPILE__
    return (f.d <= 1) ? n : (__umulhi(n, f.mul) >> f.shr);
#else
    return (f.d <= 1) ? n : (unsigned)(((unsigned long long)n * f.mul) >> 32) >> f.shr;
#endif
}

// Per-length plan passed by value to kernels.
template <typename R> struct mr_fft_len {
    int n;
    int nst;
    int radix[MR_FFT_MAXST];
    mr_fdiv mdiv[MR_FFT_MAXST];    // n / radix[s]
    mr_fdiv nsdiv[MR_FFT_MAXST];   // product of the radices before stage s
    mr_fdiv ndiv;                  // n
    const typename cvec<R>::T *tw;  // tw[q] = exp(-2 pi i q / n), q < n (device)
    // Stage-major table (device, n entries): the twiddle of butterfly input r
    // at offset k < Ns of stage (R, Ns) is tws[r * Ns + k - 1] =
    // exp(-2 pi i k r / (Ns R)), 1 <= r < R. Consecutive k are consecutive
    // entries, so a wavefront's reads are coalesced (global) or free of bank
    // conflicts (LDS); the stages' ranges [Ns - 1, Ns R - 1) tile [0, n - 1).
    const typename cvec<R>::T *tws;
};

template <typename R> inline void mr_fft_fill_divs(mr_fft_len<R> *f)
{
    int Ns = 1;
    for (int s = 0; s < f->nst; ++s) {
        f->mdiv[s] = mr_fdiv_make((unsigned)(f->n / f->radix[s]));
        f->nsdiv[s] = mr_fdiv_make((unsigned)Ns);
        Ns *= f->radix[s];
    }
    f->ndiv = mr_fdiv_make((unsigned)f->n);
}

// Factor n into the stage radices: 8s first, then 10 and 6 (prime-factor
// butterflies), 4, 2, 3, 5, 7, then the remaining primes in ascending order.
inline int mr_fft_factor(int n, int *radix)
{
    int ns = 0;
    int m = n;
    const int small[] = {8, 10, 6, 4, 2, 3, 5, 7};
    for (int r : small)
        while (m % r == 0 && ns < MR_FFT_MAXST) { radix[ns++] = r; m /= r; }
    for (int p = 11; m > 1 && ns < MR_FFT_MAXST; p += 2)
        while (m % p == 0 && ns < MR_FFT_MAXST) { radix[ns++] = p; m /= p; }
    return (m == 1) ? ns : -1;
}

// Radices with an unrolled butterfly.
inline bool mr_fft_unrolled(int r) { return (r >= 2 && r <= 8) || r == 10; }

// n is "smooth" when every radix is one of the unrolled ones.
inline bool mr_fft_smooth(int n)
{
    int r[MR_FFT_MAXST];
    int ns = mr_fft_factor(n, r);
    if (ns < 0) return false;
    for (int i = 0; i < ns; ++i)
        if (!mr_fft_unrolled(r[i])) return false;
    return true;
}

// Smallest 2^a 3^b 5^c 7^d >= n.
inline int mr_fft_next_smooth(int n)
{
    for (int m = (n < 1 ? 1 : n);; ++m) {
        int k = m;
        for (int p : {2, 3, 5, 7}) while (k % p == 0) k /= p;
        if (k == 1) return m;
    }
}

// ---------------------------------------------------------------------------
// Butterflies
// ---------------------------------------------------------------------------

template <bool INV, typename V> __device__ __forceinline__ void bfly2(V *v)
{
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV, typename V> __device__ __forceinline__ void bfly4(V *v)
{
    V a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    V c = cadd(v[1], v[3]), d = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
}

// bfly8 from the radix-4 transforms e of the even and o of the odd inputs.
template <bool INV, typename V> __device__ __forceinline__ void bfly8_join(V *v, const V *e, const V *o)
{
    using R = decltype(V::x);
    const R h = (R)0.70710678118654752440;
    // w8^q for q = 0..3: 1, (1 -+ i)/sqrt2, -+i, (-1 -+ i)/sqrt2
    V t1 = INV ? cmake<V>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
               : cmake<V>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    V t2 = cmul_mi<INV>(o[2]);
    V t3 = INV ? cmake<V>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
               : cmake<V>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);  v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], t1);    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);    v[7] = csub(e[3], t3);
}

template <bool INV, typename V> __device__ __forceinline__ void bfly8(V *v)
{
    V e[4] = {v[0], v[2], v[4], v[6]};
    V o[4] = {v[1], v[3], v[5], v[7]};
    bfly4<INV>(e);
    bfly4<INV>(o);
    bfly8_join<INV>(v, e, o);
}

// bfly4 of (a, 0, 0, b): a +- b and a +- (-+i)(-b).
template <bool INV, typename V> __device__ __forceinline__ void bfly4_ends(V *r, V a, V b)
{
    const V d = cmul_mi<INV>(cmake<V>(-b.x, -b.y));
    r[0] = cadd(a, b);
    r[2] = csub(a, b);
    r[1] = cadd(a, d);
    r[3] = csub(a, d);
}

// bfly8 with v[2] .. v[5] zero (the first stage of a zero-padded centred
// pencil): the same values as bfly8 up to the sign of an exact zero, with
// half the additions of its radix-4 halves.
template <bool INV, typename V> __device__ __forceinline__ void bfly8_ends(V *v)
{
    V e[4], o[4];
    bfly4_ends<INV>(e, v[0], v[6]);
    bfly4_ends<INV>(o, v[1], v[7]);
    bfly8_join<INV>(v, e, o);
}

// Odd prime radix with compile-time constants: V_q = sum_r v_r e^{-+2 pi i r q/R},
// written with the symmetric pairs (r, R-r) to halve the multiplies.
template <int RR> struct mr_trig;
template <> struct mr_trig<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <> struct mr_trig<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                    -0.80901699437494742410, 0.30901699437494742410};
    static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                    -0.58778525229247312917, -0.95105651629515357212};
};
template <> struct mr_trig<7> {
    static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                    -0.90096886790241912624, -0.90096886790241912624,
                                    -0.22252093395631440429, 0.62348980185873353053};
    static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                    0.43388373911755812048, -0.43388373911755812048,
                                    -0.97492791218182360702, -0.78183148246802980871};
};

template <int RR, bool INV, typename V> __device__ __forceinline__ void bfly_odd(V *v)
{
    using R = decltype(V::x);
    constexpr int H = (RR - 1) / 2;
    V sp[H], dm[H];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
        sp[r - 1] = cadd(v[r], v[RR - r]);
        dm[r - 1] = csub(v[r], v[RR - r]);
    }
    V out[RR];
    V s0 = v[0];
#pragma unroll
    for (int r = 0; r < H; ++r) s0 = cadd(s0, sp[r]);
    out[0] = s0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        // re part from symmetric sums with cos, im part from differences with sin
        R ar = v[0].x, ai = v[0].y, br = 0, bi = 0;
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            const int k = (r * q) % RR;
            const R c = (R)mr_trig<RR>::c[k];
            const R s = (R)mr_trig<RR>::s[k];
            ar += c * sp[r - 1].x;
            ai += c * sp[r - 1].y;
            br += s * dm[r - 1].x;
            bi += s * dm[r - 1].y;
        }
        // forward: e^{-i theta}: sum = A - i B  with B = sum s*dm
        if (!INV) {
            out[q] = cmake<V>(ar + bi, ai - br);
            out[RR - q] = cmake<V>(ar - bi, ai + br);
        } else {
            out[q] = cmake<V>(ar - bi, ai + br);
            out[RR - q] = cmake<V>(ar + bi, ai - br);
        }
    }
#pragma unroll
    for (int q = 0; q < RR; ++q) v[q] = out[q];
}

template <int RR, bool INV, typename V> __device__ __forceinline__ void bfly(V *v);

constexpr int mr_modinv(int a, int m)
{
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 1;
}

// Composite radix R = R1 R2 with coprime factors by the prime-factor
// (Good-Thomas) algorithm: input n = (R2 n1 + R1 n2) mod R, output
// k = (R2 <R2^-1>_R1 k1 + R1 <R1^-1>_R2 k2) mod R, so that
// W_R^{nk} = W_R1^{n1 k1} W_R2^{n2 k2}: R2 DFTs of length R1, then R1 DFTs of
// length R2, with no twiddles between them.
template <int R1, int R2, bool INV, typename V> __device__ __forceinline__ void bfly_pfa(V *v)
{
    constexpr int R = R1 * R2;
    constexpr int e1 = R2 * mr_modinv(R2 % R1, R1);
    constexpr int e2 = R1 * mr_modinv(R1 % R2, R2);
    V y[R2][R1];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) y[n2][n1] = v[(R2 * n1 + R1 * n2) % R];
        bfly<R1, INV>(y[n2]);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        V t[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) t[n2] = y[n2][k1];
        bfly<R2, INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[(e1 * k1 + e2 * k2) % R] = t[k2];
    }
}

template <int RR, bool INV, typename V> __device__ __forceinline__ void bfly(V *v)
{
    if constexpr (RR == 2) bfly2<INV>(v);
    else if constexpr (RR == 4) bfly4<INV>(v);
    else if constexpr (RR == 8) bfly8<INV>(v);
    else if constexpr (RR == 6) bfly_pfa<2, 3, INV>(v);
    else if constexpr (RR == 10) bfly_pfa<2, 5, INV>(v);
    else bfly_odd<RR, INV>(v);
}

// One Stockham stage with compile-time radix, LDS A -> LDS B, over tp pencils
// (pencil stride ns) of length n, Ns = product of earlier radices.
template <int RR, bool INV, typename V>
__device__ __forceinline__ void mr_stage(const V *__restrict__ A, V *__restrict__ B, const V *tw,
                                         int n, int ns, int tp, int Ns, const mr_fdiv &mdv,
                                         const mr_fdiv &nsdv)
{
    const int m = n / RR;
    for (int f = threadIdx.x; f < tp * m; f += blockDim.x) {
        const int p = (int)mr_fdiv_q(mdv, (unsigned)f);
        const int j = f - p * m;
        const int jn = (int)mr_fdiv_q(nsdv, (unsigned)j);
        const int k = j - jn * Ns;
        const V *a = A + p * ns + j;
        V v[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r) v[r] = a[r * m];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < RR; ++r) {
                V w = tw[r * Ns + k - 1];
                v[r] = INV ? cmulc(w, v[r]) : cmul(w, v[r]);
            }
        }
        bfly<RR, INV>(v);
        V *b = B + p * ns + jn * Ns * RR + k;
#pragma unroll
        for (int q = 0; q < RR; ++q) b[q * Ns] = v[q];
    }
}

// Any radix: each thread evaluates one output of one butterfly as a length-R
// dot product; the combined stage and DFT exponent is r (k + Ns q) stride,
// looked up in the natural-order table tw (global memory).
template <bool INV, typename V>
__device__ __forceinline__ void mr_stage_any(const V *__restrict__ A, V *__restrict__ B,
                                             const V *tw, int n, int ns, int tp, int Ns, int RR)
{
    const int m = n / RR;
    const int stride = n / (Ns * RR);
    const int per = m * RR;  // outputs per pencil
    for (int f = threadIdx.x; f < tp * per; f += blockDim.x) {
        const int p = f / per;
        const int rem = f - p * per;
        const int j = rem / RR;
        const int q = rem - j * RR;
        const int k = j % Ns;
        const V *a = A + p * ns + j;
        const long long base = (long long)(k + Ns * q) * stride;
        V acc = cmake<V>(0, 0);
        for (int r = 0; r < RR; ++r) {
            V w = tw[(int)((base * r) % n)];
            acc = cadd(acc, INV ? cmulc(w, a[r * m]) : cmul(w, a[r * m]));
        }
        B[p * ns + (j / Ns) * Ns * RR + k + q * Ns] = acc;
    }
}

// Full transform of tp pencils held in LDS buffer A (B is scratch of the
// same size). Returns the buffer that holds the result.
template <bool INV, typename R>
__device__ typename cvec<R>::T *mr_fft_lds(const mr_fft_len<R> &f, typename cvec<R>::T *A,
                                           typename cvec<R>::T *B, const typename cvec<R>::T *tw,
                                           int ns, int tp)
{
    using V = typename cvec<R>::T;
    int Ns = 1;
    for (int s = 0; s < f.nst; ++s) {
        const int RR = f.radix[s];
        const mr_fdiv &md = f.mdiv[s], &nd = f.nsdiv[s];
        switch (RR) {
        case 2: mr_stage<2, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 3: mr_stage<3, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 4: mr_stage<4, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 5: mr_stage<5, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 6: mr_stage<6, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 7: mr_stage<7, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 8: mr_stage<8, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        case 10: mr_stage<10, INV, V>(A, B, tw, f.n, ns, tp, Ns, md, nd); break;
        default: mr_stage_any<INV, V>(A, B, f.tw, f.n, ns, tp, Ns, RR); break;
        }
        __syncthreads();
        V *t = A; A = B; B = t;
        Ns *= RR;
    }
    return A;
}

// Hardware dispatches workgroup b of a 1-D grid to XCD b mod 8. Renumber so
// that every XCD receives a contiguous run of logical tiles: neighbouring
// tiles (which share cache lines and kernel rows) then meet in one L2.
__device__ __forceinline__ long long mr_xcd_block(long long bid, long long nb)
{
    const long long q = nb / 8, r = nb % 8;
    const long long x = bid % 8, i = bid / 8;
    return x * q + (x < r ? x : r) + i;
}

// LDS pencil stride: odd for even n so that strided LDS accesses of
// neighbouring pencils fall in different banks.
__host__ __device__ inline int mr_fft_ns(int n) { return (n % 2 == 0) ? n + 1 : n; }

// Functor protocol of the axis kernels.
//
// Load:  static constexpr bool kElemFast;  (true: a pencil's elements are
//        contiguous in memory, so consecutive threads walk one pencil;
//        false: consecutive threads walk consecutive pencils)
//        struct PB;                         per-pencil base, computed once
//        int nsub() const;                  load rounds (e.g. one per coil)
//        PB pencil(long long p, int sub) const;
//        V get(const PB &, int e) const;    natural element e of the pencil
// Store: static constexpr bool kElemFast; struct PB;
//        PB pencil(long long p, int sub) const;
//        void put(const PB &, int pl, int e, int sub, V v, V *acc) const;
//        void flush(long long p0, int tp, V *acc) const;   after the last round
// Pencil bases live in LDS, so the per-element address is one multiply-add.
#define MR_FFT_MAXTX 64

template <class Load, class Store, typename V>
__device__ __forceinline__ void mr_fft_load_tile(const Load &ld, typename Load::PB *lpb, V *A,
                                                 const mr_fdiv &ndiv, int n, int ns, int tp,
                                                 int lgTX)
{
    if (Load::kElemFast) {
        for (int q = threadIdx.x; q < tp * n; q += blockDim.x) {
            const int p = (int)mr_fdiv_q(ndiv, (unsigned)q), e = q - p * n;
            A[p * ns + e] = ld.get(lpb[p], e);
        }
    } else {
        const int TX = 1 << lgTX;
        for (int q = threadIdx.x; q < (n << lgTX); q += blockDim.x) {
            const int e = q >> lgTX, p = q & (TX - 1);
            if (p < tp) A[p * ns + e] = ld.get(lpb[p], e);
        }
    }
}

template <class Store, typename V>
__device__ __forceinline__ void mr_fft_store_tile(const Store &st, typename Store::PB *spb,
                                                  const V *A, V *acc, const mr_fdiv &ndiv, int n,
                                                  int ns, int tp, int lgTX, int sub)
{
    if (Store::kElemFast) {
        for (int q = threadIdx.x; q < tp * n; q += blockDim.x) {
            const int p = (int)mr_fdiv_q(ndiv, (unsigned)q), e = q - p * n;
            st.put(spb[p], p, e, sub, A[p * ns + e], acc);
        }
    } else {
        const int TX = 1 << lgTX;
        for (int q = threadIdx.x; q < (n << lgTX); q += blockDim.x) {
            const int e = q >> lgTX, p = q & (TX - 1);
            if (p < tp) st.put(spb[p], p, e, sub, A[p * ns + e], acc);
        }
    }
}

// Ping-pong kernel: any length and precision. TX = 2^lgTX pencils per
// workgroup (the last one may hold fewer).
template <typename R, bool INV, class Load, class Store>
__global__ void __launch_bounds__(256)
mr_fft_axis_kernel(mr_fft_len<R> f, Load ld, Store st, long long npencil, int lgTX,
                   long long pbase)
{
    using V = typename cvec<R>::T;
    extern __shared__ __align__(16) unsigned char mr_smem[];
    __shared__ typename Load::PB lpb[MR_FFT_MAXTX];
    __shared__ typename Store::PB spb[MR_FFT_MAXTX];
    const int TX = 1 << lgTX;
    const int n = f.n;
    const int ns = mr_fft_ns(n);
    V *tw = reinterpret_cast<V *>(mr_smem);
    V *A0 = tw + n;
    V *B0 = A0 + (size_t)TX * ns;
    V *acc = B0 + (size_t)TX * ns;
    const long long p0 = pbase + mr_xcd_block(blockIdx.x, gridDim.x) * TX;
    const int tp = (int)((npencil - p0 < TX) ? (npencil - p0) : TX);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = f.tws[i];
    const int nsub = ld.nsub();
    for (int sub = 0; sub < nsub; ++sub) {
        if (threadIdx.x < tp) {
            lpb[threadIdx.x] = ld.pencil(p0 + threadIdx.x, sub);
            spb[threadIdx.x] = st.pencil(p0 + threadIdx.x, sub);
        }
        __syncthreads();
        mr_fft_load_tile<Load, Store, V>(ld, lpb, A0, f.ndiv, n, ns, tp, lgTX);
        __syncthreads();
        V *res = mr_fft_lds<INV, R>(f, A0, B0, tw, ns, tp);
        mr_fft_store_tile<Store, V>(st, spb, res, acc, f.ndiv, n, ns, tp, lgTX, sub);
        __syncthreads();
    }
    st.flush(p0, tp, acc);
}

// ---------------------------------------------------------------------------
// Register-staged variant for smooth lengths: one LDS buffer instead of two.
// Every stage first reads all of a thread's butterfly inputs into registers,
// then (after a barrier) writes the outputs back in place, so a workgroup
// needs TX*n + n LDS values and can hold twice the pencils of the
// ping-pong kernel. Requires TX*n <= MR_FFT_EMAX and unrolled radices
// (2..8, 10).
// ---------------------------------------------------------------------------

#define MR_FFT_EMAX 6144
#define MR_FFT_FAST_THREADS 512

// NT = workgroup size, EMAX = bound on tp * n (sizes the register stage).
template <int RR, bool INV, typename V, int NT = MR_FFT_FAST_THREADS, int EMAX = MR_FFT_EMAX>
__device__ __forceinline__ void mr_stage_reg(V *A, const V *tw, int n, int ns, int tp, int Ns,
                                             const mr_fdiv &mdv, const mr_fdiv &nsdv)
{
    constexpr int KMAX = (EMAX / RR + NT - 1) / NT;
    const int m = n / RR;
    const int nb = tp * m;
    V v[KMAX][RR];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int f = threadIdx.x + k * NT;
        if (f < nb) {
            const int p = (int)mr_fdiv_q(mdv, (unsigned)f), j = f - p * m;
            const V *a = A + p * ns + j;
#pragma unroll
            for (int r = 0; r < RR; ++r) v[k][r] = a[r * m];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int f = threadIdx.x + k * NT;
        if (f < nb) {
            const int p = (int)mr_fdiv_q(mdv, (unsigned)f), j = f - p * m;
            const int jn = (int)mr_fdiv_q(nsdv, (unsigned)j);
            const int kk = j - jn * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < RR; ++r) {
                    const V w = tw[r * Ns + kk - 1];
                    v[k][r] = INV ? cmulc(w, v[k][r]) : cmul(w, v[k][r]);
                }
            }
            bfly<RR, INV>(v[k]);
            V *b = A + p * ns + jn * Ns * RR + kk;
#pragma unroll
            for (int q = 0; q < RR; ++q) b[q * Ns] = v[k][q];
        }
    }
    __syncthreads();
}

template <typename R, bool INV, class Load, class Store>
__global__ void __launch_bounds__(MR_FFT_FAST_THREADS)
mr_fft_axis_fast(mr_fft_len<R> f, Load ld, Store st, long long npencil, int lgTX,
                 long long pbase)
{
    using V = typename cvec<R>::T;
    extern __shared__ __align__(16) unsigned char mr_smem[];
    __shared__ typename Load::PB lpb[MR_FFT_MAXTX];
    __shared__ typename Store::PB spb[MR_FFT_MAXTX];
    const int TX = 1 << lgTX;
    const int n = f.n;
    const int ns = mr_fft_ns(n);
    V *tw = reinterpret_cast<V *>(mr_smem);
    V *A = tw + n;
    V *acc = A + (size_t)TX * ns;
    const long long p0 = pbase + mr_xcd_block(blockIdx.x, gridDim.x) * TX;
    const int tp = (int)((npencil - p0 < TX) ? (npencil - p0) : TX);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = f.tws[i];
    const int nsub = ld.nsub();
    for (int sub = 0; sub < nsub; ++sub) {
        if (threadIdx.x < tp) {
            lpb[threadIdx.x] = ld.pencil(p0 + threadIdx.x, sub);
            spb[threadIdx.x] = st.pencil(p0 + threadIdx.x, sub);
        }
        __syncthreads();
        mr_fft_load_tile<Load, Store, V>(ld, lpb, A, f.ndiv, n, ns, tp, lgTX);
        __syncthreads();
        int Ns = 1;
        for (int s = 0; s < f.nst; ++s) {
            const int RR = f.radix[s];
            const mr_fdiv &md = f.mdiv[s], &nd = f.nsdiv[s];
            switch (RR) {
            case 2: mr_stage_reg<2, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            case 3: mr_stage_reg<3, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            case 4: mr_stage_reg<4, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            case 5: mr_stage_reg<5, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            case 6: mr_stage_reg<6, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            case 7: mr_stage_reg<7, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            case 10: mr_stage_reg<10, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            default: mr_stage_reg<8, INV, V>(A, tw, n, ns, tp, Ns, md, nd); break;
            }
            Ns *= RR;
        }
        mr_fft_store_tile<Store, V>(st, spb, A, acc, f.ndiv, n, ns, tp, lgTX, sub);
        __syncthreads();
    }
    st.flush(p0, tp, acc);
}

// ---------------------------------------------------------------------------
// Host side: twiddle tables cached per (length, precision).
// ---------------------------------------------------------------------------

// Returns the device plan of length n in precision R (tables are created on
// first use on the current device and kept until mr_fft_release_all()).
template <typename R> int mr_fft_get(int n, mr_fft_len<R> *out);
void mr_fft_release_all();

// Workgroup tile and dynamic LDS of an axis pass.
struct mr_fft_launch {
    int TX;
    size_t lds;
    long long blocks;
};
template <typename R>
inline mr_fft_launch mr_fft_pick(int n, long long npencil, size_t acc_elems_per_pencil,
                                 int want_tx)
{
    using V = typename cvec<R>::T;
    const size_t budget = 64 * 1024;
    const int ns = mr_fft_ns(n);
    int TX = want_tx;
    auto need = [&](int tx) {
        return ((size_t)n + 2 * (size_t)tx * ns + (size_t)tx * acc_elems_per_pencil) * sizeof(V);
    };
    while (TX > 1 && need(TX) > budget) TX /= 2;
    while (TX > 1 && TX / 2 >= npencil) TX /= 2;
    mr_fft_launch l;
    l.TX = TX;
    l.lds = need(TX);
    l.blocks = (npencil + TX - 1) / TX;
    return l;
}
