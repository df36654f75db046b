// Compile-time FFT stages: the length N, its radices and every stride and
// divisor are constants, so divisions become one full-rate 24-bit multiply
// and a shift, LDS offsets become instruction immediates, and no run-time
// plan occupies scalar registers. Same self-sorting Stockham algorithm and
// radix rule as the run-time engine (mr_fft.hpp), and the same stage-major
// twiddle table (entry r*Ns + k - 1 of stage (R, Ns)).
#pragma once

#include "mr_hip.hpp"

#include <type_traits>

typedef unsigned int mr_u32x4_ct __attribute__((ext_vector_type(4)));

// Stage radices of n by the rule of mr_fft_factor (8s, then 10, 6, 4, 2, 3, 5, 7)
// and the product Ns of the radices before each stage.
struct mr_ct_fac {
    int nst;
    int r[16];
    int ns[16];
    int rest;  // 1 when n is smooth
};

constexpr mr_ct_fac mr_ct_factor(int n)
{
    mr_ct_fac f{};
    const int small[8] = {8, 10, 6, 4, 2, 3, 5, 7};
    int m = n, Ns = 1;
    for (int i = 0; i < 8; ++i)
        while (m % small[i] == 0 && f.nst < 16) {
            f.r[f.nst] = small[i];
            f.ns[f.nst] = Ns;
            Ns *= small[i];
            ++f.nst;
            m /= small[i];
        }
    f.rest = m;
    return f;
}

constexpr int mr_ct_clog2(int d)
{
    int k = 0;
    while ((1 << k) < d) ++k;
    return k;
}

// x / D for 0 <= x < 2^14 as (x * M) >> K with M = ceil(2^K / D),
// K = 14 + ceil(log2 D): the rounding error x (M D - 2^K) / (D 2^K) stays
// below 2^-ceil(log2 D) <= 1/D, so the floor is exact; x * M < 2^30 fits
// one 24-bit multiply. Powers of two shift.
template <int D> struct mr_cdiv {
    static_assert(D >= 1 && D <= (1 << 14), "divisor range");
    static constexpr bool POW2 = (D & (D - 1)) == 0;
    static constexpr int K = 14 + mr_ct_clog2(D);
    static constexpr unsigned M = POW2 ? 0u : (unsigned)(((1ull << K) + D - 1) / D);
    __device__ static __forceinline__ unsigned q(unsigned x)
    {
        if constexpr (POW2) return x >> mr_ct_clog2(D);
        else return __umul24(x, M) >> K;
    }
};

// Exhaustive compile-time check of mr_cdiv over its whole domain.
template <int D> constexpr bool mr_cdiv_exact()
{
    if (mr_cdiv<D>::POW2) return true;
    for (unsigned x = 0; x < (1u << 14); ++x)
        if ((unsigned)(((unsigned long long)x * mr_cdiv<D>::M) >> mr_cdiv<D>::K) != x / D)
            return false;
    return true;
}

// Twiddle sources: a buffer resource over the global table, or the table
// copied into LDS.
// Twiddles from global memory (buffer loads): a butterfly loads only
// w = W^k and forms W^(r k) = w^r by complex products (kPowers), since the
// loads, not the arithmetic, are the scarce resource of the fused x pass.
// Each power is a product of two lower ones, at most four roundings deep.
struct mr_tw_buf {
    static constexpr bool kPowers = true;
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ float2 get(unsigned idx, unsigned off) const
    {
        const unsigned int __attribute__((ext_vector_type(2))) v =
            __builtin_amdgcn_raw_buffer_load_b64(r, (int)(idx * 8u), (int)(off * 8u), 0);
        return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
    }
};
struct mr_tw_lds {
    static constexpr bool kPowers = false;
    const float2 *t;
    __device__ __forceinline__ float2 get(unsigned idx, unsigned off) const { return t[idx + off]; }
};

// Workgroup barrier over LDS traffic only: this wavefront's LDS operations
// complete, then s_barrier. __syncthreads() would also drain the row loads
// still streaming into the other buffer (its release fence covers every LDS
// write, and a global-to-LDS load is one), so the overlap would be lost. No
// memory access moves across it (memory clobber); the row loads are awaited
// explicitly where their buffer is read.
__device__ __forceinline__ void mr_lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// No hook between a stage's reads and its writes (mr_ct_stage).
struct mr_ct_nomid {
    static constexpr bool active = false;
    __device__ void operator()() const {}
};
// A hook run once a stage's inputs and twiddles are in registers, before its
// writes: work that may reuse the LDS slots the stage does not write.
template <class F> struct mr_ct_mid {
    static constexpr bool active = true;
    F f;
    __device__ void operator()() const { f(); }
};

// One Stockham stage (radix R, Ns = NS) in place over tp pencils of length N
// at pencil stride NSTR in LDS, NT threads. Phase 1 reads every butterfly
// input and twiddle of the thread into registers, phase 2 (after the
// barrier) writes the outputs. Rounds are skipped per wavefront (a scalar
// test on the wavefront's first butterfly); inside a round, lanes past the
// last butterfly repeat it, reading the same inputs and writing the same
// values to the same places, so no per-lane mask lives across the barrier.
// SWI / SWO: the stage reads / writes element e of pencil p at e ^ mr_ct_swz(p)
// (the wide-tile layout, whose unpadded pencils would otherwise put the 16
// pencils of one row on the same LDS banks); inner stages use plain order.
__device__ __forceinline__ unsigned mr_ct_swz(unsigned p) { return (p & 7u) << 2; }

// PR = 1: the first forward stage of a centred zero-padded pencil (image
// points at e < N/4 and e >= 3N/4, zeros between; radix 8): inputs r = 2..5
// are zero, so they are not read, and the butterfly skips them. PR = 2: the
// last inverse stage when only outputs e < N/4 and e >= 3N/4 are kept (the
// crop of the same pencil): outputs landing wholly between are not stored,
// and their arithmetic is dropped with them.
// SZ bit 1: a first stage (NS = 1, radix 8) stores output P of a pencil at
// P ^ 2((P >> 4) & 3); bit 2: the second stage reads its inputs there.
// A first-stage lane writes its 8 outputs as 64 contiguous bytes, so in plain
// order the 16-byte writes of 8 neighbouring lanes hit two of the eight
// 16-byte bank groups (4-way conflicts); the swap of 16-byte units within
// each 64-byte block by bits 4-5 of the position spreads them over all eight.
// The second stage's reads stay permutations within aligned 16-element runs.
// SZ bits 64 / 128: the same pair of stages where the second stage's N / R is
// a multiple of 8 but not of 16 (x length 320): P ^ 2((P >> 3) & 3), which
// takes the first stage's writes from 4-way to 2-way conflicts.
// SZ bit 4: a second stage with NS = 8 and radix 8 stores output P at
// P ^ 8((P >> 6) & 1); bit 8: the third stage (N / R = 64) reads there. The
// second stage's writes from lanes 8 apart (positions 64 apart) otherwise
// share banks; the flip of bit 3 by bit 6 separates them.
template <int N, int NSTR, int R, int NS, bool INV, int NT, int TPMAX, class TW, bool SWI = false,
          bool SWO = false, int PR = 0, bool LB = false, class MID = mr_ct_nomid, int SZ = 0>
__device__ __forceinline__ void mr_ct_stage(float2 *A, const TW &tw, int tp, int wave,
                                            const MID &mid = MID{})
{
    static_assert(!(SZ & 1) || (NS == 1 && R == 8 && !SWO && NSTR % 2 == 0), "swizzled first stage");
    static_assert(!(SZ & 2) || ((N / R) % 16 == 0 && !SWI && PR == 0), "swizzled second stage");
    static_assert(!(SZ & 4) || (NS == 8 && NS * R == 64 && !SWO && PR == 0), "swizzled stage-2 store");
    static_assert(!(SZ & 8) || (N / R == 64 && !SWI && PR != 1 && PR != 4), "swizzled stage-3 load");
    static_assert(!(SZ & 64) || (NS == 1 && R == 8 && !SWO && NSTR % 2 == 0), "swizzled first stage");
    static_assert(!(SZ & 128) || ((N / R) % 8 == 0 && !SWI && PR == 0), "swizzled second stage");
    static_assert((PR != 1 && PR != 4) || (R == 8 && NS == 1 && !INV && !SWI),
                  "padded first stage");
    static_assert((PR != 2 && PR != 3) || (NS * R == N && INV && !SWO && N % 4 == 0),
                  "cropped last stage");
    // LB: barriers wait for this wavefront's LDS traffic only, not for its
    // global loads and stores (global-to-LDS loads may be in flight)
    auto barrier = [] {
        if constexpr (LB) mr_lds_barrier();
        else __syncthreads();
    };
    // with LB (a persistent loop), the thread index is hidden behind an empty
    // asm so that the addressing derived from it is recomputed per stage
    // instead of held in registers across the loop
    unsigned tid = threadIdx.x;
    if constexpr (LB) asm volatile("" : "+v"(tid));
    constexpr int m = N / R;
    constexpr int KMAX = (TPMAX * m + NT - 1) / NT;
    constexpr int RW = (R > 1) ? R - 1 : 1;
    static_assert(mr_cdiv_exact<m>() && mr_cdiv_exact<NS>(), "inexact constant division");
    // Butterfly f uses twiddle index (f mod m) mod NS = f mod NS (NS divides
    // m). With NS dividing NT, that is threadIdx.x mod NS in every round of a
    // thread: its twiddles are read, and their powers formed, once (SH).
    // Only a clamped lane (f past nb - 1 in a later round) would then
    // multiply with another butterfly's twiddles; it stores nothing. A
    // wavefront whose first round has clamped lanes has no later round
    // (nb <= 64 wave + 63 < NT + 64 wave), so its twiddles are its own.
    constexpr bool SH = NS > 1 && KMAX > 1 && NT % NS == 0;
    constexpr int KW = SH ? 1 : KMAX;
    const int nb = tp * m;
    float2 v[KMAX][R];
    float2 w[KW][TW::kPowers ? 1 : RW];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k * NT + wave * 64 < nb) {
            const unsigned f = min((tid + k * NT), (unsigned)(nb - 1));
            const unsigned p = mr_cdiv<m>::q(f), j = f - p * m;
            const float2 *a = A + p * NSTR + j;
            if constexpr (SWI) {
                const unsigned c = mr_ct_swz(p);
#pragma unroll
                for (int r = 0; r < R; ++r) v[k][r] = A[p * NSTR + ((j + r * m) ^ c)];
            } else if constexpr ((SZ & 128) != 0) {
                // position j + r m, m a multiple of 8: bits 3-4 are those of
                // (j >> 3) + r m / 8
                const unsigned jh = j >> 3;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    v[k][r] = A[p * NSTR + r * m + (j ^ (2u * ((jh + (unsigned)(r * (m / 8))) & 3u)))];
            } else if constexpr ((SZ & 2) != 0) {
                // position j + r m, with m a multiple of 16: bits 4-5 are
                // those of (j >> 4) + r m / 16
                const unsigned jh = j >> 4;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    v[k][r] = A[p * NSTR + r * m + (j ^ (2u * ((jh + (unsigned)(r * (m / 16))) & 3u)))];
            } else if constexpr (PR == 1) {
                v[k][0] = a[0];
                v[k][1] = a[m];
                v[k][6] = a[6 * m];
                v[k][7] = a[7 * m];
            } else if constexpr (PR == 4) {
                // the image stored half a pencil on (e + N/2 mod N): input r
                // of the butterfly sits at slot (r + 4 mod 8) m
                v[k][0] = a[4 * m];
                v[k][1] = a[5 * m];
                v[k][6] = a[2 * m];
                v[k][7] = a[3 * m];
            } else if constexpr ((SZ & 8) != 0) {
                // position j + 64 r, j < 64: bit 6 is r's bit 0
#pragma unroll
                for (int r = 0; r < R; ++r) v[k][r] = A[p * NSTR + r * m + (j ^ (8u * (r & 1)))];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[k][r] = a[r * m];
            }
            if constexpr (NS > 1) {
                if (k < KW) {
                    const unsigned kk = j - mr_cdiv<NS>::q(j) * NS;
                    if constexpr (TW::kPowers) {
                        w[k][0] = tw.get(NS + kk - 1, 0u);
                    } else {
#pragma unroll
                        for (int r = 1; r < R; ++r)
                            w[k][r - 1] = tw.get(NS + kk - 1, (unsigned)((r - 1) * NS));
                    }
                }
            }
        }
    }
    barrier();
    // twiddle powers W^r, r = 1 .. R-1, of a shared twiddle: each a product of
    // two lower ones (the same products, in the same order, as per round)
    float2 ws[SH && TW::kPowers ? RW : 1];
    if constexpr (SH && TW::kPowers) {
        if (wave * 64 < nb) {
            ws[0] = w[0][0];
#pragma unroll
            for (int r = 2; r < R; ++r) ws[r - 1] = cmul(ws[r / 2 - 1], ws[r - r / 2 - 1]);
        }
    }
    if constexpr (MID::active) {
        // the twiddle loads complete before the hook: a wait for them after
        // it would also wait for whatever memory traffic the hook starts
        static_assert(SH && TW::kPowers, "stage hook needs the shared twiddle form");
#pragma unroll
        for (int r = 0; r < RW; ++r) asm volatile("" : "+v"(ws[r].x), "+v"(ws[r].y));
        mid();
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k * NT + wave * 64 < nb) {
            const unsigned fr = tid + k * NT;
            const unsigned f = min(fr, (unsigned)(nb - 1));
            const unsigned p = mr_cdiv<m>::q(f), j = f - p * m;
            const unsigned jn = mr_cdiv<NS>::q(j), kk = j - jn * NS;
            if constexpr (NS > 1 && SH && TW::kPowers) {
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[k][r] = INV ? cmulc(ws[r - 1], v[k][r]) : cmul(ws[r - 1], v[k][r]);
            } else if constexpr (NS > 1 && TW::kPowers) {
                // each power applied as soon as it is formed, so few are live
                float2 wp[RW];
                wp[0] = w[k][0];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    if (r >= 2) wp[r - 1] = cmul(wp[r / 2 - 1], wp[r - r / 2 - 1]);
                    v[k][r] = INV ? cmulc(wp[r - 1], v[k][r]) : cmul(wp[r - 1], v[k][r]);
                }
            } else if constexpr (NS > 1) {
                constexpr int kw = 0;
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[k][r] = INV ? cmulc(w[SH ? kw : k][r - 1], v[k][r])
                                  : cmul(w[SH ? kw : k][r - 1], v[k][r]);
            }
            if constexpr (PR == 1 || PR == 4) bfly8_ends<INV>(v[k]);
            else bfly<R, INV>(v[k]);
            // a clamped lane of a shared-twiddle stage holds wrong values
            if (!SH || k == 0 || fr < (unsigned)nb) {
                float2 *b = A + p * NSTR + jn * (NS * R) + kk;
                if constexpr (SWO) {
                    const unsigned c = mr_ct_swz(p), o = jn * (NS * R) + kk;
#pragma unroll
                    for (int q = 0; q < R; ++q) A[p * NSTR + ((o + q * NS) ^ c)] = v[k][q];
                } else if constexpr ((SZ & 4) != 0) {
                    // output q at 64 jn + kk + 8 (q ^ s), s = jn & 1 (bit 6)
                    const unsigned sft = 8u * (jn & 1u);
                    float2 *be = A + p * NSTR + jn * 64 + kk + sft;
                    float2 *bo = A + p * NSTR + jn * 64 + kk - sft;
#pragma unroll
                    for (int q = 0; q < R; ++q) ((q & 1) ? bo : be)[q * 8] = v[k][q];
                } else if constexpr ((SZ & 65) != 0) {
                    // outputs 8j .. 8j + 7 (NS = 1: jn = j, kk = 0), 16-byte
                    // pairs at 8j + (2c ^ f), f = 2 ((j >> 1) & 3) (bit 1) or
                    // 2 (j & 3) (bit 64)
                    const unsigned f = (SZ & 1) ? 2u * ((j >> 1) & 3u) : 2u * (j & 3u);
                    float2 *row = A + p * NSTR + 8 * j;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const mr_u32x4_ct w4 = {__float_as_uint(v[k][2 * c].x),
                                                __float_as_uint(v[k][2 * c].y),
                                                __float_as_uint(v[k][2 * c + 1].x),
                                                __float_as_uint(v[k][2 * c + 1].y)};
                        *reinterpret_cast<mr_u32x4_ct *>(row + ((2u * c) ^ f)) = w4;
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < R; ++q) {
                        constexpr int Q4 = N / 4;
                        const bool in = NS * q + NS <= Q4 || NS * q >= 3 * Q4;
                        const bool out = NS * q >= Q4 && NS * q + NS <= 3 * Q4;
                        if constexpr (PR == 3) {
                            // exactly the kept slots: the others may be in use
                            const unsigned e = kk + NS * q;
                            if (in || (!out && (e < (unsigned)Q4 || e >= 3u * Q4)))
                                b[q * NS] = v[k][q];
                        } else if (PR != 2 || !out) {
                            b[q * NS] = v[k][q];
                        }
                    }
                }
            }
        }
    }
    barrier();
}

// Stages S .. SE - 1 of the transform (SE = 99: to the last); LB as in
// mr_ct_stage.
// Z: the first two stages exchange data in the bank-spread order of
// mr_ct_stage's SZ (where mr_ct_can_swz allows it).
template <int N> constexpr bool mr_ct_can_swz()
{
    return mr_ct_factor(N).nst >= 2 && mr_ct_factor(N).r[0] == 8 &&
           (N / mr_ct_factor(N).r[1]) % 16 == 0 && N % 2 == 0;
}
// The second and third stages exchange data in the bank-spread order of
// SZ bits 4 and 8: factors 8, 8, R with N / R = 64.
template <int N> constexpr bool mr_ct_can_swz2()
{
    return mr_ct_factor(N).nst == 3 && mr_ct_factor(N).r[0] == 8 && mr_ct_factor(N).r[1] == 8 &&
           N / mr_ct_factor(N).r[2] == 64;
}
// SZ of stage S when the transform uses the bank-spread orders (Z).
template <int N> constexpr bool mr_ct_can_swz8()
{
    return !mr_ct_can_swz<N>() && mr_ct_factor(N).nst >= 2 && mr_ct_factor(N).r[0] == 8 &&
           (N / mr_ct_factor(N).r[1]) % 8 == 0 && N % 2 == 0;
}
template <int N, bool Z> constexpr int mr_ct_sz(int S)
{
    if (!Z) return 0;
    int sz = 0;
    if (mr_ct_can_swz<N>()) sz |= (S == 0 ? 1 : S == 1 ? 2 : 0);
    if (mr_ct_can_swz8<N>()) sz |= (S == 0 ? 64 : S == 1 ? 128 : 0);
    if (mr_ct_can_swz2<N>()) sz |= (S == 1 ? 4 : S == 2 ? 8 : 0);
    return sz;
}
template <int N, int NSTR, int S, bool INV, int NT, int TPMAX, class TW, bool SW = false, int SE = 99,
          bool LB = false, bool Z = false>
__device__ __forceinline__ void mr_ct_fft(float2 *A, const TW &tw, int tp, int wave)
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    static_assert(F.rest == 1, "length must be 2^a 3^b 5^c 7^d");
    static_assert(!SW || (N % 32 == 0 && F.nst >= 2), "swizzled order needs 32 | N, 2 stages");
    static_assert(!Z || !SW, "bank-spread orders with the wide-tile swizzle");
    if constexpr (S < F.nst && S < SE) {
        mr_ct_stage<N, NSTR, F.r[S], F.ns[S], INV, NT, TPMAX, TW, SW && S == 0,
                    SW && S == F.nst - 1, 0, LB, mr_ct_nomid, mr_ct_sz<N, Z>(S)>(A, tw, tp, wave);
        mr_ct_fft<N, NSTR, S + 1, INV, NT, TPMAX, TW, SW, SE, LB, Z>(A, tw, tp, wave);
    }
}

// Transform of centred zero-padded pencils whose result is cropped the same
// way (image points at e < N/4 and e >= 3N/4): forward with the padded first
// stage, inverse with the cropped last stage; plain order, first radix 8.
template <int N> constexpr bool mr_ct_can_trim()
{
    return mr_ct_factor(N).r[0] == 8 && N % 32 == 0 && mr_ct_factor(N).nst >= 2;
}
template <int N, int NSTR, bool INV, int NT, int TPMAX, class TW>
__device__ __forceinline__ void mr_ct_fft_trim(float2 *A, const TW &tw, int tp, int wave)
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    constexpr bool Z = F.nst >= 3;
    static_assert(mr_ct_can_trim<N>(), "trimmed transform");
    constexpr int L = F.nst - 1;
    if constexpr (!INV) {
        mr_ct_stage<N, NSTR, F.r[0], F.ns[0], false, NT, TPMAX, TW, false, false, 1, false,
                    mr_ct_nomid, mr_ct_sz<N, Z>(0)>(A, tw, tp, wave);
        mr_ct_fft<N, NSTR, 1, false, NT, TPMAX, TW, false, 99, false, Z>(A, tw, tp, wave);
    } else {
        mr_ct_fft<N, NSTR, 0, true, NT, TPMAX, TW, false, L, false, Z>(A, tw, tp, wave);
        mr_ct_stage<N, NSTR, F.r[L], F.ns[L], true, NT, TPMAX, TW, false, false, 2, false,
                    mr_ct_nomid, mr_ct_sz<N, Z>(L)>(A, tw, tp, wave);
    }
}

// The same for pencils whose image is stored half a pencil on, in slots
// [N/4, 3N/4) (so that the kept slots [0, N/4) and [3N/4, N) of the previous
// row's result and the next row's image coexist): the forward transform reads
// the image there; the inverse writes exactly the kept slots and runs mid()
// once its last stage has read its inputs, when the other slots are free.
// Barriers wait for LDS traffic only.
template <int N, int NSTR, int NT, int TPMAX, class TW>
__device__ __forceinline__ void mr_ct_fft_mid_fwd(float2 *A, const TW &tw, int tp, int wave)
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    constexpr bool Z = F.nst >= 3;
    static_assert(mr_ct_can_trim<N>(), "trimmed transform");
    mr_ct_stage<N, NSTR, F.r[0], F.ns[0], false, NT, TPMAX, TW, false, false, 4, true, mr_ct_nomid,
                mr_ct_sz<N, Z>(0)>(A, tw, tp, wave);
    mr_ct_fft<N, NSTR, 1, false, NT, TPMAX, TW, false, 99, true, Z>(A, tw, tp, wave);
}
template <int N, int NSTR, int NT, int TPMAX, class TW, class MID>
__device__ __forceinline__ void mr_ct_fft_crop_inv(float2 *A, const TW &tw, int tp, int wave,
                                                   const MID &mid)
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    constexpr bool Z = F.nst >= 3;
    static_assert(mr_ct_can_trim<N>(), "trimmed transform");
    constexpr int L = F.nst - 1;
    mr_ct_fft<N, NSTR, 0, true, NT, TPMAX, TW, false, L, true, Z>(A, tw, tp, wave);
    mr_ct_stage<N, NSTR, F.r[L], F.ns[L], true, NT, TPMAX, TW, false, false, 3, true, MID,
                mr_ct_sz<N, Z>(L)>(
        A, tw, tp, wave, mid);
}

// Axis pass of the FFT engine at compile-time length N (same functor
// protocol and tiling as mr_fft_axis_fast): TX = 2^lgTX <= TPMAX pencils per
// workgroup, 512 threads, twiddles in LDS.
template <int N> constexpr int mr_ct_axis_tpmax()
{
    int t = 64;
    while (t > 1 && t * N > MR_FFT_EMAX) t /= 2;
    return t;
}

// Threads of the 16-pencil tile: 1024 from length 512 up (32 wavefronts per
// CU at two workgroups), 512 below, where 1024 would leave half the lanes of
// the butterfly rounds idle.
template <int N> constexpr int mr_w16_threads() { return N >= 512 ? 1024 : 512; }

// W16 (wide tile): 16 pencils per workgroup, so that a pencil group spans
// whole 128-byte lines when pencils are the contiguous axis. It holds only the
// pencils in LDS (stride N, 16·N·8 bytes: two workgroups per CU at N = 640)
// and reads twiddles from the global table (powers formed by products).
template <int N, bool W16> struct mr_ct_axis_form {
    static constexpr int NSTR = W16 ? N : ((N % 2 == 0) ? N + 1 : N);
    static constexpr int TPMAX = W16 ? 16 : mr_ct_axis_tpmax<N>();
    static constexpr int NTW = W16 ? 0 : N;  // twiddles held in LDS
};

template <int N, bool INV, class Load, class Store, bool W16 = false>
__global__ void __launch_bounds__(W16 ? mr_w16_threads<N>() : MR_FFT_FAST_THREADS)
mr_fft_axis_ct(const float2 *tws, Load ld, Store st, long long npencil, int lgTX,
               long long pbase)
{
    constexpr int NT = W16 ? mr_w16_threads<N>() : MR_FFT_FAST_THREADS;
    using FM = mr_ct_axis_form<N, W16>;
    constexpr int NSTR = FM::NSTR;
    constexpr int TPMAX = FM::TPMAX;
    extern __shared__ __align__(16) unsigned char mr_smem[];
    __shared__ typename Load::PB lpb[MR_FFT_MAXTX];
    __shared__ typename Store::PB spb[MR_FFT_MAXTX];
    const int TX = 1 << lgTX;
    float2 *twl = reinterpret_cast<float2 *>(mr_smem);
    float2 *A = twl + FM::NTW;
    float2 *acc = A + (size_t)TX * NSTR;
    const long long p0 = pbase + mr_xcd_block(blockIdx.x, gridDim.x) * TX;
    const int tp = (int)((npencil - p0 < TX) ? (npencil - p0) : TX);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    using TWT = typename std::conditional<W16, mr_tw_buf, mr_tw_lds>::type;
    TWT tw;
    if constexpr (W16) {
        tw = mr_tw_buf{mr_rsrc(tws, (unsigned)N * 8u)};
    } else {
        for (int i = threadIdx.x; i < N; i += NT) twl[i] = tws[i];
        tw = mr_tw_lds{twl};
    }
    const int nsub = ld.nsub();
    if constexpr (W16) {
        // thread t always serves pencil t mod 16 (NT is a multiple of 16):
        // its pencil descriptors stay in registers, and 16 lanes read or
        // write one 128-byte line of each row
        static_assert(!Load::kElemFast && !Store::kElemFast && NT % 16 == 0, "wide tile form");
        const int p = threadIdx.x & 15;
        const unsigned c = mr_ct_swz(p);
        const bool pv = p < tp;
        const long long pp = p0 + (pv ? p : 0);
        for (int sub = 0; sub < nsub; ++sub) {
            const typename Load::PB lb = ld.pencil(pp, sub);
            const typename Store::PB sb = st.pencil(pp, sub);
            if (pv)
                for (int e = threadIdx.x >> 4; e < N; e += NT / 16)
                    A[p * NSTR + (e ^ c)] = ld.get(lb, e);
            __syncthreads();
            mr_ct_fft<N, NSTR, 0, INV, NT, TPMAX, TWT, true>(A, tw, tp, wave);
            if (pv)
                for (int e = threadIdx.x >> 4; e < N; e += NT / 16)
                    st.put(sb, p, e, sub, A[p * NSTR + (e ^ c)], acc);
            __syncthreads();
        }
        st.flush(p0, tp, acc);
        return;
    }
    if constexpr (!Load::kElemFast && !Store::kElemFast) {
        // pencils are the contiguous axis: thread t always serves pencil
        // t mod TX (NT is a multiple of TX), so its descriptors stay in
        // registers; the elements each thread moves are those of the loops
        // below
        const int p = threadIdx.x & (TX - 1);
        const bool pv = p < tp;
        const long long pp = p0 + (pv ? p : 0);
        const int e0 = threadIdx.x >> lgTX, de = NT >> lgTX;
        for (int sub = 0; sub < nsub; ++sub) {
            const typename Load::PB lb = ld.pencil(pp, sub);
            const typename Store::PB sb = st.pencil(pp, sub);
            if (pv)
                for (int e = e0; e < N; e += de) A[p * NSTR + e] = ld.get(lb, e);
            __syncthreads();
            mr_ct_fft<N, NSTR, 0, INV, NT, TPMAX>(A, tw, tp, wave);
            if (pv)
                for (int e = e0; e < N; e += de) st.put(sb, p, e, sub, A[p * NSTR + e], acc);
            __syncthreads();
        }
        st.flush(p0, tp, acc);
        return;
    }
    for (int sub = 0; sub < nsub; ++sub) {
        if (threadIdx.x < tp) {
            lpb[threadIdx.x] = ld.pencil(p0 + threadIdx.x, sub);
            spb[threadIdx.x] = st.pencil(p0 + threadIdx.x, sub);
        }
        __syncthreads();
        if (Load::kElemFast) {
            for (int q = threadIdx.x; q < tp * N; q += NT) {
                const int p = (int)mr_cdiv<N>::q((unsigned)q), e = q - p * N;
                A[p * NSTR + e] = ld.get(lpb[p], e);
            }
        } else {
            for (int q = threadIdx.x; q < (N << lgTX); q += NT) {
                const int e = q >> lgTX, p = q & (TX - 1);
                if (p < tp) A[p * NSTR + e] = ld.get(lpb[p], e);
            }
        }
        __syncthreads();
        mr_ct_fft<N, NSTR, 0, INV, NT, TPMAX>(A, tw, tp, wave);
        if (Store::kElemFast) {
            for (int q = threadIdx.x; q < tp * N; q += NT) {
                const int p = (int)mr_cdiv<N>::q((unsigned)q), e = q - p * N;
                st.put(spb[p], p, e, sub, A[p * NSTR + e], acc);
            }
        } else {
            for (int q = threadIdx.x; q < (N << lgTX); q += NT) {
                const int e = q >> lgTX, p = q & (TX - 1);
                if (p < tp) st.put(spb[p], p, e, sub, A[p * NSTR + e], acc);
            }
        }
        __syncthreads();
    }
    st.flush(p0, tp, acc);
}

// ---------------------------------------------------------------------------
// Double-buffered 16-pencil axis pass with direct global-to-LDS loads.
//
// For passes whose 16 pencils of a tile are consecutive along the contiguous
// axis (row stride I a multiple of 16), every source row of a tile is one
// 128-byte line. A workgroup owns one CU (two 16·N·8-byte buffers, 160 KB at
// N = 640) and walks tiles t, t + grid, ...: while it transforms and stores
// tile t from one buffer, the rows of tile t + grid stream into the other
// with global_load_lds_dwordx4 (16 bytes per lane straight into LDS at
// base + 16·lane, no VGPRs), eight rows per wavefront instruction. Buffers
// hold [element][16 pencils], so a wavefront's reads and writes in the
// stages are 128-byte rows of 16 pencils (no bank conflicts, no swizzle).
// The butterflies, twiddles and their products are those of mr_ct_stage,
// so results are bit-identical to the one-tile kernels.

// Source rows of a tile and the elements they fill: row s (0 <= s < rows)
// of pencil p0 lands at element (s - c) mod N.
template <class Load> struct mr_dma_rows { static constexpr bool kOk = false; };
template <> struct mr_dma_rows<mr_ld_pad<float2, false>> {
    static constexpr bool kOk = true;
    using L = mr_ld_pad<float2, false>;
    __host__ __device__ static int rows(const L &l) { return l.Ns; }
    __host__ __device__ static int shift(const L &l) { return l.c; }
    __host__ __device__ static long long stride(const L &l) { return l.I; }
    __host__ __device__ static const float2 *base(const L &l) { return l.in; }
    __device__ static const float2 *row0(const L &l, long long p0) { return l.in + l.pencil(p0, 0); }
};
template <> struct mr_dma_rows<mr_ld_axis<float2, false>> {
    static constexpr bool kOk = true;
    using L = mr_ld_axis<float2, false>;
    __host__ __device__ static int rows(const L &l) { return l.n; }
    __host__ __device__ static int shift(const L &) { return 0; }
    __host__ __device__ static long long stride(const L &l) { return l.I; }
    __host__ __device__ static const float2 *base(const L &l) { return l.in; }
    __device__ static const float2 *row0(const L &l, long long p0) { return l.in + l.pencil(p0, 0); }
};
template <class Store> struct mr_dma_store { static constexpr bool kOk = false; };
template <> struct mr_dma_store<mr_st_axis<float2, false>> { static constexpr bool kOk = true; };
template <> struct mr_dma_store<mr_st_crop<float2, false>> { static constexpr bool kOk = true; };


// Butterflies per work-item and stage, and stages, at most (the
// double-buffered passes: lengths 320, 512 and 640 with 16 pencils, 56 with
// 128 pencils).
constexpr int MR_EP_KMAX = 2;
constexpr int MR_EP_NST = 4;

constexpr int mr_ep_lg(int v) { return v <= 1 ? 0 : 1 + mr_ep_lg(v / 2); }

// The twiddle of each of a work-item's butterflies, for every stage of a
// tile of PT pencils: the same for every tile, so they are loaded once per
// workgroup and no vector memory load (which completes in order) waits
// behind the next tile's rows.
template <int N, int S, int NT, int PT, class TW>
__device__ __forceinline__ void mr_ep_twload(float2 (&w)[MR_EP_NST][MR_EP_KMAX], const TW &tw)
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    static_assert(F.rest == 1 && F.nst <= MR_EP_NST, "length must be 2^a 3^b 5^c 7^d");
    static_assert((PT & (PT - 1)) == 0, "pencils per tile: a power of two");
    if constexpr (S < F.nst) {
        constexpr int R = F.r[S], NS = F.ns[S], m = N / R, NB = PT * m;
        constexpr int KMAX = (NB + NT - 1) / NT;
        static_assert(KMAX <= MR_EP_KMAX, "butterflies per work-item");
        static_assert(mr_cdiv_exact<NS>(), "inexact constant division");
        if constexpr (NS > 1) {
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                const int f = (int)threadIdx.x + k * NT;
                if (f < NB) {
                    const unsigned j = (unsigned)(f >> mr_ep_lg(PT));
                    const unsigned kk = j - mr_cdiv<NS>::q(j) * NS;
                    w[S][k] = tw.get(NS + kk - 1, 0u);
                }
            }
        }
        mr_ep_twload<N, S + 1, NT, PT>(w, tw);
    }
}

// One Stockham stage over [element][PT pencils]: work-item f takes pencil
// f mod PT and butterfly f / PT, so neighbouring lanes touch one row. OOP:
// the inputs come from rd(pencil, element) instead of A, and no barrier
// separates the reads from the writes.
template <int N, int S, bool INV, int NT, int PT, bool OOP = false, class RD = int>
__device__ __forceinline__ void mr_ep_stage(float2 *A, const float2 (&ws)[MR_EP_KMAX],
                                            const RD &rd = 0)
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    constexpr int R = F.r[S], NS = F.ns[S];
    constexpr int m = N / R;
    constexpr int NB = PT * m;
    constexpr int KMAX = (NB + NT - 1) / NT;
    constexpr int RW = (R > 1) ? R - 1 : 1;
    constexpr int LG = mr_ep_lg(PT);
    float2 v[KMAX][R];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int f = (int)threadIdx.x + k * NT;
        if (f < NB) {
            const int p = f & (PT - 1), j = f >> LG;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (OOP) v[k][r] = rd(p, j + r * m);
                else v[k][r] = A[(j + r * m) * PT + p];
            }
        }
    }
    if constexpr (!OOP) mr_lds_barrier();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int f = (int)threadIdx.x + k * NT;
        if (f < NB) {
            const int p = f & (PT - 1), j = f >> LG;
            const unsigned jn = mr_cdiv<NS>::q((unsigned)j), kk = (unsigned)j - jn * NS;
            if constexpr (NS > 1) {
                float2 wp[RW];
                wp[0] = ws[k];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    if (r >= 2) wp[r - 1] = cmul(wp[r / 2 - 1], wp[r - r / 2 - 1]);
                    v[k][r] = INV ? cmulc(wp[r - 1], v[k][r]) : cmul(wp[r - 1], v[k][r]);
                }
            }
            bfly<R, INV>(v[k]);
#pragma unroll
            for (int q = 0; q < R; ++q) A[(jn * (NS * R) + kk + q * NS) * PT + p] = v[k][q];
        }
    }
    mr_lds_barrier();
}

// Stages S .. last, in place.
template <int N, int S, bool INV, int NT, int PT>
__device__ __forceinline__ void mr_ep_fft(float2 *A, const float2 (&w)[MR_EP_NST][MR_EP_KMAX])
{
    constexpr mr_ct_fac F = mr_ct_factor(N);
    if constexpr (S < F.nst) {
        mr_ep_stage<N, S, INV, NT, PT>(A, w[S]);
        mr_ep_fft<N, S + 1, INV, NT, PT>(A, w);
    }
}

template <int N, bool INV, class Load, class Store>
__global__ void __launch_bounds__(mr_w16_threads<N>())
mr_fft_axis_dma(const float2 *tws, Load ld, Store st, long long ntiles)
{
    constexpr int NT = mr_w16_threads<N>();
    constexpr int NW = NT / 64;
    using RW = mr_dma_rows<Load>;
    // rows land in B; the first stage reads B and writes A, the others work
    // in place on A. Two distinct LDS objects, so that the compiler sees that
    // the next tile's row loads into B cannot alias the stages' accesses to A
    // (and does not wait for them there).
    __shared__ __align__(16) float2 A[16 * N];
    __shared__ __align__(16) float2 B[16 * N];
    float2 w[MR_EP_NST][MR_EP_KMAX];
    mr_ep_twload<N, 0, NT, 16>(w, mr_tw_buf{mr_rsrc(tws, (unsigned)N * 8u)});
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ns = RW::rows(ld), c = RW::shift(ld);
    const long long rs = RW::stride(ld);
    // rows [0, na) land at elements N - c + s, rows [na, ns) at s - c; both
    // runs are whole groups of 8 rows (checked by the host)
    const int na = c < ns ? c : ns;
    const int ngrp = ns >> 3;
    long long t = blockIdx.x;
    if (t >= ntiles) return;
    // elements that no source row fills (zero padding): N - ns of them from
    // (ns - c) mod N on; the row loads never write them, so once is enough
    {
        int z0 = ns - c;
        if (z0 < 0) z0 += N;
        if (z0 >= N) z0 -= N;
        for (int i = threadIdx.x; i < (N - ns) * 16; i += NT) {
            int e = z0 + (i >> 4);
            if (e >= N) e -= N;
            B[e * 16 + (i & 15)] = make_float2(0.f, 0.f);
        }
    }
    auto fill = [&](long long tt) {
        const float2 *src = RW::row0(ld, tt * 16) + (lane & 7) * 2;
        for (int gi = wave; gi < ngrp; gi += NW) {
            const int s0 = gi * 8;
            const int e0 = s0 < na ? s0 - c + N : s0 - c;
            __builtin_amdgcn_global_load_lds(
                src + rs * (s0 + (lane >> 3)),
                (__attribute__((address_space(3))) void *)(B + e0 * 16), 16, 0, 0);
        }
    };
    fill(t);
    const int p = threadIdx.x & 15;
    for (; t < ntiles; t += gridDim.x) {
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's row loads into B
        __syncthreads();                 // everyone's (and the zero rows)
        mr_ep_stage<N, 0, INV, NT, 16, true>(A, w[0], [&](int pp, int e) { return B[e * 16 + pp]; });   // B -> A
        if (t + gridDim.x < ntiles) fill(t + gridDim.x);
        mr_ep_fft<N, 1, INV, NT, 16>(A, w);
        const typename Store::PB sb = st.pencil(t * 16 + p, 0);
        for (int e = threadIdx.x >> 4; e < N; e += NT / 16)
            st.put(sb, p, e, 0, A[e * 16 + p], nullptr);
        mr_lds_barrier();                // A is read out before the next first stage
    }
}

// Launch the double-buffered pass when the tile rows are 128-byte lines
// (returns false to use the one-tile kernels).
template <int N, bool INV, class Load, class Store>
bool mr_fft_pass_dma(const mr_fft_len<float> &f, const Load &ld, const Store &st,
                     long long npencil, hipStream_t stream)
{
    // lengths whose stages fit the registers without scratch: 320 (two
    // workgroups of 512 per CU), 512 and 640 (one of 1024)
    if constexpr (!mr_dma_rows<Load>::kOk || !mr_dma_store<Store>::kOk ||
                  !(N == 320 || N >= 512)) {
        return false;
    } else {
        using RW = mr_dma_rows<Load>;
        constexpr size_t lds = (size_t)2 * 16 * N * sizeof(float2);
        if (lds > 160 * 1024 || mr_env_ll("MR_AHA_HIP_FFT_DMA", 1) == 0) return false;
        const int ns = RW::rows(ld), c = RW::shift(ld);
        if (npencil % 16 != 0 || RW::stride(ld) % 16 != 0 || ns % 8 != 0 || c % 8 != 0 ||
            c < 0 || c >= N || ns > N || ((unsigned long long)RW::base(ld) & 15) != 0)
            return false;
        static int ncu = 0;
        if (ncu == 0) {
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                    hipSuccess || ncu <= 0)
                ncu = 256;
        }
        // workgroups per CU that fit the LDS (two at N <= 320)
        const long long ntiles = npencil / 16;
        const long long per_cu = (long long)(160 * 1024 / lds);
        const long long grid = ntiles < ncu * per_cu ? ntiles : ncu * per_cu;
        hipLaunchKernelGGL((mr_fft_axis_dma<N, INV, Load, Store>), dim3((unsigned)grid),
                           dim3(mr_w16_threads<N>()), 0, stream, f.tws, ld, st, ntiles);
        return true;
    }
}

// Host: launch the compile-time kernel for length N (tile choice as in
// mr_fft_pass).
template <int N, bool INV, class Load, class Store>
int mr_fft_pass_ct_n(const mr_fft_len<float> &f, const Load &ld, const Store &st,
                     long long npencil, size_t acc_per_pencil, int want_tx, hipStream_t stream,
                     const char *name)
{
    constexpr int NSTR = (N % 2 == 0) ? N + 1 : N;
    if (npencil <= 0) return MR_SUCCESS;
    int TX = want_tx > MR_FFT_MAXTX ? MR_FFT_MAXTX : want_tx;
    const long long txmax = mr_env_ll("MR_AHA_HIP_FFT_TXMAX", MR_FFT_MAXTX);
    if (TX > txmax) TX = (int)(txmax < 1 ? 1 : txmax);
    // 2^lgTX pencils per workgroup: round a width that is not a power of two
    // down before LDS and the grid are sized for it
    TX = mr_pow2_floor(TX);
    // a dispatch counts its work-items in 32 bits: larger grids are launched
    // in chunks of whole tiles (pbase: first pencil of the chunk)
    const long long maxb = mr_max_blocks(MR_FFT_FAST_THREADS);
    // wide tile: 16 pencils per workgroup where a tile of 16 is asked for, no
    // per-pencil accumulator is needed and 16 pencils take at most 80 KB
    constexpr size_t need16 = (size_t)16 * N * sizeof(float2);
    if constexpr (N >= 256 && need16 <= 80 * 1024 && !Load::kElemFast && !Store::kElemFast) {
        // dynamic LDS above 64 KB must be allowed per kernel, once
        static const bool w16_ok =
            hipFuncSetAttribute(reinterpret_cast<const void *>(
                                    &mr_fft_axis_ct<N, INV, Load, Store, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)need16) == hipSuccess;
        if (w16_ok && TX >= 16 && acc_per_pencil == 0 && npencil >= 16 &&
            mr_env_ll("MR_AHA_HIP_FFT_W16", 1) != 0) {
        if (mr_fft_pass_dma<N, INV>(f, ld, st, npencil, stream)) {
            MR_LAUNCH_CHECK(name);
            return MR_SUCCESS;
        }
        const long long blocks = (npencil + 15) / 16;
        const long long maxb16 = mr_max_blocks(mr_w16_threads<N>());
        for (long long b0 = 0; b0 < blocks; b0 += maxb16) {
            const long long nb = blocks - b0 < maxb16 ? blocks - b0 : maxb16;
            hipLaunchKernelGGL((mr_fft_axis_ct<N, INV, Load, Store, true>), dim3((unsigned)nb),
                               dim3(mr_w16_threads<N>()), need16, stream, f.tws, ld, st, npencil,
                               4, b0 * 16);
            MR_LAUNCH_CHECK(name);
        }
        return MR_SUCCESS;
        }
    }
    auto need = [&](int tx) {
        return ((size_t)N + (size_t)tx * NSTR + (size_t)tx * acc_per_pencil) * sizeof(float2);
    };
    while (TX > 1 && ((long long)TX * N > MR_FFT_EMAX || need(TX) > 64 * 1024)) TX /= 2;
    while (TX > 1 && TX / 2 >= npencil) TX /= 2;
    int lg = 0;
    while ((1 << lg) < TX) ++lg;
    const long long blocks = (npencil + TX - 1) / TX;
    for (long long b0 = 0; b0 < blocks; b0 += maxb) {
        const long long nb = blocks - b0 < maxb ? blocks - b0 : maxb;
        hipLaunchKernelGGL((mr_fft_axis_ct<N, INV, Load, Store>), dim3((unsigned)nb),
                           dim3(MR_FFT_FAST_THREADS), need(TX), stream, f.tws, ld, st, npencil, lg,
                           b0 * TX);
        MR_LAUNCH_CHECK(name);
    }
    return MR_SUCCESS;
}

// FP32 axis pass: the compile-time kernel for lengths 56, 256, 320, 384, 512
// and 640 (MR_AHA_HIP_FFT_CT=0 turns it off), mr_fft_pass otherwise.
template <bool INV, class Load, class Store>
int mr_fft_pass_ct(const mr_fft_len<float> &f, const Load &ld, const Store &st, long long npencil,
                   size_t acc_per_pencil, int want_tx, hipStream_t stream, const char *name)
{
    if (mr_env_ll("MR_AHA_HIP_FFT_CT", 1) != 0) {
        switch (f.n) {
#define MR_CT_LEN(n)                                                                             \
    case n:                                                                                      \
        return mr_fft_pass_ct_n<n, INV>(f, ld, st, npencil, acc_per_pencil, want_tx, stream, name);
            MR_CT_LEN(56) MR_CT_LEN(256) MR_CT_LEN(320) MR_CT_LEN(384) MR_CT_LEN(512) MR_CT_LEN(640)
#undef MR_CT_LEN
        default: break;
        }
    }
    return mr_fft_pass<float, INV>(f, ld, st, npencil, acc_per_pencil, want_tx, stream, name);
}
