// Fused x pass specialised at compile time for the common x grid lengths.
//
// Same dataflow as mr_k_xbd (one k-space row, G coils, L ranks per
// workgroup: zero-padded load, forward x-FFT, block multiply, inverse x-FFT,
// cropped store, in place), but the length N and the rank L are template
// parameters. Every stage radix, stride and divisor is then a constant:
// divisions become one full-rate 24-bit multiply and a shift, LDS offsets
// of a butterfly's inputs and outputs become instruction immediates, and no
// run-time plan occupies scalar registers. The twiddles of a stage are
// loaded with its inputs, before the stage's barrier, from the stage-major
// table (entry r*Ns + k - 1 of stage (R, Ns)).
//
// Instantiated for N in {256, 320, 384, 512, 640} and L in {4, 8, 16};
// other shapes run mr_k_xbd.

#include "mr_xbd.hpp"
#include "mr_fft_ct.hpp"

namespace {

// Coils one workgroup may take: LDS within 80 KB (two workgroups per CU)
// and G*L*N <= MR_XBD_EMAX.
constexpr int mr_xbd_ct_gmax(int N, int L)
{
    const int by_lds = 81920 / (L * N * 8);
    const int by_e = MR_XBD_EMAX / (L * N);
    const int g = by_lds < by_e ? by_lds : by_e;
    return g < 1 ? 1 : g;
}

// Locations of a coil-pair row walked per coil instead (0: none), so that the
// wavefronts of the block step load the four SIMDs evenly: in the 1024-thread
// two-coil form, N/64 pair wavefronts with N/64 = 2 mod 4 leave two SIMDs one
// walk ahead; 128 locations split into four one-coil wavefronts (half a pair
// walk each) even them out. Needs whole 64-location blocks (the 16-byte
// entry pairing) and 16 wavefronts.
constexpr int mr_xbd_split_locs(int N, int NT, int GMAX, bool masked)
{
    return (!masked && GMAX == 2 && NT == 1024 && N % 64 == 0 && (N / 64) % 4 == 2 &&
            N / 64 + 2 <= 16)
               ? 128
               : 0;
}

// Load the image points of one pencil into its N-point LDS row: source s
// lands at e = (s - cnx) mod N. Chunks of 64 points, NC of them, issued back
// to back; lanes past the Nx points (last chunk) read zero onto padding slots.
template <int N, int NC>
__device__ __forceinline__ void mr_xbd_load_chunks(float2 *Ap, __amdgpu_buffer_rsrc_t rs, int lane,
                                                   int cnx)
{
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int s = c * 64 + lane;
        int e = s - cnx;
        if (e < 0) e += N;
        Ap[e] = mr_bload(rs, (unsigned)s * 8u);
    }
}

// mr_xbd_load_chunks for nc = ceil(Nx / 64) chunks (wavefront-uniform).
template <int N, int NC = N / 64>
__device__ __forceinline__ void mr_xbd_load_src(float2 *Ap, __amdgpu_buffer_rsrc_t rs, int lane,
                                                int cnx, int nc)
{
    if constexpr (NC >= 1) {
        if (nc == NC) {
            mr_xbd_load_chunks<N, NC>(Ap, rs, lane, cnx);
            return;
        }
        mr_xbd_load_src<N, NC - 1>(Ap, rs, lane, cnx, nc);
    }
}

// The same with two points per lane and 16-byte loads and LDS writes (chunks
// of 128 points): for Nx and cnx even, points s, s+1 (s even) land in the
// adjacent slots e, e+1 (e even). Lanes past the Nx points write nothing.
typedef unsigned int mr_u32x4_ld __attribute__((ext_vector_type(4)));
template <int N, int NC>
__device__ __forceinline__ void mr_xbd_load_chunks2(float2 *Ap, __amdgpu_buffer_rsrc_t rs, int lane,
                                                    int cnx, int Nx)
{
    mr_u32x4_ld v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
        v[c] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((c * 128 + 2 * lane) * 8), 0, 0);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int s = c * 128 + 2 * lane;
        int e = s - cnx;
        if (e < 0) e += N;
        if (s < Nx) *reinterpret_cast<mr_u32x4_ld *>(Ap + e) = v[c];
    }
}

template <int N, int NC = (N + 127) / 128>
__device__ __forceinline__ void mr_xbd_load_src2(float2 *Ap, __amdgpu_buffer_rsrc_t rs, int lane,
                                                 int cnx, int Nx, int nc)
{
    if constexpr (NC >= 1) {
        if (nc == NC) {
            mr_xbd_load_chunks2<N, NC>(Ap, rs, lane, cnx, Nx);
            return;
        }
        mr_xbd_load_src2<N, NC - 1>(Ap, rs, lane, cnx, Nx, nc);
    }
}

// Block step of one k-space row (row r, gn coils of L pencils of length N in
// LDS at stride N, after the forward transform): every location's L x L
// Hermitian block times its L-vector, per coil, in place.
template <int N, int L, bool MASKED, int KF, int NT, int GMAX, bool MF>
__device__ __forceinline__ void mr_xbd_ct_block(float2 *A, const float2 *__restrict__ ker,
                                                const int *__restrict__ inv, long long kstride,
                                                const mr_dims &d, long long r, int gn,
                                                const unsigned *__restrict__ kq,
                                                const float *__restrict__ ksc, int wave,
                                                unsigned tid)
{
    constexpr int NW = NT / 64;
    const int lane = (int)(tid & 63u);
    const int My = d.m[1];
    const int jy = (int)(r % My), jz = (int)(r / My);
    int ky = jy + d.c_m[1]; if (ky >= My) ky -= My;
    int kz = jz + d.c_m[2]; if (kz >= d.m[2]) kz -= d.m[2];
    const long long rowc = (long long)N * (ky + (long long)My * kz);
    // kernel index of location e of the row (-1 outside the mask)
    auto kidx = [&](int e) -> long long {
        if (MASKED) return inv[e + (long long)N * r];
        int kx = e + d.c_m[0];
        if (kx >= N) kx -= N;
        return rowc + kx;
    };
    if constexpr (MF) {
        static_assert(!MASKED && KF == MR_KF_FP32 && N % 32 == 0, "MFMA block step");
        // a wavefront per 32 adjacent locations of a coil pair (an odd last
        // coil alone); adjacent locations have adjacent kernel entries (N and
        // the kx centering shift even), the kernel is 16-byte aligned
        constexpr int NB = N / 32;
        const int npair = (gn + 1) >> 1;
        for (int task = wave; task < npair * NB; task += NW) {
            const int qp = task / NB;
            const int e0 = (task - qp * NB) * 32;
            float2 *ap = A + 2 * qp * L * N;
            const long long kb = kidx(e0 + 2 * (lane >> 2));
            const bool two = 2 * qp + 1 < gn;
            if (d.tri_upper) {
                if (two) mr_bd_mfma32<L, true, 2>(ap, N, ker, kstride, kb, e0, lane);
                else mr_bd_mfma32<L, true, 1>(ap, N, ker, kstride, kb, e0, lane);
            } else {
                if (two) mr_bd_mfma32<L, false, 2>(ap, N, ker, kstride, kb, e0, lane);
                else mr_bd_mfma32<L, false, 1>(ap, N, ker, kstride, kb, e0, lane);
            }
        }
    } else if constexpr (KF == MR_KF_FP32 && GMAX >= 2) {
        // coils in pairs: a thread walks one location for two coils, loading
        // each kernel entry once for both; an odd last coil walks alone (the
        // 16-bit kernel, half the bytes per entry, measured no gain from it)
        const int npair = gn >> 1;
        const int nitem = ((gn + 1) >> 1) * N;
        // 16-byte entry loads over lane pairs (mr_bd_walk_pair_x2): unmasked,
        // so that a pair of adjacent locations has adjacent kernel entries
        // (N and the kx centering shift even: the pair never straddles the
        // kx wrap, and its even entry index is even) and a 16-byte aligned
        // kernel. Lane l of a 64-location block then takes location 2l
        // (l < 32) or 2(l - 32) + 1.
        const bool x2 = !MASKED && N % 64 == 0 && d.c_m[0] % 2 == 0 &&
                        ((unsigned long long)ker & 15) == 0;
        auto item = [&](int wi) {
            const int q = (int)mr_cdiv<N>::q((unsigned)wi);
            int e = wi - q * N;
            const int hi = (e & 63) >> 5;
            if (x2 && q < npair) e = (e & ~63) | (((e & 31) << 1) + hi);
            const long long b = kidx(e);
            float2 *a0 = A + 2 * q * L * N + e;
            if (q < npair) {
                float2 *a1 = a0 + L * N;
                if (b < 0) {
#pragma unroll
                    for (int l = 0; l < L; ++l) {
                        a0[l * N] = make_float2(0.f, 0.f);
                        a1[l * N] = make_float2(0.f, 0.f);
                    }
                } else if (!MASKED && x2) {
                    const unsigned bp = (unsigned)((b - hi) * 8 + (hi ? kstride * 8 : 0));
                    if (d.tri_upper)
                        mr_bd_walk_pair_x2<L, true, 4>(a0, a1, N, ker, kstride, bp, (unsigned)(b * 8));
                    else if constexpr (L % 2 == 0)
                        mr_bd_walk_pair_x2c<L, 4>(a0, a1, N, ker, kstride, bp, (unsigned)(b * 8));
                    else
                        mr_bd_walk_pair_x2<L, false, 4>(a0, a1, N, ker, kstride, bp, (unsigned)(b * 8));
                } else if (d.tri_upper) {
                    mr_bd_walk_pair<L, true, (L > 8 ? 8 : 16)>(a0, a1, N, ker, kstride, (unsigned)(b * 8));
                } else {
                    mr_bd_walk_pair<L, false, (L > 8 ? 8 : 16)>(a0, a1, N, ker, kstride, (unsigned)(b * 8));
                }
            } else if (b < 0) {
#pragma unroll
                for (int l = 0; l < L; ++l) a0[l * N] = make_float2(0.f, 0.f);
            } else if (d.tri_upper) {
                mr_bd_walk_g<L, true>(a0, N, ker, kstride, (unsigned)(b * 8));
            } else {
                mr_bd_walk_g<L, false>(a0, N, ker, kstride, (unsigned)(b * 8));
            }
        };
        if constexpr (((GMAX + 1) / 2) * N <= NT && mr_xbd_split_locs(N, NT, GMAX, MASKED) > 0) {
            // N/64 pair wavefronts would load the four SIMDs unevenly (N = 640:
            // 3, 3, 2, 2 walks). The last NS locations are walked per coil
            // instead, by 2 NS / 64 more wavefronts: wavefronts 0 .. NP/64 - 1
            // walk coil pairs, then coil 0 and coil 1 of locations NP .. N - 1,
            // so each SIMD holds the same share of FMAs (N = 640: two pair
            // walks and one one-coil walk, about half as long). The one-coil
            // wavefronts load the same entries as each other, mostly from L2.
            constexpr int NS = mr_xbd_split_locs(N, NT, GMAX, MASKED);
            constexpr int NP = N - NS;
            const int t = (int)tid;
            const bool split = x2 && gn == 2;
            if (t < (split ? NP : nitem)) {
                item(t);
            } else if (split && t < NP + 2 * NS) {
                const int u = t - NP;
                const int g = u / NS;
                int e = NP + (u - g * NS);
                const int hi = (e & 63) >> 5;
                e = (e & ~63) | (((e & 31) << 1) + hi);
                const long long b = kidx(e);
                float2 *a = A + g * L * N + e;
                const unsigned bp = (unsigned)((b - hi) * 8 + (hi ? kstride * 8 : 0));
                if (d.tri_upper)
                    mr_bd_walk_pair_x2<L, true, 8, 1>(a, a, N, ker, kstride, bp, (unsigned)(b * 8));
                else if constexpr (L % 2 == 0)
                    mr_bd_walk_pair_x2c<L, 8, 1>(a, a, N, ker, kstride, bp, (unsigned)(b * 8));
                else
                    mr_bd_walk_pair_x2<L, false, 8, 1>(a, a, N, ker, kstride, bp, (unsigned)(b * 8));
            }
        } else if constexpr (((GMAX + 1) / 2) * N <= NT) {
            // a thread per item: one round, no loop (a loop here keeps the
            // walk's column bases live across iterations and spills)
            if ((int)tid < nitem) item((int)tid);
        } else {
            for (int wi = (int)tid; wi < nitem; wi += NT) item(wi);
        }
    } else {
    // FP32, one walk per coil: entries two per 16-byte load over lane pairs
    // under the conditions of the pair walk
    const bool x1 = KF == MR_KF_FP32 && !MASKED && N % 64 == 0 && d.c_m[0] % 2 == 0 &&
                    ((unsigned long long)ker & 15) == 0;
    for (int wi = (int)tid; wi < gn * N; wi += NT) {
        const int g = (int)mr_cdiv<N>::q((unsigned)wi);
        int e = wi - g * N;
        const int hi = (e & 63) >> 5;
        if (x1) e = (e & ~63) | (((e & 31) << 1) + hi);
        const long long b = kidx(e);
        float2 *a = A + g * L * N + e;
        if (b < 0) {
#pragma unroll
            for (int l = 0; l < L; ++l) a[l * N] = make_float2(0.f, 0.f);
        } else if (KF == MR_KF_FP32 && !MASKED && x1) {
            const unsigned bp = (unsigned)((b - hi) * 8 + (hi ? kstride * 8 : 0));
            if (d.tri_upper)
                mr_bd_walk_pair_x2<L, true, 8, 1>(a, a, N, ker, kstride, bp, (unsigned)(b * 8));
            else if constexpr (L % 2 == 0)
                mr_bd_walk_pair_x2c<L, 8, 1>(a, a, N, ker, kstride, bp, (unsigned)(b * 8));
            else
                mr_bd_walk_pair_x2<L, false, 8, 1>(a, a, N, ker, kstride, bp, (unsigned)(b * 8));
        } else if constexpr (KF != MR_KF_FP32) {
            if (d.tri_upper)
                mr_bd_walk_q<L, true, KF>(a, N, kq, ksc, kstride, (unsigned)(b * 4), (unsigned)(b * 4));
            else
                mr_bd_walk_q<L, false, KF>(a, N, kq, ksc, kstride, (unsigned)(b * 4), (unsigned)(b * 4));
        } else if (d.tri_upper) {
            mr_bd_walk_g<L, true>(a, N, ker, kstride, (unsigned)(b * 8));
        } else {
            mr_bd_walk_g<L, false>(a, N, ker, kstride, (unsigned)(b * 8));
        }
    }
    }
}

// NT threads, up to GMAX coils per workgroup (the default: two workgroups of
// 512 threads per CU; the wide form: 1024 threads and two coils in 160 KB,
// one workgroup per CU). Either way a CU holds 16 wavefronts.
// MF: the block step on the matrix cores (mr_bd_mfma32; FP32 kernel,
// unmasked, 16-byte aligned kernel), a wavefront per 32 locations of a coil
// pair.
template <int N, int L, bool MASKED, int KF, int NT = MR_XBD_THREADS,
          int GMAX = mr_xbd_ct_gmax(N, L), bool MF = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
mr_k_xbd_ct(float2 *__restrict__ Y, const float2 *__restrict__ ker, const int *__restrict__ inv,
            long long kstride, mr_dims d, const float2 *tws, long long R, int G, int ncoil, int ncg,
            const unsigned *__restrict__ kq, const float *__restrict__ ksc)
{
    constexpr int NW = NT / 64;
    constexpr int TPMAX = GMAX * L;
    extern __shared__ __align__(16) unsigned char mr_smem[];
    float2 *A = reinterpret_cast<float2 *>(mr_smem);
    const int Nx = d.n[0], cnx = d.c_n[0];
    const long long q = mr_xcd_block(blockIdx.x, gridDim.x);
    const long long r = q / ncg;
    const int c0 = (int)(q - r * ncg) * G;
    const int gn = (ncoil - c0 < G) ? ncoil - c0 : G;
    const int tp = gn * L;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const mr_tw_buf twb{mr_rsrc(tws, (unsigned)N * 8u)};
    // pencil pl = g*L + l of the workgroup is volume v = L*c0 + pl; every
    // wavefront loads and stores whole pencils, so a pencil's base is uniform
    const float2 *Yr = Y + (long long)Nx * r + (long long)Nx * R * L * c0;
    const long long pstride = (long long)Nx * R;
    const unsigned rowbytes = (unsigned)Nx * 8u;

    // Only the Nx image points of a pencil are loaded (padding costs no load
    // instructions, which are what bounds this kernel); the N - Nx padding
    // slots e = (s - cnx) mod N, s >= Nx, are written with zeros.
    const int nck = (Nx + 63) >> 6;
    // two points per lane when the pencils and their slots pair up (Nx a
    // multiple of 4, so Nx and cnx = Nx/2 are even; 16-byte aligned rows)
    const bool px = (Nx & 3) == 0 && ((unsigned long long)Y & 15) == 0;
    // the image fills e < N/4 and e >= 3N/4 exactly (N = 2 Nx, centre Nx/2):
    // the transforms skip the zeros between and the cropped outputs, and the
    // padding slots, which the first stage then never reads, stay unwritten
    const bool trim = mr_ct_can_trim<N>() && 2 * Nx == N && 4 * cnx == N;
    for (int pl = wave; pl < tp; pl += NW) {
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, rowbytes);
        float2 *Ap = A + pl * N;
        if (px)
            mr_xbd_load_src2<N>(Ap, rs, lane, cnx, Nx, (Nx + 127) >> 7);
        else
            mr_xbd_load_src<N>(Ap, rs, lane, cnx, nck);
        if (!trim)
            for (int i = lane; i < N - Nx; i += 64) {
                int e = Nx + i - cnx;
                if (e >= N) e -= N;
                if (e < 0) e += N;
                Ap[e] = make_float2(0.f, 0.f);
            }
    }
    __syncthreads();
    if constexpr (mr_ct_can_trim<N>()) {
        if (trim)
            mr_ct_fft_trim<N, N, false, NT, TPMAX>(A, twb, tp, wave);
        else
            mr_ct_fft<N, N, 0, false, NT, TPMAX, mr_tw_buf, false, 99, false, true>(A, twb, tp, wave);
    } else {
        mr_ct_fft<N, N, 0, false, NT, TPMAX, mr_tw_buf, false, 99, false, true>(A, twb, tp, wave);
    }

    mr_xbd_ct_block<N, L, MASKED, KF, NT, GMAX, MF>(A, ker, inv, kstride, d, r, gn, kq, ksc, wave,
                                                   threadIdx.x);
    __syncthreads();
    if constexpr (mr_ct_can_trim<N>()) {
        if (trim)
            mr_ct_fft_trim<N, N, true, NT, TPMAX>(A, twb, tp, wave);
        else
            mr_ct_fft<N, N, 0, true, NT, TPMAX, mr_tw_buf, false, 99, false, true>(A, twb, tp, wave);
    } else {
        mr_ct_fft<N, N, 0, true, NT, TPMAX, mr_tw_buf, false, 99, false, true>(A, twb, tp, wave);
    }

    for (int pl = wave; pl < tp; pl += NW) {
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, rowbytes);
        const float2 *Ap = A + pl * N;
        if (px) {
            for (int s = 2 * lane; s < Nx; s += 128) {
                int e = s - cnx;
                if (e < 0) e += N;
                __builtin_amdgcn_raw_buffer_store_b128(
                    *reinterpret_cast<const mr_u32x4_ld *>(Ap + e), rs, s * 8, 0, 0);
            }
        } else {
            for (int s = lane; s < Nx; s += 64) {
                int e = s - cnx;
                if (e < 0) e += N;
                mr_bstore(rs, (unsigned)s * 8u, Ap[e]);
            }
        }
    }
}

// Persistent form of the two-coil 1024-thread fused pass, for a centred image
// filling half the x grid (N = 2 Nx, centre Nx/2, Nx a multiple of 4, 16-byte
// aligned rows, FP32 kernel). A workgroup per CU walks (row, coil-pair)
// items, interleaved with the other workgroups of its XCD, and keeps each item's
// image half a pencil on, in slots [N/4, 3N/4). The last inverse stage reads
// every slot and then writes only the kept ones, [0, N/4) and [3N/4, N), so
// from that point the next item's image rows stream straight from HBM into
// the middle slots (global_load_lds_dwordx4, no registers) while this item's
// last stage and its stores run: the load latency that a workgroup of the
// one-item form waits out before its first transform is hidden. The LDS
// loads are written in asm so that the compiler, which does not see them,
// adds no wait for them to the LDS reads of the kept slots; the kernel waits
// for them itself (vmcnt counts them with the stores issued after them, and
// completes in order) before the next item's transform reads them.
template <int N, int L, bool MASKED>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4)))
mr_k_xbd_pd(float2 *__restrict__ Y, const float2 *__restrict__ ker, const int *__restrict__ inv,
            long long kstride, mr_dims d, const float2 *tws, long long R, int G, int ncoil, int ncg,
            const unsigned *__restrict__ kq, const float *__restrict__ ksc)
{
    constexpr int NT = 1024, NW = NT / 64, GMAX = 2, TPMAX = GMAX * L;
    constexpr int NX = N / 2, NQ = N / 4;
    constexpr int SPP = (NX + 127) / 128;          // 16-byte stores per pencil and wavefront
    constexpr int NBLK = (NX * 8 + 1023) / 1024;   // 1 KB LDS loads per pencil
    static_assert(mr_ct_can_trim<N>() && NX % 4 == 0 && 2 * SPP <= 63, "persistent fused form");
    extern __shared__ __align__(16) unsigned char mr_smem[];
    float2 *A = reinterpret_cast<float2 *>(mr_smem);
    const unsigned lds0 = (unsigned)(size_t)A;
    // lane index read afresh where used (see tid below)
    auto lane_now = [] {
        unsigned t = threadIdx.x;
        asm volatile("" : "+v"(t));
        return (int)(t & 63u);
    };
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const mr_tw_buf twb{mr_rsrc(tws, (unsigned)N * 8u)};
    const long long pstride = (long long)NX * R;
    // XCD x (blockIdx.x mod 8) owns a contiguous share [qa, qb) of the items,
    // in proportion to its workgroups; its cx workgroups take every cx-th
    // item, so at any time they work on neighbouring items, mostly the coil
    // groups of the same few rows, whose kernel rows its L2 then holds once
    const int nxcd = 8, xcd = blockIdx.x % nxcd, iw = blockIdx.x / nxcd;
    const int g8 = gridDim.x / nxcd, g8r = gridDim.x % nxcd;
    const int cx = g8 + (xcd < g8r ? 1 : 0);
    const long long before = (long long)xcd * g8 + (xcd < g8r ? xcd : g8r);
    const long long nq = R * ncg;
    const long long qa = nq * before / gridDim.x + iw, qb = nq * (before + cx) / gridDim.x;
    if (qa >= qb) return;
    // the image rows of item q into slots [N/4, 3N/4) of its pencils
    auto fetch = [&](long long q) {
        const long long r = q / ncg;
        const int c0 = (int)(q - r * ncg) * G;
        const int tp = ((ncoil - c0 < G) ? ncoil - c0 : G) * L;
        const char *Yr = reinterpret_cast<const char *>(Y + NX * r + NX * R * L * c0);
        const int lane = lane_now();
        for (int t = wave; t < tp * NBLK; t += NW) {
            const int pl = t / NBLK, b = t - pl * NBLK;
            const unsigned off = (unsigned)(b * 1024 + lane * 16);
            if (off < NX * 8u) {
                const char *src = Yr + pstride * pl * 8 + off;
                const unsigned m0 = __builtin_amdgcn_readfirstlane(
                    lds0 + (unsigned)(pl * N + NQ) * 8u + (unsigned)b * 1024u);
                // m0 is listed as written (nothing else in this kernel keeps a
                // value in it; clang warns that it is a reserved register)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
                asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                             :
                             : "s"(m0), "v"(src)
                             : "memory", "m0");
#pragma clang diagnostic pop
            }
        }
    };
    fetch(qa);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    mr_lds_barrier();
    for (long long q = qa; q < qb; q += cx) {
        const long long r = q / ncg;
        const int c0 = (int)(q - r * ncg) * G;
        const int gn = (ncoil - c0 < G) ? ncoil - c0 : G;
        const int tp = gn * L;
        mr_ct_fft_mid_fwd<N, N, NT, TPMAX>(A, twb, tp, wave);
        // the thread index is re-read for every item (an empty asm hides that
        // it repeats), so what derives from it is recomputed, not held in
        // registers across the loop
        unsigned tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        mr_xbd_ct_block<N, L, MASKED, MR_KF_FP32, NT, GMAX, false>(A, ker, inv, kstride, d, r, gn,
                                                                  kq, ksc, wave, tid);
        mr_lds_barrier();
        const bool more = q + cx < qb;
        auto next = [&] {
            if (more) fetch(q + cx);
        };
        mr_ct_fft_crop_inv<N, N, NT, TPMAX>(A, twb, tp, wave, mr_ct_mid<decltype(next)>{next});
        const float2 *Yr = Y + NX * r + NX * R * L * c0;
        const int lane = lane_now();
        for (int pl = wave; pl < tp; pl += NW) {
            const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, (unsigned)NX * 8u);
            const float2 *Ap = A + pl * N;
#pragma unroll
            for (int i = 0; i < SPP; ++i) {
                const int s = 2 * lane + 128 * i;
                int e = s - NQ;
                if (e < 0) e += N;
                if (s < NX)
                    __builtin_amdgcn_raw_buffer_store_b128(
                        *reinterpret_cast<const mr_u32x4_ld *>(Ap + e), rs, s * 8, 0, 0);
            }
        }
        // the next item's rows have landed once at most this wavefront's own
        // stores, issued after them, are outstanding
        if (tp == 2 * NW) {
            if constexpr (SPP == 3) __builtin_amdgcn_s_waitcnt(0x0F76);
            else __builtin_amdgcn_s_waitcnt(0x0F70);
        } else if (tp == NW) {
            if constexpr (SPP == 3) __builtin_amdgcn_s_waitcnt(0x0F73);
            else __builtin_amdgcn_s_waitcnt(0x0F70);
        } else {
            __builtin_amdgcn_s_waitcnt(0x0F70);
        }
        mr_lds_barrier();
    }
}

template <int N> mr_xbd_ct_fn mr_xbd_ct_pick_n(int L, bool masked, int *gmax, bool mf)
{
    switch (L) {
#define MR_CT_CASE(l)                                                          \
    case l:                                                                    \
        *gmax = mr_xbd_ct_gmax(N, l);                                          \
        if (mf)                                                                \
            return masked ? nullptr                                            \
                          : mr_k_xbd_ct<N, l, false, MR_KF_FP32, MR_XBD_THREADS, \
                                        mr_xbd_ct_gmax(N, l), true>;           \
        return masked ? mr_k_xbd_ct<N, l, true, MR_KF_FP32>                    \
                      : mr_k_xbd_ct<N, l, false, MR_KF_FP32>;
        MR_CT_CASE(4) MR_CT_CASE(8) MR_CT_CASE(16)
#undef MR_CT_CASE
    default: return nullptr;
    }
}

// 16-bit kernels: unmasked, every compiled (N, L).
template <int KF> mr_xbd_ct_fn mr_xbd_ct_pick_q(int n, int L, int *gmax)
{
#define MR_CTQ(nn, l)                                                          \
    if (n == nn && L == l) {                                                   \
        *gmax = mr_xbd_ct_gmax(nn, l);                                         \
        return mr_k_xbd_ct<nn, l, false, KF>;                                  \
    }
    MR_CTQ(256, 4) MR_CTQ(256, 8) MR_CTQ(256, 16) MR_CTQ(320, 4) MR_CTQ(320, 8)
    MR_CTQ(320, 16) MR_CTQ(384, 4) MR_CTQ(384, 8) MR_CTQ(384, 16) MR_CTQ(512, 4)
    MR_CTQ(512, 8) MR_CTQ(512, 16) MR_CTQ(640, 4) MR_CTQ(640, 8) MR_CTQ(640, 16)
#undef MR_CTQ
    return nullptr;
}

// Quantise an FP32 kernel [T, kcol] into the 16-bit format.
template <int KF>
__global__ void mr_k_quantize(const float2 *__restrict__ ker, long long kcol, int L, int upper,
                              unsigned *__restrict__ q, float *__restrict__ sc)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kcol) return;
    float mx[10];
    const int nb = mr_kf_nblk(L);
    for (int b = 0; b < nb; ++b) mx[b] = 0.f;
    int t = 0;
    for (int c = 0; c < L; ++c) {
        const int r0 = upper ? 0 : c, r1 = upper ? c : L - 1;
        for (int r = r0; r <= r1; ++r, ++t) {
            const float2 e = ker[k + kcol * t];
            const int b = mr_kf_blk(r, c);
            mx[b] = fmaxf(mx[b], fmaxf(fabsf(e.x), fabsf(e.y)));
        }
    }
    for (int b = 0; b < nb; ++b) {
        if (!(mx[b] > 0.f)) mx[b] = 1.f;
        sc[k + kcol * b] = mx[b];
    }
    t = 0;
    for (int c = 0; c < L; ++c) {
        const int r0 = upper ? 0 : c, r1 = upper ? c : L - 1;
        for (int r = r0; r <= r1; ++r, ++t) {
            const float2 e = ker[k + kcol * t];
            const float inv = 1.f / mx[mr_kf_blk(r, c)];
            const float re = e.x * inv, im = e.y * inv;
            unsigned w;
            if (KF == MR_KF_BF16) {
                w = mr_bf16_rne(re) | (mr_bf16_rne(im) << 16);
            } else {
                typedef _Float16 h2 __attribute__((ext_vector_type(2)));
                h2 h;
                h.x = (_Float16)re;
                h.y = (_Float16)im;
                w = __builtin_bit_cast(unsigned, h);
            }
            q[k + kcol * t] = w;
        }
    }
}

}  // namespace

mr_xbd_ct_fn mr_xbd_ct_pick_wide(int n, int L, bool masked, int kf, bool mf)
{
    // FP32 only: the 16-bit walk moves half the bytes and measured slower in
    // the wide form (one workgroup per CU) than in two 80 KB workgroups
    if (L != 16 || (n != 640 && n != 512) || kf != MR_KF_FP32) return nullptr;
    if (mf) {
        if (masked) return nullptr;
        return n == 640 ? mr_k_xbd_ct<640, 16, false, MR_KF_FP32, 1024, 2, true>
                        : mr_k_xbd_ct<512, 16, false, MR_KF_FP32, 1024, 2, true>;
    }
    if (n == 640)
        return masked ? mr_k_xbd_ct<640, 16, true, MR_KF_FP32, 1024, 2>
                      : mr_k_xbd_ct<640, 16, false, MR_KF_FP32, 1024, 2>;
    return masked ? mr_k_xbd_ct<512, 16, true, MR_KF_FP32, 1024, 2>
                  : mr_k_xbd_ct<512, 16, false, MR_KF_FP32, 1024, 2>;
}

mr_xbd_ct_fn mr_xbd_ct_pick_persistent(int n, int L, bool masked, int kf)
{
    if (L != 16 || kf != MR_KF_FP32) return nullptr;
    if (n == 640) return masked ? mr_k_xbd_pd<640, 16, true> : mr_k_xbd_pd<640, 16, false>;
    if (n == 512) return masked ? mr_k_xbd_pd<512, 16, true> : mr_k_xbd_pd<512, 16, false>;
    return nullptr;
}

mr_xbd_ct_fn mr_xbd_ct_pick(int n, int L, bool masked, int kf, int *gmax, bool mf)
{
    *gmax = 0;
    if (kf == MR_KF_BF16)
        return (masked || mf) ? nullptr : mr_xbd_ct_pick_q<MR_KF_BF16>(n, L, gmax);
    if (kf == MR_KF_FP16)
        return (masked || mf) ? nullptr : mr_xbd_ct_pick_q<MR_KF_FP16>(n, L, gmax);
    switch (n) {
    case 256: return mr_xbd_ct_pick_n<256>(L, masked, gmax, mf);
    case 320: return mr_xbd_ct_pick_n<320>(L, masked, gmax, mf);
    case 384: return mr_xbd_ct_pick_n<384>(L, masked, gmax, mf);
    case 512: return mr_xbd_ct_pick_n<512>(L, masked, gmax, mf);
    case 640: return mr_xbd_ct_pick_n<640>(L, masked, gmax, mf);
    default: return nullptr;
    }
}

int mr_kf_quantize(int kf, const float2 *ker, long long kcol, int L, int upper, unsigned *q,
                   float *sc, hipStream_t stream)
{
    MR_REQUIRE(L >= 1 && L <= 16, "16-bit kernel format: L = %d outside [1, 16]", L);
    const unsigned blocks = (unsigned)((kcol + 255) / 256);
    if (kf == MR_KF_BF16)
        hipLaunchKernelGGL(mr_k_quantize<MR_KF_BF16>, dim3(blocks), dim3(256), 0, stream, ker, kcol,
                           L, upper, q, sc);
    else
        hipLaunchKernelGGL(mr_k_quantize<MR_KF_FP16>, dim3(blocks), dim3(256), 0, stream, ker, kcol,
                           L, upper, q, sc);
    MR_LAUNCH_CHECK("mr_k_quantize");
    return MR_SUCCESS;
}
