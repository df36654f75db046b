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

template <int N, int L, bool MASKED>
__global__ void __launch_bounds__(MR_XBD_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
mr_k_xbd_ct(float2 *__restrict__ Y, const float2 *__restrict__ ker, const int *__restrict__ inv,
            long long kstride, mr_dims d, const float2 *tws, long long R, int G, int ncoil, int ncg)
{
    constexpr int NW = MR_XBD_THREADS / 64;
    constexpr int TPMAX = mr_xbd_ct_gmax(N, L) * L;
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

    for (int pl = wave; pl < tp; pl += NW) {
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, rowbytes);
        float2 *Ap = A + pl * N;
#pragma unroll
        for (int e0 = 0; e0 < N; e0 += 64) {
            const int e = e0 + lane;
            if (e < N) {
                int s = e + cnx;
                if (s >= N) s -= N;
                Ap[e] = mr_bload(rs, (unsigned)s * 8u);  // s >= Nx reads zero: the padding
            }
        }
    }
    __syncthreads();
    mr_ct_fft<N, N, 0, false, MR_XBD_THREADS, TPMAX>(A, twb, tp, wave);

    const int My = d.m[1];
    const int jy = (int)(r % My), jz = (int)(r / My);
    int ky = jy + d.c_m[1]; if (ky >= My) ky -= My;
    int kz = jz + d.c_m[2]; if (kz >= d.m[2]) kz -= d.m[2];
    const long long rowc = (long long)N * (ky + (long long)My * kz);
    for (int wi = threadIdx.x; wi < gn * N; wi += MR_XBD_THREADS) {
        const int g = (int)mr_cdiv<N>::q((unsigned)wi), e = wi - g * N;
        long long b;
        if (MASKED) {
            b = inv[e + (long long)N * r];
        } else {
            int kx = e + d.c_m[0];
            if (kx >= N) kx -= N;
            b = rowc + kx;
        }
        float2 *a = A + g * L * N + e;
        if (b < 0) {
#pragma unroll
            for (int l = 0; l < L; ++l) a[l * N] = make_float2(0.f, 0.f);
        } else if (d.tri_upper) {
            mr_bd_walk_g<L, true>(a, N, ker, kstride, (unsigned)(b * 8));
        } else {
            mr_bd_walk_g<L, false>(a, N, ker, kstride, (unsigned)(b * 8));
        }
    }
    __syncthreads();
    mr_ct_fft<N, N, 0, true, MR_XBD_THREADS, TPMAX>(A, twb, tp, wave);

    for (int pl = wave; pl < tp; pl += NW) {
        const __amdgpu_buffer_rsrc_t rs = mr_rsrc(Yr + pstride * pl, rowbytes);
        const float2 *Ap = A + pl * N;
        for (int s = lane; s < Nx; s += 64) {
            int e = s - cnx;
            if (e < 0) e += N;
            mr_bstore(rs, (unsigned)s * 8u, Ap[e]);
        }
    }
}

template <int N> mr_xbd_ct_fn mr_xbd_ct_pick_n(int L, bool masked, int *gmax)
{
    switch (L) {
#define MR_CT_CASE(l)                                                          \
    case l:                                                                    \
        *gmax = mr_xbd_ct_gmax(N, l);                                          \
        return masked ? mr_k_xbd_ct<N, l, true> : mr_k_xbd_ct<N, l, false>;
        MR_CT_CASE(4) MR_CT_CASE(8) MR_CT_CASE(16)
#undef MR_CT_CASE
    default: return nullptr;
    }
}

}  // namespace

mr_xbd_ct_fn mr_xbd_ct_pick(int n, int L, bool masked, int *gmax)
{
    *gmax = 0;
    switch (n) {
    case 256: return mr_xbd_ct_pick_n<256>(L, masked, gmax);
    case 320: return mr_xbd_ct_pick_n<320>(L, masked, gmax);
    case 384: return mr_xbd_ct_pick_n<384>(L, masked, gmax);
    case 512: return mr_xbd_ct_pick_n<512>(L, masked, gmax);
    case 640: return mr_xbd_ct_pick_n<640>(L, masked, gmax);
    default: return nullptr;
    }
}
