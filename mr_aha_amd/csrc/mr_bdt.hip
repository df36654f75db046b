// Block multiply for ranks above the compiled ones (L = 17 .. 512), in place
// over the padded k-space grids of the separate-pass route.
//
// Same operation as the reference's tiled kernels
// (src/cuda/mr_kernels_bdt.cu:218-416): at every grid location
// y_o = sum_r v_r A_ro with A Hermitian from its packed triangle, for each
// coil of the group. Layout here: K [P, L, ncoil] (natural order), kernel
// [T, kstride] centered (or [T, num_masked] with the inverse mask map).
//
// A workgroup of 512 work-items (eight wavefronts) owns a tile of TP = 16
// consecutive locations and a batch of NCB coils:
//  * the batch's inputs v[b][r][p] are staged in LDS once (location fastest:
//    the global reads and the LDS reads are both unit-stride over p);
//  * the stored triangle is walked in panels of whole stored columns (a
//    contiguous range of packed offsets); a panel is copied once from HBM
//    into LDS for the TP locations, tile[e][p], 16 consecutive kernel entries
//    (128 bytes) per packed offset;
//  * work-item (p, g) owns location p and outputs o = g, g + 32, ...; from
//    each panel it adds the column o (y_o += A_ro v_r, rows r of column o)
//    and the mirrors of the panel's other columns (y_o += conj(A_oc) v_c);
//    every output touches exactly L entries, so the work-items stay
//    balanced, and every stored entry is read from HBM once per coil batch
//    and from LDS by the two work-items whose outputs it feeds;
//  * each entry read from LDS serves the NCB coils of the batch from
//    registers.
// OPT (outputs per work-item, a power of two >= ceil(L/32)) and NCB are
// template parameters so that the accumulators stay in registers:
// NCB * OPT <= 16. The panel capacity keeps a workgroup within half of the
// LDS where the rank allows, so that two workgroups (16 wavefronts) share a
// CU and hide each other's LDS and HBM latency.

#include "mr_hip.hpp"

#include <algorithm>

namespace {

constexpr int BDT_THREADS = 512;
constexpr int BDT_TP = 16;                      // locations per tile
constexpr int BDT_NG = BDT_THREADS / BDT_TP;    // work-items per location

__host__ __device__ __forceinline__ int bdt_colstart(int c, int L, int upper)
{
    return upper ? c * (c + 1) / 2 : c * L - c * (c - 1) / 2;
}

__host__ __device__ __forceinline__ int bdt_collen(int c, int L, int upper)
{
    return upper ? c + 1 : L - c;
}

template <int OPT, int NCB, bool MASKED>
__global__ void __launch_bounds__(BDT_THREADS)
mr_k_bdt(float2 *__restrict__ K, const float2 *__restrict__ ker, const int *__restrict__ inv,
         long long kstride, mr_dims d, int ncoil, int E)
{
    extern __shared__ __align__(16) unsigned char mr_smem[];
    const int L = d.L;
    const int upper = d.tri_upper;
    float2 *vt = reinterpret_cast<float2 *>(mr_smem);             // [NCB][L][TP]
    float2 *pt = vt + (size_t)NCB * L * BDT_TP;                   // [E][TP]
    __shared__ long long kb[BDT_TP];
    const long long P = d.pad_size;
    const long long loc0 = mr_xcd_block(blockIdx.x, gridDim.x) * BDT_TP;
    const int tid = threadIdx.x;
    if (tid < BDT_TP) {
        const long long loc = loc0 + tid;
        long long b = -1;
        if (loc < P) b = MASKED ? (long long)inv[loc] : mr_nat_to_cent_lin(loc, d);
        kb[tid] = b;
    }
    const int p = tid % BDT_TP;
    const int g = tid / BDT_TP;
    const long long loc = loc0 + p;
    for (int cb = 0; cb < ncoil; cb += NCB) {
        const int nb = ncoil - cb < NCB ? ncoil - cb : NCB;
        __syncthreads();   // kb ready; the previous batch is done with vt and pt
        for (int f = tid; f < nb * L * BDT_TP; f += BDT_THREADS) {
            const int q = f / BDT_TP, pp = f - q * BDT_TP;       // q = b * L + r
            const long long l2 = loc0 + pp;
            vt[f] = (l2 < P) ? K[l2 + P * ((long long)L * cb + q)] : make_float2(0.f, 0.f);
        }
        float2 y[NCB][OPT];
#pragma unroll
        for (int b = 0; b < NCB; ++b)
#pragma unroll
            for (int j = 0; j < OPT; ++j) y[b][j] = make_float2(0.f, 0.f);
        for (int c0 = 0; c0 < L;) {
            // panel: whole stored columns c0 .. c1-1, at most E entries
            int c1 = c0, n = 0;
            while (c1 < L && n + bdt_collen(c1, L, upper) <= E) n += bdt_collen(c1++, L, upper);
            const int t0 = bdt_colstart(c0, L, upper);
            __syncthreads();   // the previous panel is consumed
            for (int f = tid; f < n * BDT_TP; f += BDT_THREADS) {
                const int e = f / BDT_TP, pp = f - e * BDT_TP;
                const long long b = kb[pp];
                pt[f] = (b >= 0) ? ker[b + kstride * (long long)(t0 + e)] : make_float2(0.f, 0.f);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < OPT; ++j) {
                const int o = g + BDT_NG * j;
                if (o >= L) continue;
                // column o: y_o += A_ro v_r over its stored rows r
                if (o >= c0 && o < c1) {
                    const int r0 = upper ? 0 : o;
                    const int len = bdt_collen(o, L, upper);
                    const float2 *e = pt + (size_t)(bdt_colstart(o, L, upper) - t0) * BDT_TP + p;
#pragma unroll 2
                    for (int i = 0; i < len; ++i) {
                        const float2 a = e[(size_t)i * BDT_TP];
#pragma unroll
                        for (int b = 0; b < NCB; ++b)
                            y[b][j] = cmac(y[b][j], a, vt[((size_t)b * L + r0 + i) * BDT_TP + p]);
                    }
                }
                // mirrors: entry (o, c) of stored column c != o in the panel
                const int m0 = upper ? (o + 1 > c0 ? o + 1 : c0) : c0;
                const int m1 = upper ? c1 : (o < c1 ? o : c1);
                for (int c = m0; c < m1; ++c) {
                    const int row = upper ? o : o - c;   // position of row o in column c
                    const float2 a = pt[(size_t)(bdt_colstart(c, L, upper) - t0 + row) * BDT_TP + p];
#pragma unroll
                    for (int b = 0; b < NCB; ++b)
                        y[b][j] = cmacc(y[b][j], a, vt[((size_t)b * L + c) * BDT_TP + p]);
                }
            }
            c0 = c1;
        }
        // the inputs of the batch are in LDS: the outputs go straight to HBM
        if (loc < P) {
            const bool out = MASKED ? kb[p] >= 0 : true;
#pragma unroll
            for (int j = 0; j < OPT; ++j) {
                const int o = g + BDT_NG * j;
                if (o >= L) continue;
#pragma unroll
                for (int b = 0; b < NCB; ++b)
                    if (b < nb)
                        K[loc + P * ((long long)L * (cb + b) + o)] =
                            out ? y[b][j] : make_float2(0.f, 0.f);
            }
        }
    }
}

typedef void (*mr_bdt_fn)(float2 *, const float2 *, const int *, long long, mr_dims, int, int);

template <int OPT, bool MASKED> mr_bdt_fn bdt_pick_ncb(int ncb)
{
    if constexpr (OPT <= 4) {
        if (ncb >= 4) return mr_k_bdt<OPT, 4, MASKED>;
    }
    if constexpr (OPT <= 8) {
        if (ncb >= 2) return mr_k_bdt<OPT, 2, MASKED>;
    }
    return mr_k_bdt<OPT, 1, MASKED>;
}

template <bool MASKED> mr_bdt_fn bdt_pick(int opt, int ncb)
{
    switch (opt) {
    case 1: return bdt_pick_ncb<1, MASKED>(ncb);
    case 2: return bdt_pick_ncb<2, MASKED>(ncb);
    case 4: return bdt_pick_ncb<4, MASKED>(ncb);
    case 8: return bdt_pick_ncb<8, MASKED>(ncb);
    default: return bdt_pick_ncb<16, MASKED>(ncb);
    }
}

}  // namespace

// Launch the tiled block multiply over K [P, L, nc]; returns MR_* status.
// MR_AHA_HIP_BDT_NCB caps the coil batch (1, 2 or 4; tests).
int mr_bdt_launch(const mr_dims &d, float2 *K, const float2 *ker, const int *inv,
                  long long kstride, int nc, bool masked, hipStream_t s)
{
    const int L = d.L;
    MR_REQUIRE(L >= 1 && L <= 512, "block multiply: rank L = %d outside 1..512", L);
    int opt = 1;
    while (opt * BDT_NG < L) opt *= 2;
    int ncb = 16 / opt;
    if (ncb > 4) ncb = 4;
    if (ncb > nc) ncb = nc;
    const long long cap = mr_env_ll("MR_AHA_HIP_BDT_NCB", 4);
    if (ncb > cap) ncb = (int)(cap < 1 ? 1 : cap);
    ncb = ncb >= 4 ? 4 : (ncb >= 2 ? 2 : 1);
    int dev = 0, maxlds = 0;
    MR_HIP_CHECK(hipGetDevice(&dev));
    MR_HIP_CHECK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const size_t vbytes = (size_t)ncb * L * BDT_TP * sizeof(float2);
    const size_t room = (size_t)maxlds - 256;   // static kb[]
    MR_REQUIRE(vbytes + (size_t)L * BDT_TP * sizeof(float2) <= room,
               "block multiply: L = %d does not fit the %d bytes of LDS", L, maxlds);
    // panel capacity: within half of the LDS (two workgroups per CU) when a
    // whole column still fits, else the rest of it
    const size_t half = (size_t)maxlds / 2 - 256;
    const size_t pbytes =
        (vbytes + (size_t)L * BDT_TP * sizeof(float2) <= half ? half : room) - vbytes;
    int E = (int)(pbytes / (BDT_TP * sizeof(float2)));
    const int T = L * (L + 1) / 2;
    if (E > T) E = T;
    const size_t lds = vbytes + (size_t)E * BDT_TP * sizeof(float2);
    mr_bdt_fn fn = masked ? bdt_pick<true>(opt, ncb) : bdt_pick<false>(opt, ncb);
    if (lds > 65536 &&
        hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        MR_FAIL(MR_ERR_UNSUPPORTED, "block multiply: %zu bytes of LDS refused", lds);
    }
    const long long blocks = mr_ceil_div(d.pad_size, BDT_TP);
    MR_GRID_CHECK(blocks, BDT_THREADS, "mr_k_bdt");
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(BDT_THREADS), lds, s, K, ker, inv, kstride,
                       d, nc, E);
    MR_LAUNCH_CHECK("mr_k_bdt");
    return MR_SUCCESS;
}
