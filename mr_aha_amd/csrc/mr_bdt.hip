// Block multiply for ranks above the compiled ones (L = 17 .. 512), in place
// over the padded k-space grids of the separate-pass route.
//
// Same operation as the reference's tiled kernels
// (src/cuda/mr_kernels_bdt.cu:218-416): at every grid location
// y_o = sum_r v_r A_ro with A Hermitian from its packed triangle, for each
// coil of the group. Layout here: K [P, L, ncoil] (natural order), kernel
// [T, kstride] centered (or [T, num_masked] with the inverse mask map).
//
// Register blocking. A workgroup owns a tile of TP = 16 consecutive
// locations and a batch of NCB coils; work-item (p, g) owns location p and
// the OPT consecutive outputs o = g*OPT .. g*OPT+OPT-1 of all NCB coils.
//  * The batch's inputs are staged in LDS once, vt[r][b][p]: location
//    fastest, so the global reads are 128-byte lines and the LDS reads of the
//    four output groups of a wavefront are one broadcast line.
//  * The work-item walks r = 0 .. L-1 in order. For each r it loads the OPT
//    entries A_{r,o} of its outputs straight from the kernel (for a fixed
//    packed offset the 16 locations of the tile are one 128-byte line) and
//    the NCB inputs v_r from LDS, and does OPT x NCB complex multiply-adds:
//    one LDS read per OPT and one kernel load per NCB multiply-adds.
//  * Every work-item does exactly L steps, so the main loop has no barrier
//    and no idle lanes. A_{r,o} is row r of the stored column o on the stored
//    side of the diagonal, else the conjugate of row o of the stored column
//    r: three runs of r per output group (the rows before the group, the
//    OPT x OPT triangle at the diagonal, the rows after it); within a run the
//    addresses advance by a uniform amount per step.
//  * The coil batches of a tile are consecutive workgroups (and so on one
//    XCD): they run together and read the same kernel lines, which the first
//    brings into L2 for the others.
// Each output's sum runs over r in increasing order whatever the batch size,
// so results do not depend on it.

#include "mr_hip.hpp"

#include <algorithm>

namespace {

constexpr int BDT_TP = 16;   // locations per tile: one 128-byte line per entry

// packed offset of stored column c
__host__ __device__ __forceinline__ long long bdt_colstart(int c, int L, bool upper)
{
    return upper ? (long long)c * (c + 1) / 2 : (long long)c * L - (long long)c * (c - 1) / 2;
}

// packed offset of the stored entry (row, col): row >= col in the lower
// triangle, row <= col in the upper
__device__ __forceinline__ long long bdt_off(int row, int col, int L, bool upper)
{
    return bdt_colstart(col, L, upper) + (upper ? row : row - col);
}

// y[b][j] += a[j] v[b] (CJ: conj(a[j]) v[b]) for the NCB inputs at vr
template <int OPT, int NCB, bool CJ>
__device__ __forceinline__ void bdt_step(float2 (&y)[NCB][OPT], const float2 (&a)[OPT],
                                         const float2 *vr)
{
    float2 v[NCB];
#pragma unroll
    for (int b = 0; b < NCB; ++b) v[b] = vr[b * BDT_TP];
#pragma unroll
    for (int b = 0; b < NCB; ++b)
#pragma unroll
        for (int j = 0; j < OPT; ++j)
            y[b][j] = CJ ? cmacc(y[b][j], a[j], v[b]) : cmac(y[b][j], a[j], v[b]);
}

// Rows [r0, r1) whose OPT entries sit at kp[qo[j]], qo[j] advancing by
// `adv` per row (a run along the stored columns of the outputs). Two entry
// sets alternate, so the next row's loads are in flight while a row is
// multiplied.
template <int OPT, int NCB>
__device__ __forceinline__ void bdt_run_cols(float2 (&y)[NCB][OPT], const float2 *kp,
                                             long long (&qo)[OPT], long long adv,
                                             const float2 *vr, int r0, int r1)
{
    if (r0 >= r1) return;
    float2 a0[OPT], a1[OPT];
#pragma unroll
    for (int j = 0; j < OPT; ++j) a0[j] = kp[qo[j]];
    int r = r0;
#pragma unroll 1
    for (; r + 1 < r1; r += 2) {
#pragma unroll
        for (int j = 0; j < OPT; ++j) {
            qo[j] += adv;
            a1[j] = kp[qo[j]];
        }
        bdt_step<OPT, NCB, false>(y, a0, vr + (size_t)r * NCB * BDT_TP);
        if (r + 2 < r1) {
#pragma unroll
            for (int j = 0; j < OPT; ++j) {
                qo[j] += adv;
                a0[j] = kp[qo[j]];
            }
        }
        bdt_step<OPT, NCB, false>(y, a1, vr + (size_t)(r + 1) * NCB * BDT_TP);
    }
    if (r < r1) bdt_step<OPT, NCB, false>(y, a0, vr + (size_t)r * NCB * BDT_TP);
}

// Rows [r0, r1) whose OPT entries are rows oc[j] of the stored column r
// (conjugated): kp[cbase(r) + ko[j]]; loads alternate like the column runs.
template <int OPT, int NCB>
__device__ __forceinline__ void bdt_run_rows(float2 (&y)[NCB][OPT], const float2 *kp,
                                             const long long (&ko)[OPT], long long kstride,
                                             int L, bool upper, const float2 *vr, int r0, int r1)
{
    if (r0 >= r1) return;
    auto cbase = [&](int r) {
        return kstride * (bdt_colstart(r, L, upper) - (upper ? 0 : r));
    };
    float2 a0[OPT], a1[OPT];
    long long c = cbase(r0);
#pragma unroll
    for (int j = 0; j < OPT; ++j) a0[j] = kp[c + ko[j]];
    int r = r0;
#pragma unroll 1
    for (; r + 1 < r1; r += 2) {
        c = cbase(r + 1);
#pragma unroll
        for (int j = 0; j < OPT; ++j) a1[j] = kp[c + ko[j]];
        bdt_step<OPT, NCB, true>(y, a0, vr + (size_t)r * NCB * BDT_TP);
        if (r + 2 < r1) {
            c = cbase(r + 2);
#pragma unroll
            for (int j = 0; j < OPT; ++j) a0[j] = kp[c + ko[j]];
        }
        bdt_step<OPT, NCB, true>(y, a1, vr + (size_t)(r + 1) * NCB * BDT_TP);
    }
    if (r < r1) bdt_step<OPT, NCB, true>(y, a0, vr + (size_t)r * NCB * BDT_TP);
}

// Work-items per workgroup at most: 1024, or 512 for the 12-coil batch, whose
// accumulators need more than the 128 VGPRs a 1024-wide workgroup allows.
constexpr int bdt_max_threads(int ncb) { return ncb >= 12 ? 512 : 1024; }

template <int OPT, int NCB, bool MASKED>
__global__ void __launch_bounds__(bdt_max_threads(NCB))
mr_k_bdt(float2 *__restrict__ K, const float2 *__restrict__ ker, const int *__restrict__ inv,
         long long kstride, mr_dims d, int ncoil, int ng, int nbat)
{
    extern __shared__ __align__(16) unsigned char mr_smem[];
    float2 *vt = reinterpret_cast<float2 *>(mr_smem);   // [L][NCB][TP]
    const int L = d.L;
    const bool upper = d.tri_upper != 0;
    const long long P = d.pad_size;
    const long long lb = mr_xcd_block(blockIdx.x, gridDim.x);
    const long long tile = lb / nbat;
    const int cb = (int)(lb - tile * nbat) * NCB;
    const int nb = ncoil - cb < NCB ? ncoil - cb : NCB;
    const long long loc0 = tile * BDT_TP;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    for (int f = tid; f < L * NCB * BDT_TP; f += nthr) {
        const int pp = f & (BDT_TP - 1);
        const int q = f >> 4;                       // q = r * NCB + b
        const int r = q / NCB, b = q - r * NCB;
        const long long l2 = loc0 + pp;
        vt[f] = (l2 < P && b < nb) ? K[l2 + P * ((long long)L * (cb + b) + r)]
                                   : make_float2(0.f, 0.f);
    }
    const int p = tid & (BDT_TP - 1);
    const int g = tid >> 4;
    const long long loc = loc0 + p;
    long long kb = -1;
    if (loc < P) kb = MASKED ? (long long)inv[loc] : mr_nat_to_cent_lin(loc, d);
    __syncthreads();
    if (g >= ng) return;   // no barrier follows

    const int o0 = g * OPT;
    // outputs past L (last group) walk the entries of output L-1 and are not stored
    int oc[OPT];
    long long ko[OPT];
#pragma unroll
    for (int j = 0; j < OPT; ++j) {
        oc[j] = o0 + j < L ? o0 + j : L - 1;
        ko[j] = kstride * oc[j];
    }
    const float2 *kp = ker + (kb >= 0 ? kb : 0);
    const float2 *vr = vt + p;
    float2 y[NCB][OPT];
#pragma unroll
    for (int b = 0; b < NCB; ++b)
#pragma unroll
        for (int j = 0; j < OPT; ++j) y[b][j] = make_float2(0.f, 0.f);

    const int tr0 = o0, tr1 = o0 + OPT < L ? o0 + OPT : L;   // rows of the diagonal triangle
    long long qo[OPT];
    if (!upper) {
        // rows r < o0: conj(A_{o,r}), rows o of the stored column r
        bdt_run_rows<OPT, NCB>(y, kp, ko, kstride, L, false, vr, 0, tr0);
    } else {
        // rows r < o0: A_{r,o}, rows r of the stored column o
#pragma unroll
        for (int j = 0; j < OPT; ++j) qo[j] = kstride * bdt_colstart(oc[j], L, true);
        bdt_run_cols<OPT, NCB>(y, kp, qo, kstride, vr, 0, tr0);
    }
    // the diagonal triangle: rows o0 .. o0+OPT-1
    for (int r = tr0; r < tr1; ++r) {
        float2 a[OPT];
#pragma unroll
        for (int j = 0; j < OPT; ++j) {
            const int o = oc[j];
            const bool stored = upper ? r <= o : r >= o;
            const float2 e =
                kp[kstride * (stored ? bdt_off(r, o, L, upper) : bdt_off(o, r, L, upper))];
            a[j] = stored ? e : make_float2(e.x, -e.y);
        }
        bdt_step<OPT, NCB, false>(y, a, vr + (size_t)r * NCB * BDT_TP);
    }
    if (!upper) {
        // rows r >= o0+OPT: A_{r,o}, rows r of the stored column o
#pragma unroll
        for (int j = 0; j < OPT; ++j) qo[j] = kstride * bdt_off(tr1, oc[j], L, false);
        bdt_run_cols<OPT, NCB>(y, kp, qo, kstride, vr, tr1, L);
    } else {
        // rows r >= o0+OPT: conj(A_{o,r}), rows o of the stored column r
        bdt_run_rows<OPT, NCB>(y, kp, ko, kstride, L, true, vr, tr1, L);
    }

    if (loc < P) {
        const bool out = kb >= 0;
#pragma unroll
        for (int j = 0; j < OPT; ++j) {
            if (o0 + j >= L) continue;
#pragma unroll
            for (int b = 0; b < NCB; ++b)
                if (b < nb)
                    K[loc + P * ((long long)L * (cb + b) + o0 + j)] =
                        out ? y[b][j] : make_float2(0.f, 0.f);
        }
    }
}

typedef void (*mr_bdt_fn)(float2 *, const float2 *, const int *, long long, mr_dims, int, int,
                          int);

// Compiled (OPT, NCB) forms: OPT outputs x NCB coils per work-item (the
// accumulators take 4 x OPT x NCB VGPRs)
template <bool MASKED> mr_bdt_fn bdt_fn(int opt, int ncb)
{
    switch (opt * 100 + ncb) {
    case 401: return mr_k_bdt<4, 1, MASKED>;
    case 402: return mr_k_bdt<4, 2, MASKED>;
    case 404: return mr_k_bdt<4, 4, MASKED>;
    case 206: return mr_k_bdt<2, 6, MASKED>;
    case 208: return mr_k_bdt<2, 8, MASKED>;
    case 212: return mr_k_bdt<2, 12, MASKED>;
    case 801: return mr_k_bdt<8, 1, MASKED>;
    default: return nullptr;
    }
}

}  // namespace

// Launch the tiled block multiply over K [P, L, nc]; returns MR_* status.
// MR_AHA_HIP_BDT_NCB caps the coil batch and MR_AHA_HIP_BDT_OPT sets the
// outputs per work-item (tests and tuning); results do not depend on either.
int mr_bdt_launch(const mr_dims &d, float2 *K, const float2 *ker, const int *inv,
                  long long kstride, int nc, bool masked, hipStream_t s)
{
    const int L = d.L;
    MR_REQUIRE(L >= 1 && L <= 512, "block multiply: rank L = %d outside 1..512", L);
    MR_REQUIRE(nc >= 1, "block multiply: %d coils", nc);
    const long long cap = mr_env_ll("MR_AHA_HIP_BDT_NCB", 12);
    const long long want_opt = mr_env_ll("MR_AHA_HIP_BDT_OPT", 0);
    int opt, ncb;
    if (L > 256) {
        opt = 8;   // at most 64 output groups (1024 work-items)
        ncb = 1;   // 64 KB of LDS per coil
    } else {
        // the largest compiled batch within the cap, the coil count and 80 KB
        // of LDS (two workgroups per CU; MR_AHA_HIP_BDT_LDS sets the budget)
        const long long budget = std::min<long long>(mr_env_ll("MR_AHA_HIP_BDT_LDS", 81920),
                                                     160 * 1024);
        const int fit = std::max(1, (int)(budget / ((long long)L * BDT_TP * sizeof(float2))));
        const int lim = (int)std::min<long long>(std::min<long long>(cap, nc), fit);
        const int batches[] = {12, 8, 6, 4, 2, 1};
        ncb = 1;
        for (int b : batches)
            if (b <= lim) { ncb = b; break; }
        // fewest batches for that size, then the smallest compiled size
        const int nbat0 = (nc + ncb - 1) / ncb;
        for (int b : batches)
            if (b >= (nc + nbat0 - 1) / nbat0 && b <= ncb) ncb = b;
        opt = ncb > 4 ? 2 : 4;
        if (want_opt == 2 || want_opt == 4) opt = (int)want_opt;
        // a compiled form whose workgroup holds all output groups
        while (!bdt_fn<false>(opt, ncb) ||
               BDT_TP * (((L + opt - 1) / opt + 3) / 4 * 4) > bdt_max_threads(ncb)) {
            if (opt == 2 && bdt_fn<false>(4, ncb)) opt = 4;
            else if (ncb > 1) { ncb = ncb > 4 ? 4 : ncb / 2; opt = 4; }
            else break;
        }
    }
    mr_bdt_fn fn = masked ? bdt_fn<true>(opt, ncb) : bdt_fn<false>(opt, ncb);
    MR_REQUIRE(fn != nullptr, "block multiply: no compiled form for %d outputs x %d coils", opt,
               ncb);
    const int ng = (L + opt - 1) / opt;
    const int threads = BDT_TP * ((ng + 3) / 4 * 4);   // whole wavefronts
    MR_REQUIRE(threads <= bdt_max_threads(ncb), "block multiply: %d work-items", threads);
    const size_t lds = (size_t)ncb * L * BDT_TP * sizeof(float2);
    if (lds > 65536 &&
        hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        MR_FAIL(MR_ERR_UNSUPPORTED, "block multiply: %zu bytes of LDS refused", lds);
    }
    const int nbat = (nc + ncb - 1) / ncb;
    const long long blocks = mr_ceil_div(d.pad_size, BDT_TP) * nbat;
    MR_GRID_CHECK(blocks, threads, "mr_k_bdt");
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(threads), lds, s, K, ker, inv, kstride, d,
                       nc, ng, nbat);
    MR_LAUNCH_CHECK("mr_k_bdt");
    return MR_SUCCESS;
}
